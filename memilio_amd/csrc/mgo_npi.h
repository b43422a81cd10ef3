// mgo_npi.h - damping lists of contact matrices that change during a run (dynamic NPIs), shared by the
// host engine (initial lists and caches) and the node kernel (NPI implementation at run time).
//
// A node with dynamic NPIs keeps its contact matrix's damping list on the device, a node whose out-edges have
// coefficient dampings or NPIs keeps their (shared) list; n coefficients per entry (G*G for the contact
// matrix, G for the per-group mobility dampings):
//   list   [cap][3 + n]     time, level, type, coefficients, sorted by (time, type, level)
//   cache  [cap + 2]        accumulated damping coefficients and their times (Dampings::update_cache), the
//                           representation the kernels read (NodeP::damp_time / damp_mat for the contact matrix)
// The functions below restate cpp/memilio/epidemiology/damping.h (add_, remove, update_cache) and
// cpp/memilio/epidemiology/dynamic_npis.h (implement_dynamic_npis) on these flat arrays. They run on one
// thread (the host, or one lane of the node's wavefront group) and never allocate.
#pragma once

#include "mgo_internal.h"

#include <cfloat>
#include <cmath>

#if defined(__HIPCC__)
#define MGO_HD __host__ __device__
#else
#define MGO_HD
#endif

namespace mgo
{

constexpr int NPI_LEVELS = 8; // distinct damping levels per contact matrix (update_cache's sum_by_level)

MGO_HD inline int npi_entry_size(int n)
{
    return 3 + n;
}

// workspace doubles one list needs: value, active, active_end, candidate, combined [5][n], per-level sums
// [NPI_LEVELS][n], active_by_type [cap][3]
MGO_HD inline int npi_work_size(int n, int cap)
{
    return (5 + NPI_LEVELS) * n + 3 * cap;
}

// key order of Dampings::add_ (damping.h:602-613): (time, type, level)
MGO_HD inline bool npi_key_less(double ta, int ya, int la, double tb, int yb, int lb)
{
    if (ta != tb) {
        return ta < tb;
    }
    if (ya != yb) {
        return ya < yb;
    }
    return la < lb;
}

// Dampings::add_ with insert_sorted_replace (utils/stl_util.h:82-96): an entry with an equal key is
// replaced, else the new entry is inserted at its sorted position. false: the list is full.
MGO_HD inline bool npi_list_add(double* list, int& count, int cap, int n, double time, int level, int type,
                                const double* coeffs)
{
#pragma clang fp contract(off)
    const int es = npi_entry_size(n);
    int pos      = 0;
    while (pos < count) {
        const double* e = list + pos * es;
        if (!npi_key_less(e[0], int(e[2]), int(e[1]), time, type, level)) {
            break;
        }
        ++pos;
    }
    double* e = list + pos * es;
    if (pos < count && !npi_key_less(time, type, level, e[0], int(e[2]), int(e[1]))) {
        for (int k = 0; k < n; ++k) {
            e[3 + k] = coeffs[k];
        }
        return true;
    }
    if (count >= cap) {
        return false;
    }
    for (int q = count; q > pos; --q) {
        for (int k = 0; k < es; ++k) {
            list[q * es + k] = list[(q - 1) * es + k];
        }
    }
    e[0] = time;
    e[1] = double(level);
    e[2] = double(type);
    for (int k = 0; k < n; ++k) {
        e[3 + k] = coeffs[k];
    }
    ++count;
    return true;
}

// Dampings::remove (damping.h:336-341)
MGO_HD inline void npi_list_remove(double* list, int& count, int n, int idx)
{
    const int es = npi_entry_size(n);
    for (int q = idx; q + 1 < count; ++q) {
        for (int k = 0; k < es; ++k) {
            list[q * es + k] = list[(q + 1) * es + k];
        }
    }
    --count;
}

// get_active_damping (dynamic_npis.h:291-303): the last entry of (level, type) with time <= t, or -1
MGO_HD inline int npi_list_active(const double* list, int count, int n, int level, int type, double t)
{
    const int es = npi_entry_size(n);
    for (int q = count - 1; q >= 0; --q) {
        const double* e = list + q * es;
        if (int(e[1]) == level && int(e[2]) == type && e[0] <= t) {
            return q;
        }
    }
    return -1;
}

// floating_point_equal(a, b, 1e-15, 1e-15) (math/floating_point.h:39-62)
MGO_HD inline bool npi_time_equal(double a, double b)
{
    const double diff = fabs(a - b);
    return diff <= 1e-15 || diff <= fmax(fabs(a), fabs(b)) * 1e-15;
}

// Dampings::update_cache (damping.h:570-600) with update_active_dampings (damping.h:616-658) and
// inclusive_exclusive_sum (damping.h:549-566). Returns the number of cache entries, or -1 when the list
// has more than NPI_LEVELS distinct levels.
MGO_HD inline int npi_build_cache(const double* list, int count, int n, int cap, double* cache_t, double* cache_m,
                                  double* work)
{
#pragma clang fp contract(off)
    const int es = npi_entry_size(n);
    double* combined = work + 4 * n;
    double* sums     = work + 5 * n;                   // [NPI_LEVELS][n]
    double* abt      = work + (5 + NPI_LEVELS) * n;    // active_by_type: [cap][level, type, entry]
    int levels[NPI_LEVELS];
    int nlev = 0, nabt = 0, m = 1;
    cache_t[0] = -DBL_MAX;
    for (int k = 0; k < n; ++k) {
        cache_m[k] = 0.0;
    }
    for (int q = 0; q < count && q < cap; ++q) {
        const double* e = list + q * es;
        const int lev = int(e[1]), typ = int(e[2]);
        int a = 0;
        while (a < nabt && !(int(abt[3 * a]) == lev && int(abt[3 * a + 1]) == typ)) {
            ++a;
        }
        int l = 0;
        while (l < nlev && levels[l] != lev) {
            ++l;
        }
        if (a < nabt) {
            // replace the active damping of the same type and level
            const double* old = list + int(abt[3 * a + 2]) * es;
            for (int k = 0; k < n; ++k) {
                sums[l * n + k] += e[3 + k] - old[3 + k];
            }
            abt[3 * a + 2] = double(q);
        }
        else {
            abt[3 * nabt]     = double(lev);
            abt[3 * nabt + 1] = double(typ);
            abt[3 * nabt + 2] = double(q);
            ++nabt;
            if (l < nlev) {
                for (int k = 0; k < n; ++k) {
                    sums[l * n + k] += e[3 + k];
                }
            }
            else {
                if (nlev == NPI_LEVELS) {
                    return -1;
                }
                levels[nlev] = lev;
                for (int k = 0; k < n; ++k) {
                    sums[nlev * n + k] = e[3 + k];
                }
                ++nlev;
            }
        }
        for (int k = 0; k < n; ++k) {
            combined[k] = sums[k];
        }
        for (int j = 1; j < nlev; ++j) {
            for (int k = 0; k < n; ++k) {
                const double prod = combined[k] * sums[j * n + k];
                combined[k]       = combined[k] + sums[j * n + k] - prod;
            }
        }
        const int slot = npi_time_equal(e[0], cache_t[m - 1]) ? m - 1 : m;
        for (int k = 0; k < n; ++k) {
            cache_m[slot * n + k] = combined[k];
        }
        if (slot == m) {
            cache_t[m] = e[0];
            ++m;
        }
    }
    cache_t[m] = DBL_MAX;
    for (int k = 0; k < n; ++k) {
        cache_m[m * n + k] = cache_m[(m - 1) * n + k];
    }
    return m + 1;
}

// implement_dynamic_npis (dynamic_npis.h:318-392) on one damping list with the dampings of one threshold
// (damp: [ndamp][3 + G] value, level, type, group weights). contact: the list is a contact matrix's
// (n = G*G, make_contact_damping_matrix of value * weights, damping_sampling.h:282-288); else per-group
// mobility dampings (n = G, value * weights, the distinct values of make_mobility_damping_vector).
// false: the list is full.
MGO_HD inline bool npi_implement(double* list, int& count, int cap, int G, bool contact, const double* damp,
                                 int ndamp, double begin, double end, double* work)
{
#pragma clang fp contract(off)
    const int n = contact ? G * G : G, es = npi_entry_size(n);
    double* value      = work;
    double* active     = work + n;
    double* active_end = work + 2 * n;
    double* cand       = work + 3 * n;
    for (int d = 0; d < ndamp; ++d) {
        const double* s = damp + d * (3 + G);
        const int lev = int(s[1]), typ = int(s[2]);
        if (contact) {
            for (int i = 0; i < G; ++i) {
                for (int j = 0; j < G; ++j) {
                    const double gi = s[0] * s[3 + i];
                    const double gj = s[0] * s[3 + j];
                    value[i * G + j] = 1.0 - sqrt((1.0 - gi) * (1.0 - gj));
                }
            }
        }
        else {
            for (int i = 0; i < G; ++i) {
                value[i] = s[0] * s[3 + i];
            }
        }
        const int ia = npi_list_active(list, count, n, lev, typ, begin);
        const int ie = npi_list_active(list, count, n, lev, typ, end);
        for (int k = 0; k < n; ++k) {
            active[k]     = ia >= 0 ? list[ia * es + 3 + k] : 0.0;
            active_end[k] = ie >= 0 ? list[ie * es + 3 + k] : 0.0;
        }
        bool implemented = false;
        bool any_less    = false;
        for (int k = 0; k < n; ++k) {
            any_less = any_less || (active[k] < value[k]);
        }
        if (any_less) {
            for (int k = 0; k < n; ++k) {
                cand[k] = (value[k] < active[k]) ? active[k] : value[k]; // max(value, active)
            }
            if (!npi_list_add(list, count, cap, n, begin, lev, typ, cand)) {
                return false;
            }
            implemented = true;
        }
        // dampings of the same (level, type) that start inside (begin, end): at least the NPI's value
        for (int q = 0; q < count; ++q) {
            double* e = list + q * es;
            if (int(e[1]) == lev && int(e[2]) == typ && e[0] > begin && e[0] < end) {
                for (int k = 0; k < n; ++k) {
                    e[3 + k] = (e[3 + k] < value[k]) ? value[k] : e[3 + k]; // max(coeffs, value)
                }
                implemented = true;
            }
        }
        if (implemented && !npi_list_add(list, count, cap, n, end, lev, typ, active_end)) {
            return false;
        }
    }
    // a damping equal to the previous one of its (level, type) is dropped (dynamic_npis.h:373-391; the
    // reference's index walk skips the entry after each removal, kept here)
    for (int i = 0; i < count - 1; ++i) {
        const int cur    = count - 1 - i;
        const double* ec = list + cur * es;
        for (int p = cur - 1; p >= 0; --p) {
            const double* ep = list + p * es;
            if (int(ep[1]) == int(ec[1]) && int(ep[2]) == int(ec[2])) {
                bool same = true;
                for (int k = 0; k < n; ++k) {
                    same = same && (ep[3 + k] == ec[3 + k]);
                }
                if (same) {
                    npi_list_remove(list, count, n, cur);
                }
                break;
            }
        }
    }
    return true;
}

} // namespace mgo
