"""Thin Python driver over the C ABI of libmgo.so (include/mgo.h), used by tests and bench.py.

Every call goes to the HIP engine; a missing library or device is an error, never a fallback.
"""
import ctypes as C

import numpy as np

from . import capi


class MgoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mgo status {status}: {msg}")
        self.status = status


class Context:
    def __init__(self, device=0, hbm_budget_bytes=0, flags=0):
        self.lib = capi.load()
        self.ptr = C.c_void_p()
        cfg = capi.DeviceCfg(device, hbm_budget_bytes, flags)
        st = self.lib.mgo_context_create(C.byref(cfg), C.byref(self.ptr))
        if st != capi.OK:
            raise MgoError(st, self.lib.mgo_last_error(None).decode())

    def check(self, st):
        if st != capi.OK:
            raise MgoError(st, self.lib.mgo_last_error(self.ptr).decode())

    def close(self):
        if self.ptr:
            self.lib.mgo_context_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_flags(self, flags):
        """MGO_CTX_* flags for the runs started afterwards."""
        self.check(self.lib.mgo_context_set_flags(self.ptr, flags))

    def ensemble_mean_values(self, values):
        """ensemble_mean over the first axis of host values [run, ...], computed on the device; returns the
        shape of one run (mgo_ensemble_mean_values)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.ndim < 1 or v.shape[0] < 1 or v[0].size < 1:
            raise ValueError("ensemble_mean_values needs at least one run with values")
        out = np.empty(v.shape[1:], dtype=np.float64)
        self.check(self.lib.mgo_ensemble_mean_values(self.ptr, capi.dptr(v), v.shape[0], v[0].size, capi.dptr(out)))
        return out

    def ensemble_percentiles_values(self, values, ps):
        """ensemble_percentile over the first axis of host values [run, ...] for each p in ps, computed on the
        device; returns [len(ps), ...] (mgo_ensemble_percentiles_values)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.ndim < 1 or v.shape[0] < 1 or v[0].size < 1:
            raise ValueError("ensemble_percentiles_values needs at least one run with values")
        p = np.ascontiguousarray(np.atleast_1d(ps), dtype=np.float64)
        out = np.empty((p.size,) + v.shape[1:], dtype=np.float64)
        self.check(self.lib.mgo_ensemble_percentiles_values(self.ptr, capi.dptr(v), v.shape[0], v[0].size,
                                                            capi.dptr(p), p.size, capi.dptr(out)))
        return out

    def set_tile(self, tile):
        """Runs per node-major tile for ensembles created afterwards (0 = autotuned)."""
        self.check(self.lib.mgo_context_set_tile(self.ptr, tile))

    def clock_ms(self):
        return self.lib.mgo_clock_ms(self.ptr)

    # ------------------------------------------------------------------ single model
    def simulate_single(self, spec, cfg, tmax):
        nd = spec.num_days(cfg.t0, tmax)
        out = np.zeros((max(nd, 1), spec.num_nodes, spec.state_size()))
        got = C.c_int32(0)
        self.check(self.lib.mgo_simulate_single(self.ptr, C.byref(spec.desc()), C.byref(cfg), tmax, nd,
                                                capi.dptr(out), C.byref(got)))
        return out[:got.value]

    # ------------------------------------------------------------------ ensemble
    def _ensemble_desc(self, specs, cfgs, tmaxs):
        nr = len(specs)
        graphs = (capi.GraphDesc * nr)(*[s.desc() for s in specs])
        cfg_arr = (capi.SimCfg * nr)(*cfgs)
        tm = np.ascontiguousarray(tmaxs, dtype=np.float64)
        desc = capi.EnsembleDesc(nr, graphs, cfg_arr, capi.dptr(tm))
        keep = (graphs, cfg_arr, tm)
        return desc, keep

    def ensemble_run(self, specs, cfgs, tmaxs, want_output=True):
        desc, keep = self._ensemble_desc(specs, cfgs, tmaxs)
        sizes = [s.num_days(c.t0, t) * s.num_nodes * s.state_size() for s, c, t in zip(specs, cfgs, tmaxs)]
        out = np.zeros(sum(sizes)) if want_output else None
        info = capi.RunInfo()
        self.check(self.lib.mgo_ensemble_run(self.ptr, C.byref(desc), capi.dptr(out) if want_output else None,
                                             C.byref(info)))
        del keep
        if not want_output:
            return None, info
        res, o = [], 0
        for s, c, t, sz in zip(specs, cfgs, tmaxs, sizes):
            res.append(out[o:o + sz].reshape(s.num_days(c.t0, t), s.num_nodes, s.state_size()))
            o += sz
        return res, info


class Batch:
    """Device-resident ensemble: create once, run, read back (mgo_batch_*)."""

    def __init__(self, ctx, specs, cfgs, tmaxs):
        self.ctx = ctx
        self.specs, self.cfgs, self.tmaxs = specs, cfgs, list(tmaxs)
        desc, keep = ctx._ensemble_desc(specs, cfgs, tmaxs)
        self.id = C.c_int64(0)
        ctx.check(ctx.lib.mgo_batch_create(ctx.ptr, C.byref(desc), C.byref(self.id)))
        del keep

    def run(self):
        info = capi.RunInfo()
        self.ctx.check(self.ctx.lib.mgo_batch_run(self.ctx.ptr, self.id, C.byref(info)))
        return info

    def reset(self):
        self.ctx.check(self.ctx.lib.mgo_batch_reset(self.ctx.ptr, self.id))

    def read_daily(self, first=0, count=None):
        count = len(self.specs) - first if count is None else count
        sel = list(zip(self.specs, self.cfgs, self.tmaxs))[first:first + count]
        sizes = [s.num_days(c.t0, t) * s.num_nodes * s.state_size() for s, c, t in sel]
        out = np.zeros(sum(sizes))
        self.ctx.check(self.ctx.lib.mgo_batch_read_daily(self.ctx.ptr, self.id, first, count, capi.dptr(out)))
        res, o = [], 0
        for (s, c, t), sz in zip(sel, sizes):
            res.append(out[o:o + sz].reshape(s.num_days(c.t0, t), s.num_nodes, s.state_size()))
            o += sz
        return res

    def _shape(self, first, count):
        count = len(self.specs) - first if count is None else count
        s, c, t = self.specs[first], self.cfgs[first], self.tmaxs[first]
        return count, s.num_days(c.t0, t), s.num_nodes, s.state_size()

    def ensemble_mean(self, first=0, count=None):
        """ensemble_mean over runs [first, first+count): [day][node][10G], computed on the device."""
        count, nd, n, w = self._shape(first, count)
        out = np.zeros((nd, n, w))
        self.ctx.check(self.ctx.lib.mgo_batch_ensemble_mean(self.ctx.ptr, self.id, first, count, capi.dptr(out)))
        return out

    def ensemble_percentiles(self, ps, first=0, count=None):
        """ensemble_percentile for each p in ps: [len(ps)][day][node][10G], computed on the device."""
        count, nd, n, w = self._shape(first, count)
        p = np.ascontiguousarray(ps, dtype=np.float64)
        out = np.zeros((p.size, nd, n, w))
        self.ctx.check(self.ctx.lib.mgo_batch_ensemble_percentiles(self.ctx.ptr, self.id, first, count, capi.dptr(p),
                                                                   p.size, capi.dptr(out)))
        return out

    def sum_nodes(self, first=0, count=None):
        """sum_nodes of runs [first, first+count): [run][day][10G], computed on the device."""
        count, nd, n, w = self._shape(first, count)
        out = np.zeros((count, nd, w))
        self.ctx.check(self.ctx.lib.mgo_batch_sum_nodes(self.ctx.ptr, self.id, first, count, capi.dptr(out)))
        return out

    def close(self):
        if self.id.value:
            self.ctx.lib.mgo_batch_destroy(self.ctx.ptr, self.id)
            self.id = C.c_int64(0)


class GraphSim:
    """make_mobility_sim(t0, dt, graph) on the device; advance(tmax) == GraphSimulation::advance."""

    def __init__(self, ctx, spec, cfg, horizon):
        self.ctx, self.spec = ctx, spec
        cfg = capi.SimCfg(cfg.t0, cfg.dt_graph, cfg.dt_node, horizon, cfg.integrator)
        self.cfg = cfg
        self.id = C.c_int64(0)
        ctx.check(ctx.lib.mgo_graph_create(ctx.ptr, C.byref(spec.desc()), C.byref(cfg), C.byref(self.id)))

    def advance(self, tmax):
        self.ctx.check(self.ctx.lib.mgo_graph_advance(self.ctx.ptr, self.id, tmax))

    def get_t(self):
        t = C.c_double(0)
        self.ctx.check(self.ctx.lib.mgo_graph_get_time(self.ctx.ptr, self.id, C.byref(t)))
        return t.value

    def state(self):
        y = np.zeros((self.spec.num_nodes, self.spec.state_size()))
        self.ctx.check(self.ctx.lib.mgo_graph_read_state(self.ctx.ptr, self.id, capi.dptr(y)))
        return y

    def daily(self):
        cap = self.spec.num_days(self.cfg.t0, self.cfg.horizon)
        out = np.zeros((max(cap, 1), self.spec.num_nodes, self.spec.state_size()))
        n = C.c_int32(0)
        self.ctx.check(self.ctx.lib.mgo_graph_read_daily(self.ctx.ptr, self.id, cap, capi.dptr(out), C.byref(n)))
        return out[:n.value]

    def stats(self):
        s = np.zeros(self.spec.num_nodes, np.int64)
        a = np.zeros(self.spec.num_nodes, np.int64)
        self.ctx.check(self.ctx.lib.mgo_graph_read_stats(self.ctx.ptr, self.id,
                                                         s.ctypes.data_as(C.POINTER(C.c_int64)),
                                                         a.ctypes.data_as(C.POINTER(C.c_int64))))
        return s, a

    def close(self):
        if self.id.value:
            self.ctx.lib.mgo_graph_destroy(self.ctx.ptr, self.id)
            self.id = C.c_int64(0)


class Server:
    """Continuous-batching request loop of a context (mgo_serve_*)."""

    def __init__(self, ctx, max_resident=0, hbm_budget_bytes=0, flags=0, result_pool_bytes=0):
        self.ctx = ctx
        cfg = capi.ServeCfg(max_resident, hbm_budget_bytes, flags, result_pool_bytes)
        ctx.check(ctx.lib.mgo_serve_start(ctx.ptr, C.byref(cfg)))
        self.requests = {}
        self.finished = {}   # polled, not yet taken: ticket -> ResultInfo
        self.running = True

    def submit(self, spec, cfg, tmax):
        req = capi.Request(spec.desc(), cfg, tmax)
        t = C.c_int64(0)
        self.ctx.check(self.ctx.lib.mgo_serve_submit(self.ctx.ptr, C.byref(req), C.byref(t)))
        self.requests[t.value] = (spec, cfg, tmax)
        return t.value

    def update(self, ticket, scale, t_start=-1.0, level=0, type=0):
        u = capi.DampingUpdate(scale, t_start, level, type)
        self.ctx.check(self.ctx.lib.mgo_serve_update(self.ctx.ptr, ticket, C.byref(u)))

    def poll(self, capacity=256):
        arr = (capi.ResultInfo * capacity)()
        n = C.c_int32(0)
        self.ctx.check(self.ctx.lib.mgo_serve_poll(self.ctx.ptr, arr, capacity, C.byref(n)))
        return [arr[k] for k in range(n.value)]

    def take(self, ticket):
        """Daily values [day][node][10G] of a finished request."""
        spec, cfg, tmax = self.requests.pop(ticket)
        self.finished.pop(ticket, None)
        nd = spec.num_days(cfg.t0, tmax)
        out = np.zeros((nd, spec.num_nodes, spec.state_size()))
        self.ctx.check(self.ctx.lib.mgo_serve_take(self.ctx.ptr, ticket, capi.dptr(out), out.size))
        return out

    def wait(self, tickets, timeout_s=60.0):
        """Poll until every ticket in tickets has finished; returns {ticket: ResultInfo}."""
        import time
        want = set(tickets)
        t_end = time.monotonic() + timeout_s
        while want - set(self.finished):
            for r in self.poll():
                self.finished[r.ticket] = r
            if want - set(self.finished):
                if time.monotonic() > t_end:
                    raise TimeoutError(f"requests {sorted(want - set(self.finished))} not finished in {timeout_s} s")
                time.sleep(0.0005)
        return {t: self.finished[t] for t in want}

    def updates_applied(self, ticket):
        """Scenario updates a finished request ran with: [(scale, t_start, level, type)], t_start as applied."""
        cap = 64
        while True:
            arr = (capi.DampingUpdate * cap)()
            n = C.c_int32(0)
            self.ctx.check(self.ctx.lib.mgo_serve_updates_applied(self.ctx.ptr, ticket, arr, cap, C.byref(n)))
            if n.value <= cap:
                return [(arr[k].scale, arr[k].t_start, arr[k].level, arr[k].type) for k in range(n.value)]
            cap = n.value

    def stats(self):
        """Counters of the serving loop (mgo_serve_get_stats) as a dict."""
        st = capi.ServeStats()
        self.ctx.check(self.ctx.lib.mgo_serve_get_stats(self.ctx.ptr, C.byref(st)))
        return {name: getattr(st, name) for name, _ in capi.ServeStats._fields_ if name != "reserved"}

    def stop(self):
        if self.running:
            self.ctx.check(self.ctx.lib.mgo_serve_stop(self.ctx.ptr))
            self.running = False
