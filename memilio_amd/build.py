"""Builds memilio_amd/libmgo.so (host engine + gfx950 kernels) with hipcc, in tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmgo.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _stale(obj, deps):
    return not os.path.exists(obj) or any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def build(verbose=False, only_changed=False):
    """Compile both translation units and link libmgo.so; with only_changed, an object newer than its
    sources and headers is reused."""
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    headers = [os.path.join(CSRC, "mgo_internal.h"), os.path.join(HERE, "..", "include", "mgo.h")]
    for src, extra, deps in (("mgo_kernels.hip", [], []),
                             ("mgo_host.cpp", ["-x", "hip", "-ffp-contract=off"], ["mgo_serve.inc"])):
        obj = os.path.join(HERE, "_obj", src + ".o")
        objs_deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, d) for d in deps] + headers
        if only_changed and not _stale(obj, objs_deps):
            objs.append(obj)
            continue
        cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, only_changed="--changed" in sys.argv))
