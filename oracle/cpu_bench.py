"""CPU baseline of the bake-off operators: the reference's own
SumFactorization<double, nq> (oracle/_ref/libbkref*.so, compiled from the
reference sources, openmp/bk1_sumfact.h:16-159, openmp/bk3_sumfact.h:17-187,
openmp/BK5.cpp:18-103) timed on this host's cores at the full workload sizes.

Test and measurement infrastructure: run by bench.py's CPU-baseline leg only.

    python oracle/cpu_bench.py [--one-core] [--budget S] CONFIG...

CONFIG is bk1 (p = 3, E = 4096), bk3 (p = 5, E = 32768) or bk5 (p = 7,
E = 65536).  Each apply is timed with a wall clock and the minimum over
repeated applies is kept, as the reference's bk::time_min does
(include/bk_common.h:181-192); GDoF/s = E n^3 / t (include/bk_common.h:
196-202).  The inputs are uniform random values of the workload's shapes
(the loops' time does not depend on the values; the geometric factors lie
in [0.5, 1.5] like the reference's test data, include/bk_common.h:92-107).

The reference's `target teams loop` runs on the host as OpenMP teams; the
cores it keeps busy are measured as process CPU time over wall time.  With
--one-core the process pins itself to one core before the OpenMP runtime
starts, so that the runtime sees a single CPU.  Prints one JSON object.
"""
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))

CONFIGS = {"bk1": (1, 3, 4096), "bk3": (3, 5, 32768), "bk5": (5, 7, 65536)}
AVX512 = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def cpu_info():
    """Model name, logical CPUs this process may use, and the ISA level of
    the reference build that runs here (x86-64-v4 needs the AVX-512 F, BW,
    CD, DQ and VL extensions, else x86-64-v3)."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    isa = "x86-64-v4" if all(x in flags for x in AVX512) else "x86-64-v3"
    return {"model": model, "cpus_available": len(os.sched_getaffinity(0)), "isa": isa}


def library(isa):
    name = "libbkref_v4.so" if isa == "x86-64-v4" else "libbkref.so"
    path = os.path.join(HERE, "_ref", name)
    if not os.path.exists(path):
        path = os.path.join(HERE, "_ref", "libbkref.so")
    return path


def time_config(lib, name, budget_s, max_reps):
    import numpy as np
    op, p, E = CONFIGS[name]
    rng = np.random.default_rng(7)
    n = p + 1
    q = p + 2 if op != 5 else n
    u = rng.uniform(-1, 1, E * n ** 3)
    v = np.empty_like(u)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    if op == 1:
        B, J = rng.uniform(-1, 1, n * q), rng.uniform(0.5, 1.5, E * q ** 3)
        f = lib.bkref_bk1_sumfact
        args = (ptr(B), ptr(J), ptr(u), ptr(v))
    elif op == 3:
        B, D, G = rng.uniform(-1, 1, n * q), rng.uniform(-1, 1, q * q), rng.uniform(0.5, 1.5, E * 6 * q ** 3)
        f = lib.bkref_bk3_sumfact
        args = (ptr(B), ptr(D), ptr(G), ptr(u), ptr(v))
    else:
        D, G = rng.uniform(-1, 1, n * n), rng.uniform(0.5, 1.5, E * 6 * n ** 3)
        f = lib.bkref_bk5_sumfact
        args = (ptr(D), ptr(G), ptr(u), ptr(v))
    f.argtypes = [ctypes.c_int, ctypes.c_size_t] + [ctypes.c_void_p] * len(args)
    f.restype = ctypes.c_int
    if f(p, E, *args) != 0:   # warm-up apply (first touch of v, thread start-up)
        raise RuntimeError("reference kernel rejected p = %d" % p)
    best, reps = float("inf"), 0
    wall0, cpu0 = time.perf_counter(), time.process_time()
    while reps < 3 or (time.perf_counter() - wall0 < budget_s and reps < max_reps):
        t0 = time.perf_counter()
        f(p, E, *args)
        best = min(best, time.perf_counter() - t0)
        reps += 1
    busy = (time.process_time() - cpu0) / (time.perf_counter() - wall0)
    dofs = E * n ** 3
    return {"p": p, "nelmt": E, "dofs": dofs, "GDoF/s": round(dofs / best * 1e-9, 4), "ms": round(best * 1e3, 3),
            "applies": reps, "cores_busy": round(busy, 1)}


def main(argv):
    one = "--one-core" in argv
    budget = 3.0
    if "--budget" in argv:
        budget = float(argv[argv.index("--budget") + 1])
    names = [a for a in argv if a in CONFIGS]
    if one:
        # before the OpenMP runtime is loaded: it sizes its teams and thread
        # pool from the affinity mask and OMP_NUM_THREADS when it starts
        os.sched_setaffinity(0, {min(os.sched_getaffinity(0))})
        os.environ["OMP_NUM_THREADS"] = "1"
    info = cpu_info()
    lib = ctypes.CDLL(library(info["isa"]))
    out = {"cpu": info, "one_core": one, "configs": {}}
    for name in names:
        out["configs"][name] = time_config(lib, name, budget, 20 if not one else 5)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
