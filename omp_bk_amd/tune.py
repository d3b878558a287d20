"""Autotuner for the operators' launch variants.

Each operator form has a set of compiled launch variants (thread mapping,
wavefronts or elements per workgroup, load kind, occupancy target;
``bkx_tune_count``).  The tuner times every variant on a synthetic workload
of the requested order and size on the device, checks that each variant's
output matches variant 0 (max-relative difference <= 1e-12, like the
reference harness's comparison of every variant with its first kernel,
openmp/offload/bk1_offload.cpp:184-196), selects the fastest with
``bkx_tune_set`` and returns the timings.  The reference chooses its
variants offline with a median-of-repetitions picker
(scripts/study/bk3_pick.sh); here the choice is made per order at run time
and can be saved to and loaded from a JSON table.

    python -m omp_bk_amd.tune --op 5 --form f64 --p 7 --nelmt 65536
"""
import argparse
import json
import statistics

import torch

from . import _lib
from . import geometry as geo
from . import ops

FORMS = {"f64": 0, "f32c": 1, "f32w": 2, "bf16s": 3, "ff": 4}


def _workload(op, form, p, nelmt, device):
    torch.manual_seed(p)
    dt = torch.float64 if form in ("f64", "ff") else torch.float32
    if op == 5:
        n = p + 1
        D = torch.as_tensor(geo.bk5_basis(p), device=device)
        G = geo.geometric_factors(5, p, nelmt, shape=_shape(nelmt), device=device).reshape(-1).to(dt)
        u = (torch.rand(nelmt * n ** 3, dtype=torch.float64, device=device) * 2 - 1).to(dt)
        return lambda out: ops.bk5(p, D, G, u, out=out, form=form), nelmt * n ** 3
    n = p + 1
    B, D = (torch.as_tensor(a, device=device) for a in geo.bk13_bases(p))
    u = (torch.rand(nelmt * n ** 3, dtype=torch.float64, device=device) * 2 - 1).to(dt)
    if op == 3:
        G = geo.geometric_factors(3, p, nelmt, shape=_shape(nelmt), device=device).reshape(-1).to(dt)
        return lambda out: ops.bk3(p, B, D, G, u, out=out, form=form), nelmt * n ** 3
    J = geo.geometric_factors(1, p, nelmt, shape=_shape(nelmt), device=device).reshape(-1).to(dt)
    return lambda out: ops.bk1(p, B, J, u, out=out, form=form), nelmt * n ** 3


def _shape(nelmt):
    try:
        return geo.box_shape(nelmt)
    except ValueError:
        return (nelmt, 1, 1)


def _time(fn, out, steps, warmup):
    for _ in range(warmup):
        fn(out)
    torch.cuda.synchronize()
    times = []
    s = torch.cuda.current_stream()
    for _ in range(steps):
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(s)
        fn(out)
        t1.record(s)
        t1.synchronize()
        times.append(t0.elapsed_time(t1))
    return statistics.median(times)


def tune(op, form, p, nelmt, steps=20, warmup=3, device="cuda:0", apply=True):
    """Time every variant of (op, form) at order p; returns
    {"best": v, "ms": {v: median ms}, "relerr": {v: err vs variant 0}}."""
    L = _lib.lib()
    f = FORMS[form]
    count = L.bkx_tune_count(op, f)
    if count < 1:
        raise ValueError("no variants for op %d form %s" % (op, form))
    fn, ndof = _workload(op, form, p, nelmt, device)
    ref = torch.empty(ndof, dtype=torch.float64, device=device)
    out = torch.empty_like(ref)
    prev = L.bkx_tune_get(op, f, p)
    ms, err = {}, {}
    for v in range(count):
        _lib.call("bkx_tune_set", op, f, p, v)
        out.fill_(float("nan"))
        fn(out)
        torch.cuda.synchronize()
        if v == 0:
            ref.copy_(out)
            err[v] = 0.0
        else:
            err[v] = float(((out - ref).abs().max() / ref.abs().max()).item())
        if not (err[v] <= 1e-12):
            ms[v] = float("inf")
            continue
        ms[v] = _time(fn, out, steps, warmup)
    best = min(ms, key=ms.get)
    _lib.call("bkx_tune_set", op, f, p, best if apply else prev)
    return {"best": best, "ms": ms, "relerr": err, "nelmt": nelmt, "ndof": ndof}


def save_table(table, path):
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)


def load_table(path):
    """Apply a saved {"op/form/p": variant} table with bkx_tune_set."""
    with open(path) as fh:
        table = json.load(fh)
    for key, v in table.items():
        op, form, p = key.split("/")
        _lib.call("bkx_tune_set", int(op), FORMS[form], int(p), int(v["best"] if isinstance(v, dict) else v))
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--op", type=int, default=5)
    ap.add_argument("--form", default="f64")
    ap.add_argument("--p", type=int, nargs="+", default=[7])
    ap.add_argument("--nelmt", type=int, default=65536)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--save", default=None)
    a = ap.parse_args()
    table = {}
    for p in a.p:
        r = tune(a.op, a.form, p, a.nelmt, steps=a.steps)
        key = "%d/%s/%d" % (a.op, a.form, p)
        table[key] = r
        rates = {v: (r["ndof"] / (t * 1e-3) * 1e-9 if t != float("inf") else 0.0) for v, t in r["ms"].items()}
        print(json.dumps({"key": key, "best": r["best"], "gdofs": {k: round(x, 2) for k, x in rates.items()},
                          "relerr": r["relerr"]}), flush=True)
    if a.save:
        save_table(table, a.save)


if __name__ == "__main__":
    main()
