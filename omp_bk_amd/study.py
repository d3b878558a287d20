"""Study tools: the reference's scripts/study and scripts/plot workflows over
this package's harnesses (bin/BKn, bin/BKn_variants, bin/stream).

    python -m omp_bk_amd.study pick --op 3 --orders 5 7 [--dofs 2e7] [--reps 3] [out.csv]
    python -m omp_bk_amd.study sweep [--progs BK1 BK3 BK5] [--orders 2 4 6] [--nelmt N] [out.csv]
    python -m omp_bk_amd.study dof-sweep --op 1 --kernels 2:v7 4:v6 [out.csv]
    python -m omp_bk_amd.study roofline sweep.csv [--bw GBs] [--peak GFLOPs]
    python -m omp_bk_amd.study stream

pick       scripts/study/bk3_pick.sh: every variant of bin/BK<op>_variants
           (BK_KERNEL="-name ..." leaves some out), REPS runs each, CSV
           api,p,variant,rep,gdofs,gbs,relerr; then the fastest variant per
           order by median GDoF/s, FAIL rows excluded, as a share of the
           copy roof and of the read-mix roof of the operator (read9 for BK1
           and BK3, the reference's BK3 roof; read7 for BK5 in FP64).
sweep      scripts/study/sweep.sh: bin/BK1, BK3, BK5 per order, CSV
           prog,p,arg,nelmt,gdofs,gbs,peak_mb (peak_mb: the child's peak
           resident host memory, where the reference reads its cgroup's).
dof-sweep  scripts/study/dof_sweep.sh: GDoF/s over problem size for one
           variant per order, DoFs log-spaced, CSV
           api,nq,kernel,nelmt,dofs,gdofs,gbs,gdofs_dev,gbs_dev (no device
           columns: the harness times on the host, as the reference's omp rows).
roofline   scripts/plot/plot_roofline.py's table (no plot: matplotlib is not
           required here): per point the arithmetic intensity, GFLOP/s, the
           roof in GDoF/s and the share of it, from the reference's work and
           traffic model per DoF (bk1_cost, bk5_cost; bk3_cost is the same
           count applied to BK3) with the word size of the operators (8 B,
           FP64; the reference's floats are 4 B).  Reads the sweep and
           dof-sweep CSVs.
stream     bin/stream (roofline/stream_omp.cpp): the copy, read7 and read9 roofs.

The harnesses run on the GPU; the cost model and the CSV parsing do not.
"""

import argparse
import csv
import math
import os
import statistics
import subprocess
import sys
from collections import defaultdict

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")

# MI355X FP64 vector peak (GFLOP/s) and HBM3E bandwidth (GB/s), published
FP64_PEAK = 78.6e3
HBM_PEAK = 8.0e3


# ---------------------------------------------------------------------------
# cost model (scripts/plot/plot_roofline.py:46-53), per DoF
# ---------------------------------------------------------------------------

def bk1_cost(p, word=8):
    """BK1 flops and bytes per DoF (a mode): nm = p + 1, nq = p + 2; three
    interpolation stages there and back (an FMA counts 2), the JxW product;
    in + out + JxW (plot_roofline.py:46-49)."""
    nm, nq = p + 1, p + 2
    flops = (4 * (nq * nm ** 3 + nq ** 2 * nm ** 2 + nq ** 3 * nm) + nq ** 3) / nm ** 3
    return flops, word * (2 + (nq / nm) ** 3)


def bk3_cost(p, word=8):
    """BK3 flops and bytes per DoF (a mode): BK1's interpolation there and
    back, and at each of the nq^3 points BK5's per-point work, 12 nq + 20
    (three derivatives, G, three transposed, the sum); in + out + six
    factors of G per point."""
    nm, nq = p + 1, p + 2
    flops = (4 * (nq * nm ** 3 + nq ** 2 * nm ** 2 + nq ** 3 * nm) + (12 * nq + 20) * nq ** 3) / nm ** 3
    return flops, word * (2 + 6 * (nq / nm) ** 3)


def bk5_cost(p, F=6, word=8):
    """BK5 flops and bytes per DoF (a point, nq = p + 1): 12 nq + 20;
    in + out + F factors of G (plot_roofline.py:52-53)."""
    return 12 * (p + 1) + 20, word * (2 + F)


COST = {"BK1": bk1_cost, "BK3": bk3_cost, "BK5": bk5_cost}


def roof_point(prog, p, gdofs, bw=HBM_PEAK, peak=FP64_PEAK, word=8):
    """(flop/B, GFLOP/s, roof GDoF/s, share of roof) of one measured point."""
    fl, by = COST[prog](p, word=word)
    roof = min(bw / by, peak / fl)
    return fl / by, gdofs * fl, roof, gdofs / roof


# ---------------------------------------------------------------------------
# harness output
# ---------------------------------------------------------------------------

def parse_variant_lines(text):
    """[(variant, gdofs, gbs, relerr)] of the variant harnesses' lines: the
    pick scripts' awk rule, $2 == "GDoF/s" -> $1, $4, $7, $NF (relerr is the
    string "FAIL" for a failed variant)."""
    rows = []
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 7 and f[1] == "GDoF/s":
            rows.append((f[0], float(f[3]), float(f[6]), f[-1]))
    return rows


def parse_driver_line(text):
    """(gdofs, gbs) of a per-order driver's 'SumFactorization -> ... GDoF/s
    = x GB/s = y' line (sweep.sh's sed rule), None if there is none."""
    for line in text.splitlines():
        if line.startswith("SumFactorization") and "GDoF/s" in line:
            f = line.replace("=", " = ").split()
            return float(f[f.index("GDoF/s") + 2]), float(f[f.index("GB/s") + 2])
    return None


def parse_stream(text):
    """{name: GB/s} of bin/stream's lines ($1 -> $NF)."""
    out = {}
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 3 and f[-3] == "GB/s":
            out[f[0]] = float(f[-1])
    return out


def _run(args, env=None):
    """(stdout + stderr, peak resident MB) of one harness run."""
    e = dict(os.environ)
    e.update(env or {})
    proc = subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, env=e, text=True)
    out = proc.stdout.read()
    proc.stdout.close()
    _, status, ru = os.wait4(proc.pid, 0)
    proc.returncode = os.waitstatus_to_exitcode(status)
    if proc.returncode != 0:
        raise RuntimeError("%s exited with %d:\n%s" % (" ".join(args), proc.returncode, out[-2000:]))
    return out, ru.ru_maxrss // 1024


def stream():
    out, _ = _run([os.path.join(BIN, "stream")])
    return parse_stream(out)


# ---------------------------------------------------------------------------
# workflows
# ---------------------------------------------------------------------------

def pick(op, orders, dofs=2e7, reps=3, ntests=10, out="bk_pick.csv", exclude=""):
    """scripts/study/bk3_pick.sh for BK<op>: writes the CSV, returns its rows."""
    rows = []
    for p in orders:
        ne = max(1, round(dofs / (p + 1) ** 3))   # p + 1 modes (points, BK5) a side
        for rep in range(1, reps + 1):
            text, _ = _run([os.path.join(BIN, "BK%d_variants" % op), str(p), str(ne), str(ntests)],
                           env={"BK_KERNEL": exclude} if exclude else None)
            for v, g, b, r in parse_variant_lines(text):
                rows.append({"api": "hip", "p": p, "variant": v, "rep": rep, "gdofs": g, "gbs": b, "relerr": r})
        print("p=%d done" % p, file=sys.stderr)
    with open(out, "w", newline="") as f:
        w = csv.DictWriter(f, ["api", "p", "variant", "rep", "gdofs", "gbs", "relerr"])
        w.writeheader()
        w.writerows(rows)
    return rows


def best_by_median(rows):
    """{(api, p): (variant, median GDoF/s, median GB/s)} over variants with
    no FAIL row, and the set of failed (api, p, variant)."""
    gd, gb, bad = defaultdict(list), defaultdict(list), set()
    for r in rows:
        k = (r["api"], int(r["p"]), r["variant"])
        gd[k].append(float(r["gdofs"]))
        gb[k].append(float(r["gbs"]))
        if r["relerr"] == "FAIL":
            bad.add(k)
    best = {}
    for k, g in gd.items():
        m = statistics.median(g)
        if k not in bad and ((k[0], k[1]) not in best or m > best[(k[0], k[1])][1]):
            best[(k[0], k[1])] = (k[2], m, statistics.median(gb[k]))
    return best, bad


def pick_summary(rows, op, roofs=None):
    """The pick scripts' summary lines."""
    best, bad = best_by_median(rows)
    mix = "read7" if op == 5 else "read9"
    lines = []
    if roofs and "copy" in roofs and mix in roofs:
        copy, rd = roofs["copy"], roofs[mix]
        lines.append("roofs: copy %.0f GB/s, %s %.0f GB/s" % (copy, mix, rd))
    else:
        copy = rd = None
        lines.append("roofs: unknown (no bin/stream run)")
    for (api, p), (v, g, b) in sorted(best.items()):
        pct = "  (%3.0f%% of copy, %3.0f%% of %s)" % (100 * b / copy, 100 * b / rd, mix) if copy else ""
        lines.append("%-5s p=%d %-12s %6.2f GDoF/s  %5.0f GB/s%s" % (api, p, v, g, b, pct))
    if bad:
        lines.append("FAIL: %s" % sorted(bad))
    return lines


def sweep(progs=("BK1", "BK3", "BK5"), orders=(2, 4, 6), nelmt=524288, ntests=10, out="gpu_results.csv"):
    """scripts/study/sweep.sh: writes the CSV, returns its rows."""
    rows = []
    for prog in progs:
        for p in orders:
            text, peak = _run([os.path.join(BIN, prog), str(p), str(nelmt), str(ntests)])
            g = parse_driver_line(text)
            row = {"prog": prog, "p": p, "arg": p, "nelmt": nelmt, "gdofs": g[0] if g else "NaN",
                   "gbs": g[1] if g else "NaN", "peak_mb": peak}
            print(",".join(str(row[k]) for k in row))
            rows.append(row)
    with open(out, "w", newline="") as f:
        w = csv.DictWriter(f, ["prog", "p", "arg", "nelmt", "gdofs", "gbs", "peak_mb"])
        w.writeheader()
        w.writerows(rows)
    return rows


def dof_nelmts(nm, dof_min=1e3, dof_max=1e7, ppd=4):
    """The distinct element counts for DoF targets log-spaced from dof_min to
    dof_max, ppd per decade (dof_sweep.sh's nelmts)."""
    lo, hi = math.log10(dof_min), math.log10(dof_max)
    n = round((hi - lo) * ppd)
    return sorted({max(1, round(10 ** (lo + i / ppd) / nm ** 3)) for i in range(n + 1)})


def dof_sweep(op, kernels, dof_min=1e3, dof_max=1e7, ppd=4, ntests=20, out="dof_sweep.csv"):
    """scripts/study/dof_sweep.sh for BK<op>: kernels = {p: variant}."""
    rows = []
    for p, k in sorted(kernels.items()):
        nm = p + 1
        nq = p + 1 if op == 5 else p + 2
        for ne in dof_nelmts(nm, dof_min, dof_max, ppd):
            text, _ = _run([os.path.join(BIN, "BK%d_variants" % op), str(p), str(ne), str(ntests)],
                           env={"BK_KERNEL": k})
            got = [r for r in parse_variant_lines(text) if r[0] == k]
            g, b = (got[0][1], got[0][2]) if got else ("NaN", "NaN")
            rows.append({"api": "hip", "nq": nq, "kernel": k, "nelmt": ne, "dofs": ne * nm ** 3, "gdofs": g,
                         "gbs": b, "gdofs_dev": "", "gbs_dev": ""})
    with open(out, "w", newline="") as f:
        w = csv.DictWriter(f, ["api", "nq", "kernel", "nelmt", "dofs", "gdofs", "gbs", "gdofs_dev", "gbs_dev"])
        w.writeheader()
        w.writerows(rows)
    return rows


def roofline_table(rows, bw=HBM_PEAK, peak=FP64_PEAK, prog=None):
    """plot_roofline.py's table from sweep rows (prog, p, gdofs) or dof-sweep
    rows (nq, dofs, gdofs; prog given): the largest problem per (prog, p)."""
    best = {}
    for r in rows:
        if r.get("gdofs") in ("", "NaN", None):
            continue
        if "prog" in r:
            key, size = (r["prog"], int(r["p"])), int(r["nelmt"])
        else:
            nq = int(r["nq"])
            key, size = (prog, nq - 1 if prog == "BK5" else nq - 2), int(r["dofs"])
        if key not in best or size > best[key][0]:
            best[key] = (size, float(r["gdofs"]))
    lines = ["roofs: memory %.0f GB/s, FP64 peak %.0f GFLOP/s, ridge %.1f flop/B" % (bw, peak, peak / bw),
             "%-5s %2s %7s %8s %7s %11s %7s %7s" % ("prog", "p", "flop/B", "GFLOP/s", "GDoF/s", "roof GDoF/s",
                                                  "of roof", "of peak")]
    for (pg, p), (_, g) in sorted(best.items()):
        ai, gf, roof, frac = roof_point(pg, p, g, bw, peak)
        lines.append("%-5s %2d %7.2f %8.0f %7.2f %11.2f %6.0f%% %6.1f%%" % (pg, p, ai, gf, g, roof, 100 * frac,
                                                                           100 * gf / peak))
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m omp_bk_amd.study", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("pick")
    a.add_argument("--op", type=int, choices=(1, 3, 5), required=True)
    a.add_argument("--orders", type=int, nargs="+", default=list(range(1, 8)))
    a.add_argument("--dofs", type=float, default=2e7)
    a.add_argument("--reps", type=int, default=3)
    a.add_argument("--ntests", type=int, default=10)
    a.add_argument("--exclude", default="", help='BK_KERNEL value, e.g. "-v3 -v20"')
    a.add_argument("out", nargs="?", default="bk_pick.csv")
    a = sub.add_parser("sweep")
    a.add_argument("--progs", nargs="+", default=["BK1", "BK3", "BK5"])
    a.add_argument("--orders", type=int, nargs="+", default=[2, 4, 6])
    a.add_argument("--nelmt", type=int, default=524288)
    a.add_argument("--ntests", type=int, default=10)
    a.add_argument("out", nargs="?", default="gpu_results.csv")
    a = sub.add_parser("dof-sweep")
    a.add_argument("--op", type=int, choices=(1, 3, 5), required=True)
    a.add_argument("--kernels", nargs="+", required=True, help="p:variant, e.g. 2:v7 4:v6")
    a.add_argument("--dof-min", type=float, default=1e3)
    a.add_argument("--dof-max", type=float, default=1e7)
    a.add_argument("--ppd", type=int, default=4)
    a.add_argument("--ntests", type=int, default=20)
    a.add_argument("out", nargs="?", default="dof_sweep.csv")
    a = sub.add_parser("roofline")
    a.add_argument("csv")
    a.add_argument("--prog", choices=sorted(COST), help="operator of a dof-sweep CSV")
    a.add_argument("--bw", type=float, default=None, help="GB/s (default: bin/stream's copy, else 8000)")
    a.add_argument("--peak", type=float, default=FP64_PEAK, help="GFLOP/s")
    sub.add_parser("stream")
    args = ap.parse_args(argv)

    if args.cmd == "pick":
        rows = pick(args.op, args.orders, args.dofs, args.reps, args.ntests, args.out, args.exclude)
        print("\n".join(pick_summary(rows, args.op, stream())))
    elif args.cmd == "sweep":
        sweep(args.progs, args.orders, args.nelmt, args.ntests, args.out)
    elif args.cmd == "dof-sweep":
        kernels = {int(s.split(":")[0]): s.split(":")[1] for s in args.kernels}
        dof_sweep(args.op, kernels, args.dof_min, args.dof_max, args.ppd, args.ntests, args.out)
    elif args.cmd == "roofline":
        bw = args.bw or stream().get("copy", HBM_PEAK)
        with open(args.csv) as f:
            rows = list(csv.DictReader(f))
        print("\n".join(roofline_table(rows, bw, args.peak, args.prog)))
    else:
        for k, v in stream().items():
            print("%-6s GB/s = %8.2f" % (k, v))


if __name__ == "__main__":
    main()
