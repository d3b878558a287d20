"""The command-line harnesses on the device: the per-order drivers (bin/BK1,
BK3, BK5: openmp/BK*.cpp), the variant-table harnesses (bin/BKn_variants:
openmp/offload/bk*_offload.cpp), bin/stream (roofline/stream_omp.cpp) and
the study tools over them.

The drivers' norm is checked against the oracle's on the same data: the
reference drivers' constant data (cos basis, in = 3, JxW = 1 or G = 2), for
which every element is the same, and the LCG data of BK_RANDOM=1."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import study

pytestmark = pytest.mark.gpu
BIN = study.BIN


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    print("run:", " ".join(args), env or "", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, env=e, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _oracle_out(kind, p, u, Q):
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    if kind == 5:
        return O.bk5(nq, O.test_basis(nq, nq), Q, u)
    B = O.test_basis(nm, nq)
    if kind == 1:
        return O.bk1(nq, B, Q, u)
    return O.bk3(nq, B, O.test_basis(nq, nq), Q, u)


def _norm(text):
    line = [ln for ln in text.splitlines() if ln.startswith("norm = ")]
    assert len(line) == 1, text
    return float(line[0].split()[-1])


@pytest.mark.parametrize("kind,p", [(1, 2), (1, 8), (3, 2), (3, 7), (5, 3), (5, 7)])
def test_driver_norm_constant_data(kind, p):
    E = 1000
    text = _run([os.path.join(BIN, "BK%d" % kind), str(p), str(E), "2"])
    assert study.parse_driver_line(text) is not None
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    nf = 1 if kind == 1 else 6
    one = _oracle_out(kind, p, np.full(nm ** 3, 3.0), np.full(nf * nq ** 3, 1.0 if kind == 1 else 2.0))
    want = np.sqrt(E) * np.linalg.norm(one)
    assert _norm(text) == pytest.approx(want, rel=1e-5)   # printed to 6 digits


@pytest.mark.parametrize("kind,p", [(1, 3), (3, 5), (5, 7)])
def test_driver_norm_random_data(kind, p):
    E = 77
    text = _run([os.path.join(BIN, "BK%d" % kind), str(p), str(E), "1"], env={"BK_RANDOM": "1"})
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    nf = 1 if kind == 1 else 6
    u, Q = O.fill_random(E * nm ** 3, E * nf * nq ** 3)
    want = np.linalg.norm(_oracle_out(kind, p, u, Q))
    assert _norm(text) == pytest.approx(want, rel=1e-5)


def test_driver_rejects_unsupported_order():
    r = subprocess.run([os.path.join(BIN, "BK5"), "8", "10", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unsupported polynomial order p = 8" in r.stderr


@pytest.mark.parametrize("kind,p", [(1, 4), (3, 7), (5, 7)])
def test_variants_harness_every_variant(kind, p):
    text = _run([os.path.join(BIN, "BK%d_variants" % kind), str(p), "3001", "2"])
    assert "host reference on 3001 elements" in text
    rows = study.parse_variant_lines(text)
    count = _lib.lib().bkx_tune_count(kind, 0)
    assert [r[0] for r in rows] == ["v%d" % v for v in range(count)]
    for name, gdofs, gbs, relerr in rows:
        assert relerr != "FAIL" and float(relerr) <= 1e-13, name
        assert gdofs > 0 and gbs > gdofs


def test_variants_harness_strided_reference():
    # beyond 8192 elements every ceil(nelmt / 8192)-th element is checked
    text = _run([os.path.join(BIN, "BK1_variants"), "2", "20000", "1"], env={"BK_KERNEL": "v0"})
    assert "host reference on %d elements" % len(range(0, 20000, 3)) in text
    [(name, _, _, relerr)] = study.parse_variant_lines(text)
    assert name == "v0" and relerr != "FAIL" and float(relerr) <= 1e-13


def test_variants_harness_selection():
    # BK_KERNEL names variants to run, a leading "-" leaves one out
    # (include/bk_harness.h:20-40); every variant is compared with the
    # harness's host FP64 reference
    text = _run([os.path.join(BIN, "BK5_variants"), "3", "500", "1"], env={"BK_KERNEL": "v2 v7 -v7"})
    assert [r[0] for r in study.parse_variant_lines(text)] == ["v2"]
    text = _run([os.path.join(BIN, "BK5_variants"), "3", "500", "1"], env={"BK_KERNEL": "-v0 -v1"})
    names = [r[0] for r in study.parse_variant_lines(text)]
    assert "v0" not in names and "v1" not in names and "v2" in names


def test_stream_roofs():
    roofs = study.parse_stream(_run([os.path.join(BIN, "stream"), str(1 << 24), "3"]))
    assert set(roofs) == {"copy", "read7", "read9"}
    assert all(v > 100.0 for v in roofs.values())


def test_study_pick(tmp_path):
    out = str(tmp_path / "pick.csv")
    rows = study.pick(3, [2], dofs=2e5, reps=2, ntests=1, out=out, exclude="-v3")
    with open(out) as f:
        assert f.readline().strip() == "api,p,variant,rep,gdofs,gbs,relerr"
    assert {r["rep"] for r in rows} == {1, 2}
    assert "v3" not in {r["variant"] for r in rows}
    best, bad = study.best_by_median(rows)
    assert not bad and ("hip", 2) in best


def _run_rc(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    print("run:", " ".join(args), env or "", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, env=e, timeout=120)
    return r.returncode, r.stdout + r.stderr


def _ff_errors(text):
    """The gated errors of the ok/FAIL line, by form name."""
    line = [ln for ln in text.splitlines() if ln.startswith(("ok:", "FAIL:"))]
    assert len(line) == 1, text
    body = line[0].split(":", 1)[1].rsplit("(tol", 1)[0]
    out = {}
    for part in body.split(","):
        name, val = part.split(" rel. error vs double = ")
        out[name.strip()] = float(val)
    return line[0].split(":")[0], out


@pytest.mark.parametrize("kind,p", [(1, 1), (1, 3), (1, 8), (3, 2), (3, 5), (3, 8), (5, 3), (5, 7)])
def test_precision_driver_ok(kind, p):
    """bin/BKn_ff (openmp/BK1_ff.cpp:767-859): every form on the same LCG
    data, the float-float forms within the reference's default BK_TOL of
    1e-10 of the FP64 apply (in fact ~1e-14), exit status 0; the FP64 row's
    norm equals the oracle's."""
    nelmt = 1000
    rc, out = _run_rc([os.path.join(BIN, "BK%d_ff" % kind), str(p), str(nelmt), "2"], {"BK_RANDOM": "1"})
    assert rc == 0, out
    status, errs = _ff_errors(out)
    assert status == "ok"
    expect = {"f32c", "f32w"} | ({"float-float"} if kind != 5 else set())
    assert set(errs) == expect, errs
    assert all(e <= 1e-12 for e in errs.values()), errs
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    u, Q = O.fill_random(nelmt * nm ** 3, nelmt * (1 if kind == 1 else 6) * nq ** 3)
    ref = np.linalg.norm(_oracle_out(kind, p, u, Q))
    row = [ln for ln in out.splitlines() if ln.startswith("double")][0]
    assert abs(float(row.split("norm =")[1]) - ref) <= 1e-9 * ref
    if kind == 5 and p == 7:
        assert "bf16s" in out


def test_precision_driver_fail_exit_status():
    """A tolerance below the float-float error level makes the driver print
    FAIL and exit 1, as the reference's does."""
    rc, out = _run_rc([os.path.join(BIN, "BK1_ff"), "3", "100", "1"], {"BK_RANDOM": "1", "BK_TOL": "1e-300"})
    assert rc == 1, out
    status, _ = _ff_errors(out)
    assert status == "FAIL"


@pytest.mark.parametrize("kind,p", [(1, p) for p in range(1, 9)] + [(3, p) for p in range(1, 9)]
                         + [(5, p) for p in range(1, 8)])
def test_cpp_shim_caller_golden(kind, p, tmp_path):
    """The C++ drop-in (include/bkx_sumfact.hpp) held in the reference's own
    kernel pointer types (sumfact_kernel<double>, openmp/bk1_sumfact.h:
    162-163; BK3's kernel_fn, openmp/offload/bk3_offload.cpp:226) and called
    synchronously on hipMalloc'd golden inputs (bin/bk_shim_check): the
    output against the reference's SumFactorization<double, nq>."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bk%d_p%d.npz" % (kind, p)))
    names = {1: ("basis", "JxW", "u"), 3: ("basis", "dbasis", "G", "u"), 5: ("dbasis", "G", "u")}[kind]
    n = p + 1
    nelmt = g["u"].size // n ** 3
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([np.ascontiguousarray(g[k], dtype=np.float64).reshape(-1) for k in names]).tofile(src)
    rc, out = _run_rc([os.path.join(BIN, "bk_shim_check"), str(kind), str(p), str(nelmt), str(src), str(dst)])
    assert rc == 0, out
    assert "SumFactorization<double, %d>" % (p + 1 if kind == 5 else p + 2) in out
    got = np.fromfile(dst, dtype=np.float64)
    assert got.size == g["sumfact"].size
    assert O.rel_maxdiff(got, g["sumfact"]) <= 1e-13
