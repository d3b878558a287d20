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
    rows = study.parse_variant_lines(text)
    count = _lib.lib().bkx_tune_count(kind, 0)
    assert [r[0] for r in rows] == ["v%d" % v for v in range(count)]
    for name, gdofs, gbs, relerr in rows:
        assert relerr != "FAIL" and float(relerr) <= 1e-13, name
        assert gdofs > 0 and gbs > gdofs


def test_variants_harness_selection():
    # BK_KERNEL names variants to run, a leading "-" leaves one out
    # (include/bk_harness.h:20-40); variant 0 runs regardless as the comparison
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
