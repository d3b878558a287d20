"""The study tools on the CPU: the cost model, the parsing of the harnesses'
lines (in the reference harness's format), the median picker and the
roofline table."""
import math

import pytest

from omp_bk_amd import study


def test_cost_model_matches_reference_counts():
    # scripts/plot/plot_roofline.py:46-53 in floats (word = 4): BK1 p = 2 has
    # nm = 3, nq = 4: 4 (4*27 + 16*9 + 64*3) + 64 = 1840 flops per 27 modes
    fl, by = study.bk1_cost(2, word=4)
    assert fl == pytest.approx(1840 / 27, rel=1e-15)
    assert by == pytest.approx(4 * (2 + 64 / 27), rel=1e-15)
    # BK5 p = 7: 12 * 8 + 20 = 116 flops, (2 + 6) words
    assert study.bk5_cost(7, word=4) == (116, 32)
    assert study.bk5_cost(7) == (116, 64)
    assert study.bk5_cost(7, F=9) == (116, 88)


def test_bk3_cost_by_brute_count():
    # count the sum-factorised BK3 flops stage by stage for one element: three
    # interpolation stages there and back, three collocated derivatives and
    # their transposes (nq FMAs per point each), 9 FMAs of G and 2 adds
    for p in (1, 4, 7):
        nm, nq = p + 1, p + 2
        fma = 0
        dims = [nm, nm, nm]
        for axis in range(3):          # B: one direction at a time, nm -> nq
            fma += math.prod(dims) * nq
            dims[axis] = nq
        fma *= 2                       # and B^T back
        fma += 2 * 3 * nq ** 3 * nq    # D and D^T in three directions
        fma += 9 * nq ** 3             # G (symmetric 3 x 3) times the gradient
        flops = 2 * fma + 2 * nq ** 3  # an FMA is 2; the three-term sum
        fl, by = study.bk3_cost(p)
        assert fl == pytest.approx(flops / nm ** 3, rel=1e-14)
        assert by == pytest.approx(8 * (2 + 6 * nq ** 3 / nm ** 3), rel=1e-15)


def test_roof_point_memory_bound():
    ai, gf, roof, frac = study.roof_point("BK5", 7, 99.7)
    assert ai == pytest.approx(116 / 64)
    assert roof == pytest.approx(8000 / 64)   # HBM-bound: 125 GDoF/s
    assert frac == pytest.approx(99.7 / 125)
    assert gf == pytest.approx(99.7 * 116)


SAMPLE = """BK3 hip: p = 5, nq = 7, nelmt = 92593 (20000088 DoFs), ntests = 10
v0         GDoF/s =  41.250  GB/s = 3900.1  time =    0.485 ms  relerr = 0.0e+00
v7         GDoF/s =  60.512  GB/s = 5721.2  time =    0.331 ms  relerr = 3.1e-16
v20        GDoF/s = 120.000  GB/s = 9999.0  time =    0.167 ms  relerr = 2.0e-02  FAIL
"""


def test_parse_variant_lines():
    rows = study.parse_variant_lines(SAMPLE)
    assert rows == [("v0", 41.25, 3900.1, "0.0e+00"), ("v7", 60.512, 5721.2, "3.1e-16"),
                    ("v20", 120.0, 9999.0, "FAIL")]


def test_parse_driver_and_stream():
    text = "SumFactorization -> nelmt = 1000 GDoF/s = 12.5 GB/s = 1300.25\nnorm = 42\n"
    assert study.parse_driver_line(text) == (12.5, 1300.25)
    assert study.parse_driver_line("norm = 1\n") is None
    text = ("device 0 (gfx950), n = 2304 doubles (0.00 GB per array)\n"
            "copy   bytes =  3.686e+04  time =     0.010 ms  GB/s =  3686.40\n"
            "read9  bytes =  1.843e+04  time =     0.005 ms  GB/s =  3686.40\n")
    assert study.parse_stream(text) == {"copy": 3686.4, "read9": 3686.4}


def test_best_by_median_skips_failed_variants():
    rows = []
    for rep, (a, b, c) in enumerate([(1.0, 2.0, 9.0), (1.1, 2.2, 9.0), (0.9, 1.8, 9.0)]):
        rows.append({"api": "hip", "p": "5", "variant": "v0", "gdofs": a, "gbs": 10 * a, "relerr": "0"})
        rows.append({"api": "hip", "p": "5", "variant": "v7", "gdofs": b, "gbs": 10 * b, "relerr": "1e-16"})
        rows.append({"api": "hip", "p": "5", "variant": "v20", "gdofs": c, "gbs": 10 * c,
                     "relerr": "FAIL" if rep == 1 else "1e-16"})
    best, bad = study.best_by_median(rows)
    assert best == {("hip", 5): ("v7", 2.0, 20.0)}
    assert bad == {("hip", 5, "v20")}
    lines = study.pick_summary(rows, 3, {"copy": 40.0, "read9": 20.0})
    assert lines[0] == "roofs: copy 40 GB/s, read9 20 GB/s"
    assert "v7" in lines[1] and "50% of copy" in lines[1] and "100% of read9" in lines[1]
    assert lines[-1].startswith("FAIL")


def test_dof_nelmts():
    # dof_sweep.sh: 1e3..1e7 at 4 per decade, nm = 3 -> 17 targets, rounded and deduplicated
    ne = study.dof_nelmts(3)
    assert ne == sorted(set(ne))
    assert ne[0] == round(1e3 / 27) and ne[-1] == round(1e7 / 27)
    assert len(ne) == 17
    assert study.dof_nelmts(9, 1, 10, 1) == [1]


def test_roofline_table_largest_point():
    rows = [{"prog": "BK5", "p": "7", "nelmt": "1000", "gdofs": "50"},
            {"prog": "BK5", "p": "7", "nelmt": "2000", "gdofs": "100"},
            {"prog": "BK3", "p": "5", "nelmt": "2000", "gdofs": "NaN"}]
    lines = study.roofline_table(rows, bw=8000.0)
    assert len(lines) == 3
    assert lines[2].split()[:2] == ["BK5", "7"] and "80%" in lines[2]
    rows = [{"api": "hip", "nq": "4", "dofs": "27000", "gdofs": "10"}]
    assert study.roofline_table(rows, prog="BK1")[2].split()[:2] == ["BK1", "2"]
