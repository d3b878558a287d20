"""The roof probes behind the bench's measured roofs, through the C ABI:
the copy and the read-heavy mixes give the values they are defined to
give, and the FP64 rate probe's vector chains match the same recurrence
evaluated on the host (its matrix-core products are checked for being
finite only: the probe times them, its values carry no meaning)."""
import numpy as np
import pytest
import torch

from omp_bk_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _st():
    return torch.cuda.current_stream(DEV).cuda_stream


def test_copy_and_read_mix():
    n = 256 * 1000
    a = torch.arange(n * 9, dtype=torch.float64, device=DEV) * 0.5
    c = torch.full((n * 9,), float("nan"), dtype=torch.float64, device=DEV)
    _lib.call("bkx_probe_copy", a.data_ptr(), c.data_ptr(), n * 9, _st())
    torch.cuda.synchronize()
    assert torch.equal(a, c)
    ah = a.cpu().numpy()
    for r in (1, 3, 7, 9):
        nout = n * 9 // r // 256 * 256
        c.fill_(float("nan"))
        _lib.call("bkx_probe_read_mix", r, a.data_ptr(), c.data_ptr(), nout, _st())
        torch.cuda.synchronize()
        want = ah[: nout * r].reshape(nout // 256, r, 256).sum(axis=1).reshape(-1)
        got = c[:nout].cpu().numpy()
        assert np.array_equal(got, want), r


def test_read_probe():
    n = 16384 * 5 + 2   # a partial last workgroup
    a = torch.rand(n, dtype=torch.float64, device=DEV)
    nc = -(-n // 16384) * 8
    c = torch.full((nc + 8,), float("nan"), dtype=torch.float64, device=DEV)
    _lib.call("bkx_probe_read", a.data_ptr(), n, c.data_ptr(), _st())
    torch.cuda.synchronize()
    got = c.cpu().numpy()
    assert np.isnan(got[nc:]).all()          # nothing written past the documented extent
    assert np.isfinite(got[:nc]).all()
    ah = a.cpu().numpy()
    # lane pair t of wavefront w of workgroup g sums its 32 vectors; the
    # wavefront's sums are added by a butterfly: check the totals
    assert abs(got[:nc].sum() - ah.sum()) <= 1e-9 * ah.sum()
    assert abs(got[:nc:2].sum() - ah[0::2].sum()) <= 1e-9 * ah.sum()


def test_fp64_rate_probe():
    n, iters = 512, 3
    out = torch.full((n,), float("nan"), dtype=torch.float64, device=DEV)
    _lib.call("bkx_probe_fp64_rate", 0, iters, out.data_ptr(), n, _st())
    torch.cuda.synchronize()
    tid = np.arange(n) % 256
    a = 1.0 + 1e-6 * tid
    b = 1.0 - 1e-6 * (tid & 7)
    f = [np.full(n, 1e-3 * (k + 1)) for k in range(16)]
    for _ in range(iters * 4):   # 64 FMAs per iteration over 16 chains
        f = [x * b + a for x in f]
    # + the untouched MFMA accumulators and FP32 chains
    want = sum(f) + sum(1e-3 * k for k in range(4))
    want = want + sum(float(np.float32(1e-3) * np.float32(k + 1)) for k in range(16))
    got = out.cpu().numpy()
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    for kind in (1, 2, 3):
        out.fill_(float("nan"))
        _lib.call("bkx_probe_fp64_rate", kind, iters, out.data_ptr(), n, _st())
        torch.cuda.synchronize()
        assert torch.isfinite(out).all(), kind
