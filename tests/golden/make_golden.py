"""Writes the golden vectors of tests/golden/ from the reference's own
kernels, compiled from /root/reference by oracle/Makefile into
oracle/_ref/libbkref.so.  Run in the build container (where the reference
sources exist):  make -C oracle ref && python tests/golden/make_golden.py

Each file holds the inputs and the reference's outputs of one operator at
one order: the cos test basis in float (bk_common.h:63-86), LCG data
(bk_common.h:92-107, seed 12345) rounded to float so that the harness
references (float inputs) see the same values, and the outputs of
  sumfact    bk::SumFactorization<double, nq>           (openmp/*sumfact.h, BK5.cpp)
  harness    bk::reference_bk1/3/5 (FP64 on float inputs) (include/bk_harness.h)
  ff         BK1_ff's float-float kernels (BK1 only)     (openmp/BK1_ff.cpp)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402

NELMT = 3


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libbkref.so is missing: run `make -C oracle ref` first")
    for p in range(1, 9):
        nq, nm = p + 2, p + 1
        B = O.test_basis(nm, nq, f32=True)
        D = O.test_basis(nq, nq, f32=True)
        u, J = O.fill_random(NELMT * nm ** 3, NELMT * nq ** 3)
        J = f32(J)
        np.savez_compressed(os.path.join(HERE, "bk1_p%d.npz" % p), basis=B, JxW=J, u=u,
                            sumfact=O.ref_bk1(p, B, J, u), harness=O.ref_harness(1, nq, B, J, u),
                            ff=O.ref_bk1_ff(1, p, B, J, u))
        u, G = O.fill_random(NELMT * nm ** 3, NELMT * 6 * nq ** 3)
        G = f32(G)
        np.savez_compressed(os.path.join(HERE, "bk3_p%d.npz" % p), basis=B, dbasis=D, G=G, u=u,
                            sumfact=O.ref_bk3(p, B, D, G, u), harness=O.ref_harness(3, nq, B, D, G, u))
    for p in range(1, 8):
        nq = p + 1
        D = O.test_basis(nq, nq, f32=True)
        u, G = O.fill_random(NELMT * nq ** 3, NELMT * 6 * nq ** 3)
        G = f32(G)
        np.savez_compressed(os.path.join(HERE, "bk5_p%d.npz" % p), dbasis=D, G=G, u=u,
                            sumfact=O.ref_bk5(p, D, G, u), harness=O.ref_harness(5, nq, D, G, u))
    print("wrote golden vectors to", HERE)


if __name__ == "__main__":
    main()
