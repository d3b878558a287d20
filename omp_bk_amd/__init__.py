"""omp_bk_amd -- MI355X (gfx950) sum-factorised CEED bake-off operators.

BK1 (mass), BK3 (Poisson) and BK5 (collocated Laplacian) on element-local
E-vectors, as hand-written HIP kernels behind the C ABI of include/bkx.h.
``omp_bk_amd.ops`` mirrors the reference's operator interface on torch
tensors, ``omp_bk_amd.geometry`` builds bases, meshes and geometric factors
in FP64."""
