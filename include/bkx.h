/*
 * bkx.h -- C ABI of the MI355X (gfx950) sum-factorised bake-off operators.
 *
 * Every entry point takes device pointers owned by the caller, a polynomial
 * order p chosen at run time and a HIP stream (bkx_stream_t is a
 * hipStream_t; NULL is the null stream), enqueues its kernels on that
 * stream and returns a bkx_status without waiting for them.  The layouts are
 * the reference's element-local E-vector layouts (row-major, last index
 * fastest):
 *
 *   BK1/BK3  nm = p + 1 modes, nq = p + 2 Gauss points per direction;
 *            in/out(e, i, j, k) of nm^3 per element;
 *            basis B(i, p) = basis[i * nq + p] (nm x nq);
 *            BK1 JxW(e, p, q, r) (nq^3 per element);
 *            BK3 dbasis D(n, p) = dbasis[n * nq + p] (nq x nq),
 *                G(e, f, p, q, r) factor-major, 6 factors rr rs rt ss st tt.
 *   BK5      nq = p + 1 collocated points per direction;
 *            in/out(e, i, j, k) of nq^3 per element;
 *            dbasis D(i, n) = dbasis[i * nq + n] (row = output point);
 *            G(e, i, j, f, k) interleaved, 6 factors rr rs rt ss st tt.
 *
 * The plain FP64 entry points are the drop-in for the reference's
 * SumFactorization<double, nq> (the reference kernels return void and run
 * synchronously; the C++ shim in bkx_sumfact.hpp restores that).
 *
 * Outputs must not overlap inputs (no in-place apply): the kernels read
 * their inputs while other workgroups write outputs, and some launch
 * variants write an output value more than once (the same value).
 */
#ifndef BKX_H
#define BKX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* bkx_stream_t;

enum bkx_status {
    BKX_SUCCESS = 0,
    BKX_ERROR_ORDER = 1,   /* p outside the operator's orders (the reference prints
                              "unsupported polynomial order p", bk_common.h:163-178) */
    BKX_ERROR_NULL = 2,    /* a required pointer is NULL while nelmt > 0 */
    BKX_ERROR_LAUNCH = 3,  /* the kernel launch failed (hipGetLastError) */
    BKX_ERROR_SIZE = 4,    /* nelmt beyond the largest grid */
    BKX_ERROR_ARG = 5,     /* another invalid argument */
    BKX_ERROR_DEVICE = 6   /* no usable HIP device or a HIP runtime call failed */
};

/* Arithmetic forms of an operator (the tuning table is kept per form). */
enum bkx_form {
    BKX_FORM_F64 = 0,    /* FP64 storage and arithmetic (the drop-in) */
    BKX_FORM_F32C = 1,   /* FP32 storage, float-float compensated arithmetic */
    BKX_FORM_F32W = 2,   /* FP32 storage, FP64 arithmetic */
    BKX_FORM_BF16S = 3,  /* FP32 storage, split-BF16 MFMA products, FP32 accumulation */
    BKX_FORM_FF = 4      /* FP64 storage split into float-float, float-float arithmetic
                            (the reference's BK1_ff kernels, openmp/BK1_ff.cpp:73-723) */
};

/* Human-readable name of a bkx_status. */
const char* bkx_status_string(int status);

/* 100 * major + minor of this library. */
int bkx_version(void);

/* Lowest and highest p of an operator: op = 1, 3 or 5. */
int bkx_min_order(int op);
int bkx_max_order(int op);

/* Launch variants (workgroup shape, elements per workgroup, load kind) of an
 * operator form, and the one its entry points use at order p (0 until set).
 * The autotuner (omp_bk_amd/tune.py) times every variant and sets the best;
 * the reference picks variants with its study scripts instead
 * (scripts/study/bk3_pick.sh).  Setting is process-wide; it is not
 * synchronised with launches already enqueued. */
int bkx_tune_count(int op, int form);
int bkx_tune_set(int op, int form, int p, int variant);
int bkx_tune_get(int op, int form, int p);

/* ---------------------------------------------------------------------- */
/* Plain FP64 operators (the drop-in)                                      */
/* ---------------------------------------------------------------------- */

/* BK1 mass operator, v = B^T diag(JxW) B u per element, p = 1..8.
 * Replaces bk::SumFactorization<double, nq>(nelmt, basis, JxW, in, out)
 * of openmp/bk1_sumfact.h:16-22 (kernel type sumfact_kernel<T>,
 * openmp/bk1_sumfact.h:162-163). */
int bkx_bk1_apply_f64(int p, size_t nelmt, const double* basis, const double* JxW, const double* in,
                      double* out, bkx_stream_t stream);

/* BK3 Poisson operator, v = B^T D^T G D B u per element, p = 1..8.
 * Replaces bk::SumFactorization<double, nq>(nelmt, basis, dbasis, G, in,
 * out) of openmp/bk3_sumfact.h:17-24 (kernel_fn of
 * openmp/offload/bk3_offload.cpp:226). */
int bkx_bk3_apply_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                      const double* in, double* out, bkx_stream_t stream);

/* BK5 collocated Laplacian, v = D^T G D u per element, p = 1..7.
 * Replaces bk::SumFactorization<double, nq>(nelmt, dbasis, G, in, out) of
 * openmp/BK5.cpp:18-24 (kernel_fn of openmp/offload/bk5_offload.cpp:716). */
int bkx_bk5_apply_f64(int p, size_t nelmt, const double* dbasis, const double* G, const double* in,
                      double* out, bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Reduced-storage forms of BK5 (no reference counterpart: the reference    */
/* runs BK5 in float or double only).  u and G are FP32, v is FP64.        */
/* ---------------------------------------------------------------------- */

/* FP32-compensated: float-float arithmetic throughout (D split into
 * hi + lo, Dot2 compensated sums, the scheme of openmp/BK1_ff.cpp); v is
 * joined to FP64.  Error vs the FP64 operator on the same (FP32) inputs is
 * ~1e-14 relative. */
int bkx_bk5_apply_f32c(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                       bkx_stream_t stream);

/* FP32 storage with FP64 arithmetic: the operator of bkx_bk5_apply_f64 on
 * FP32 inputs (36 instead of 64 bytes per DoF). */
int bkx_bk5_apply_f32w(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                       bkx_stream_t stream);

/* Split-BF16 MFMA form, p = 7 only (n = 8 fills the 16 x 16 x 16 MFMA tiles):
 * every contraction operand is split into two BF16 parts (about 16
 * significant bits) and the products are accumulated in FP32 by
 * v_mfma_f32_16x16x16_bf16; fluxes in FP32.  u, G and v are FP32 (32 bytes
 * per DoF).  A reduced-precision form: about 1e-5 relative error. */
int bkx_bk5_apply_bf16s(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, float* out,
                        bkx_stream_t stream);

/* FP32-compensated on the matrix cores, p = 5, 6, 7 (n = 6, 7, 8 points per
 * direction, where the n x n contractions fill the MFMA tiles; the tiles are
 * 8 deep, so n < 8 is zero-padded; BKX_ERROR_ORDER otherwise): every contraction is
 * evaluated from six 8-bit integer digits of each operand (D per row, u and
 * each flux field per element, each digit exact in BF16) whose products are
 * summed per digit level by v_mfma_f32_32x32x16_bf16 (two digits of u or of
 * a flux field per instruction), exactly (every level sum is an integer below
 * 2^24); the levels are recombined once per output in FP64, as are the
 * pointwise fluxes.  u and G FP32, v FP64 (36 bytes per DoF).  Error vs the FP64 operator on the same FP32 inputs ~1e-14 relative.
 * No reference counterpart. */
int bkx_bk5_apply_oz(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                     bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Float-float and reduced-storage forms of BK1 and BK3 (v is FP64)         */
/* ---------------------------------------------------------------------- */

/* Float-float arithmetic on FP64 storage: u and JxW are split into float-
 * float (hi + lo) when loaded, every contraction is a Dot2 compensated sum,
 * v is joined to FP64.  Replaces the reference's float-float BK1 kernels
 * SumFactorizationFF / SumFactorizationFFTeams / SumFactorizationFFPlane
 * (openmp/BK1_ff.cpp:73-241, 265-445, 501-723), which take the FP64
 * interface of bk::SumFactorization<double, nq>.  Error vs the FP64
 * operator ~1e-15 relative. */
int bkx_bk1_apply_ff(int p, size_t nelmt, const double* basis, const double* JxW, const double* in, double* out,
                     bkx_stream_t stream);

/* BK1 with u and JxW stored in FP32: float-float arithmetic (f32c) or FP64
 * arithmetic (f32w); no reference counterpart. */
int bkx_bk1_apply_f32c(int p, size_t nelmt, const double* basis, const float* JxW, const float* in, double* out,
                       bkx_stream_t stream);
int bkx_bk1_apply_f32w(int p, size_t nelmt, const double* basis, const float* JxW, const float* in, double* out,
                       bkx_stream_t stream);

/* BK3 in the same three forms: float-float on FP64 storage (the BK1_ff
 * scheme applied to openmp/bk3_sumfact.h:17-187), and with u and G stored
 * in FP32 with float-float (f32c) or FP64 (f32w) arithmetic. */
int bkx_bk3_apply_ff(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                     const double* in, double* out, bkx_stream_t stream);
int bkx_bk3_apply_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                       const float* in, double* out, bkx_stream_t stream);
int bkx_bk3_apply_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                       const float* in, double* out, bkx_stream_t stream);

/* BK3 with the bases in even-odd form.  For centro-symmetric bases,
 * B(i, p) = B(n-1-i, q-1-p) and D(m, p) = -D(q-1-m, q-1-p) (the GLL-to-Gauss
 * interpolation and the Gauss collocation derivative), every contraction
 * splits into even and odd halves with about 30% fewer operations.  No
 * reference counterpart: the reference contracts with B and D directly.
 *   bkx_bk3_basis_size(p): doubles of the prepared form (0: unsupported p);
 *   bkx_bk3_basis_prepare: the form of the HOST arrays basis and dbasis
 *     (layouts as for bkx_bk3_apply_f64) into the host array prepared;
 *     BKX_ERROR_ARG when the bases are not centro-symmetric to 1e-14 of
 *     their largest entry (the form then represents the symmetrised bases);
 *   bkx_bk3_apply_prepared_{f64,f32w}: as bkx_bk3_apply_{f64,f32w}, with
 *     the prepared form copied to the device in place of basis and dbasis;
 *     the results agree with theirs to rounding (about 1e-16 relative). */
size_t bkx_bk3_basis_size(int p);
int bkx_bk3_basis_prepare(int p, const double* basis, const double* dbasis, double* prepared);
int bkx_bk3_apply_prepared_f64(int p, size_t nelmt, const double* prepared, const double* G, const double* in,
                               double* out, bkx_stream_t stream);
int bkx_bk3_apply_prepared_f32w(int p, size_t nelmt, const double* prepared, const float* G, const float* in,
                                double* out, bkx_stream_t stream);
/* bkx_bk3_adjoint_{f64,f32w} with the bases in the prepared even-odd form. */
int bkx_bk3_adjoint_prepared_f64(int p, size_t nelmt, const double* prepared, const double* G, const double* kappa,
                                 const double* u, const double* b, double* dldu, double* dldkappa, double* loss,
                                 bkx_stream_t stream);
int bkx_bk3_adjoint_prepared_f32w(int p, size_t nelmt, const double* prepared, const float* G, const float* kappa,
                                  const float* u, const float* b, double* dldu, double* dldkappa, double* loss,
                                  bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* BK3 with a coefficient and its adjoint (no reference counterpart; the    */
/* operator is that of bkx_bk3_apply_f64 with each point's six factors      */
/* scaled by kappa(e, p, q, r), laid out like BK1's JxW).                  */
/* ---------------------------------------------------------------------- */

/* out = A(kappa) in. */
int bkx_bk3_apply_kappa_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                            const double* kappa, const double* in, double* out, bkx_stream_t stream);

/* For L = 1/2 |A(kappa) u - b|^2: dldu = A(kappa) r and
 * dldkappa(e,p,q,r) = (grad r) . G (grad u) at each Gauss point, with
 * r = A(kappa) u - b and grad the reference gradient (qr, qs, qt); loss[e] =
 * 1/2 |r_e|^2 per element (sum them with bkx_sum_f64 for L).  One fused
 * kernel; no quadrature-point data is stored. */
int bkx_bk3_adjoint_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                        const double* kappa, const double* u, const double* b, double* dldu, double* dldkappa,
                        double* loss, bkx_stream_t stream);

/* Vector-Jacobian product of v = A(kappa) u for a cotangent w:
 * grad_u = A(kappa) w, grad_kappa = (grad w) . G (grad u). */
int bkx_bk3_vjp_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                    const double* kappa, const double* u, const double* w, double* grad_u, double* grad_kappa,
                    bkx_stream_t stream);

/* The same three with u, b (or w), kappa and G stored in FP32 and FP64
 * arithmetic: the FP64 operator applied exactly to FP32 inputs (half the
 * input bytes); outputs and loss in FP64. */
int bkx_bk3_apply_kappa_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                             const float* kappa, const float* in, double* out, bkx_stream_t stream);
int bkx_bk3_adjoint_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                         const float* kappa, const float* u, const float* b, double* dldu, double* dldkappa,
                         double* loss, bkx_stream_t stream);
int bkx_bk3_vjp_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                     const float* kappa, const float* u, const float* w, double* grad_u, double* grad_kappa,
                     bkx_stream_t stream);

/* The same three with FP32 storage and float-float compensated arithmetic
 * (FP32-compensated): B and D split once into hi + lo floats, every
 * contraction, flux and coefficient product a Dot2 sum or TwoProd of exact
 * floats and float-float values (the scheme of the reference's float-float
 * BK1 kernels, openmp/BK1_ff.cpp:73-241, float_float.h:121-134, applied to
 * the operator, the residual, the loss and both gradients); outputs joined
 * to FP64.  Error vs the FP64 entry points on the same FP32 inputs ~1e-14
 * relative. */
int bkx_bk3_apply_kappa_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                             const float* kappa, const float* in, double* out, bkx_stream_t stream);
int bkx_bk3_adjoint_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                         const float* kappa, const float* u, const float* b, double* dldu, double* dldkappa,
                         double* loss, bkx_stream_t stream);
int bkx_bk3_vjp_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                     const float* kappa, const float* u, const float* w, double* grad_u, double* grad_kappa,
                     bkx_stream_t stream);

/* *out = sum of x[0..n-1] in a fixed order (one workgroup; deterministic). */
int bkx_sum_f64(const double* x, size_t n, double* out, bkx_stream_t stream);

/* bkx_bk3_adjoint_f64 followed by bkx_sum_f64 of its per-element losses, with
 * nothing allocated: L (one double) = 1/2 |A(kappa) u - b|^2, and dldu and
 * dldkappa as bkx_bk3_adjoint_f64, the same bits as that pair of calls.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  The per-element losses live in a library-owned device array of
 * BKX_ADJOINT_LOSS_MAX doubles (8 MiB) on each GPU the library runs on, held
 * until the process ends; nelmt above that limit returns BKX_ERROR_SIZE.
 * Calls that overlap on one GPU share the array and race.  The stream's
 * device (the current device for the NULL stream) is the GPU used. */
#define BKX_ADJOINT_LOSS_MAX ((size_t)1 << 20)
int bkx_bk3_adjoint_loss_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                             const double* kappa, const double* u, const double* b, double* dldu, double* dldkappa,
                             double* L, bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Geometric factors from element nodal coordinates (the step before the    */
/* operators; no reference counterpart: the reference's operators take     */
/* synthetic factors, include/bk_common.h:92-107).                          */
/*                                                                          */
/* coords(e, c, i, j, k): x, y, z (c = 0, 1, 2) at the element's n^3 GLL    */
/* nodes, n = p + 1.  J = d x / d xi of their degree-p interpolant, xi_1,   */
/* xi_2, xi_3 the directions of the first, second, third tensor index;      */
/* G = w detJ J^-1 J^-T (rr rs rt ss st tt) and JxW = w detJ, w the        */
/* product of the three 1D weights.  FP64.                                  */
/* ---------------------------------------------------------------------- */

/* At the q = p + 2 Gauss points: G factor-major, as bkx_bk3_apply_f64 reads
 * it, and JxW(e, p, q, r), as bkx_bk1_apply_f64 reads it; either may be NULL
 * (not both).  basis and dbasis as for bkx_bk3_apply_f64, qweights the q
 * Gauss weights.  p = 1..8. */
int bkx_geo_bk3_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* qweights,
                    const double* coords, double* G, double* JxW, bkx_stream_t stream);

/* At the n GLL nodes (collocated): G interleaved, G(e, i, j, f, k), as
 * bkx_bk5_apply_f64 reads it.  dbasis as for bkx_bk5_apply_f64, qweights the
 * n GLL weights.  p = 1..7. */
int bkx_geo_bk5_f64(int p, size_t nelmt, const double* dbasis, const double* qweights, const double* coords,
                    double* G, bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Layouts of the geometric factors (the reference harness's helpers,      */
/* include/bk_harness.h:87-102, 207-244, on the device; pure copies, so    */
/* the results are bit-identical to the host helpers).                      */
/* ---------------------------------------------------------------------- */

/* BK5, p = 1..7 (nq = p + 1): from the interleaved six factors G6i(e, i, j,
 * s, k) that bkx_bk5_apply_* read, the factors with F = 6 (rr rs rt ss st tt)
 * or F = 9 (the full symmetric tensor, row-major: rr rs rt sr ss st tr ts tt)
 * per point in layout 0 (factor-major G(e, f, i, j, k)), 1 (interleaved
 * G(e, i, j, f, k)) or 2 (column G(e, f, k, i, j)).  Replaces
 * bk::make_bk5_G(G6i, nelmt, nq, F, layout) of include/bk_harness.h:217-244.
 * BKX_ERROR_ARG for another F or layout. */
int bkx_bk5_make_G_f64(int p, size_t nelmt, const double* G6i, int F, int layout, double* G, bkx_stream_t stream);
int bkx_bk5_make_G_f32(int p, size_t nelmt, const float* G6i, int F, int layout, float* G, bkx_stream_t stream);
/* The inverse: factors in (F, layout) as above -> G6i, so that a caller
 * holding factor-major, column or nine-factor G can apply BK5 (for F = 9 the
 * entries on and above the diagonal are taken).  No reference counterpart. */
int bkx_bk5_interleave_G_f64(int p, size_t nelmt, int F, int layout, const double* G, double* G6i,
                             bkx_stream_t stream);
int bkx_bk5_interleave_G_f32(int p, size_t nelmt, int F, int layout, const float* G, float* G6i,
                             bkx_stream_t stream);
/* BK1, p = 1..8 (nq = p + 2): JxW(e, p, q, r) -> JxW_t(e, r, q, p).  Replaces
 * bk::transpose_jxw(JxW, nelmt, nq) of include/bk_harness.h:87-102. */
int bkx_bk1_transpose_jxw_f64(int p, size_t nelmt, const double* JxW, double* JxW_t, bkx_stream_t stream);
int bkx_bk1_transpose_jxw_f32(int p, size_t nelmt, const float* JxW, float* JxW_t, bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Element restriction (no reference counterpart: the reference applies     */
/* its operators to E-vectors only).  offsets[e * elemsize + t] is the     */
/* global node (0 .. lsize - 1) of element e's local node t, as libCEED's   */
/* element-restriction offsets; gather: uE = R uL, scatter: vL = R^T vE.    */
/* R^T A R is then the assembled (continuous Galerkin) operator.           */
/* ---------------------------------------------------------------------- */

typedef struct bkx_restrict bkx_restrict;

/* Copies the host array offsets (nelmt * elemsize) to the device with its
 * transpose (each node's E-vector positions in ascending order).  Checks
 * every offset against lsize. */
int bkx_restrict_create(size_t nelmt, int elemsize, const int32_t* offsets, size_t lsize, bkx_restrict** plan);
int bkx_restrict_destroy(bkx_restrict* plan);
int bkx_restrict_info(const bkx_restrict* plan, size_t* nelmt, int* elemsize, size_t* lsize);
/* uE[x] = uL[offsets[x]] (uL: lsize, uE: nelmt * elemsize values). */
int bkx_restrict_gather_f64(const bkx_restrict* plan, const double* uL, double* uE, bkx_stream_t stream);
/* vL[g] = sum of vE[x] over offsets[x] = g, x ascending: deterministic, and
 * equal bit for bit to the sequential accumulation (vL is overwritten). */
int bkx_restrict_scatter_f64(const bkx_restrict* plan, const double* vE, double* vL, bkx_stream_t stream);
/* BK5 of the restricted L-vector, vE = A_E (R uL), with the gather fused
 * into the operator's loads of u (no E-vector of u is written): the same
 * bits as bkx_restrict_gather_f64 followed by bkx_bk5_apply_f64.  The plan's
 * elements are (p + 1)^3 GLL nodes (else BKX_ERROR_ARG); G and vE as for
 * bkx_bk5_apply_f64 over its nelmt elements.  No reference counterpart. */
int bkx_bk5_apply_restricted_f64(const bkx_restrict* plan, int p, const double* dbasis, const double* G,
                                 const double* uL, double* vE, bkx_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Mixed-order BK3 batches (no reference counterpart: the reference runs    */
/* one order per binary, include/bk_common.h:143-178).                     */
/*                                                                          */
/* A batch holds nelmt elements in the caller's order, element e of order   */
/* p_e in 1..8 (n = p + 1, q = p + 2): its n^3 values of u / v start at     */
/* uoff[e] = sum_{e' < e} n_e'^3 and its 6 q^3 factors of G (factor-major,  */
/* as bkx_bk3_apply_f64) at goff[e] = sum_{e' < e} 6 q_e'^3 (or, for a      */
/* grouped plan, grouped by order).  `bases` packs B_p (n x q) then D_p     */
/* (q x q) for p = 1, 2, ..., 8 in turn.                                    */
/* ---------------------------------------------------------------------- */

typedef struct bkx_mixed_plan bkx_mixed_plan;
typedef struct bkx_graph bkx_graph;

/* Builds a plan (host work plus a few small device arrays): the elements
 * sorted by order and cut into tiles of rounds_by_order[p] x epb elements,
 * epb = floor(256 / q^2) / div_by_order[p] (div 1, 2 or 4, at least one
 * element; BKX_ERROR_ARG for another value), kept both grouped by order
 * and in launch order (by the memory position of the tile's first element,
 * last first).  div_by_order and rounds_by_order are indexed by the order:
 * each holds 9 entries, [0] unused and [p] for p = 1..8 (a shorter array is
 * read out of bounds); either may be NULL for all 1.  grid <= 0 picks the
 * number of resident workgroups of the device.  The single launch is
 * persistent (grid workgroups, each taking the next tile from a queue held
 * by the plan) when some rounds_by_order[p] > 1, or when grid > 0 and
 * rounds_by_order is NULL; otherwise it runs one workgroup per tile.  The
 * queue is two counters that the launch leaves at zero: persistent launches
 * with one plan (direct or replayed from a graph) must not overlap. */
int bkx_mixed_plan_create(const int32_t* order, size_t nelmt, const int32_t* div_by_order,
                          const int32_t* rounds_by_order, int grid, bkx_mixed_plan** plan);
/* The same plan for the grouped layout: each order's elements contiguous
 * (ascending element index, orders ascending), each order's range of u/v
 * and of G starting on a 256-byte boundary; the offsets (with the padding
 * between the ranges) are those of bkx_mixed_plan_offsets. */
int bkx_mixed_plan_create_grouped(const int32_t* order, size_t nelmt, const int32_t* div_by_order,
                                  const int32_t* rounds_by_order, int grid, bkx_mixed_plan** plan);
int bkx_mixed_plan_destroy(bkx_mixed_plan* plan);
/* Sizes of u/v (ndof) and G (ngeo) in doubles, padding included, the tile
 * count and grid. */
int bkx_mixed_plan_info(const bkx_mixed_plan* plan, size_t* ndof, size_t* ngeo, int* ntiles, int* grid);
/* The element offsets uoff, goff (host arrays of nelmt). */
int bkx_mixed_plan_offsets(const bkx_mixed_plan* plan, int64_t* uoff, int64_t* goff);
/* The launch shape of order p's kernel in the per-order launches
 * (bkx_bk3_mixed_apply_by_order*, _apply_order, by-order graphs made after
 * the call): div, the elements-per-workgroup divisor of floor(256 / q^2)
 * (1, 2 or 4), and gpf, the rows of G kept in flight (2, 3 or 4); the
 * defaults are div_by_order[p] and 3.  The tuner (mixed.tune) searches them
 * per order.  The results are the same bits for every shape.  Host-side
 * only; BKX_ERROR_ARG for another div or gpf. */
int bkx_mixed_plan_set_order_shape(bkx_mixed_plan* plan, int p, int div, int gpf);
int bkx_mixed_plan_get_order_shape(const bkx_mixed_plan* plan, int p, int* div, int* gpf);

/* One launch over every tile: one workgroup per tile (default), or
 * persistent when the plan was made so (see bkx_mixed_plan_create).  The
 * array sizes cannot be checked here: bases must hold the 708 doubles of the
 * packed bases (sum over p = 1..8 of (p+1)(p+2) + (p+2)^2), u and v ndof and
 * G ngeo doubles of bkx_mixed_plan_info. */
int bkx_bk3_mixed_apply_f64(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in,
                            double* out, bkx_stream_t stream);
/* One launch per order, each by a kernel compiled for that order alone. */
int bkx_bk3_mixed_apply_by_order_f64(const bkx_mixed_plan* plan, const double* bases, const double* G,
                                     const double* in, double* out, bkx_stream_t stream);
/* The launch of one order's elements only (used by the tuner). */
int bkx_bk3_mixed_apply_order_f64(const bkx_mixed_plan* plan, int p, const double* bases, const double* G,
                                  const double* in, double* out, bkx_stream_t stream);
/* The single launch (by_order = 0) or the per-order launches captured in a
 * hipGraph for the given arrays; replay with bkx_graph_launch. */
int bkx_bk3_mixed_graph_create(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in,
                               double* out, int by_order, bkx_graph** graph);
/* The same with the bases of every order in even-odd form (see
 * bkx_bk3_basis_prepare): bkx_bk3_mixed_prepare builds the forms of the
 * packed HOST bases (layout as above) into the HOST array prepared of
 * bkx_bk3_mixed_prepared_size() doubles, BKX_ERROR_ARG if an order's bases
 * are not centro-symmetric; the caller copies it to the device. */
size_t bkx_bk3_mixed_prepared_size(void);
int bkx_bk3_mixed_prepare(const double* bases, double* prepared);
int bkx_bk3_mixed_apply_prepared_f64(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                     const double* in, double* out, bkx_stream_t stream);
int bkx_bk3_mixed_apply_by_order_prepared_f64(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                              const double* in, double* out, bkx_stream_t stream);
int bkx_bk3_mixed_graph_create_prepared(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                        const double* in, double* out, int by_order, bkx_graph** graph);
int bkx_graph_launch(bkx_graph* graph, bkx_stream_t stream);
int bkx_graph_destroy(bkx_graph* graph);

/* ---------------------------------------------------------------------- */
/* Probes (the roofs the operators are measured against)                   */
/* ---------------------------------------------------------------------- */

/* c = a for n doubles (n even).  HIP counterpart of the copy roof of
 * roofline/stream_omp.cpp. */
int bkx_probe_copy(const double* a, double* c, size_t n, bkx_stream_t stream);

/* c[b*256 + l] = sum_{j<R} a[(b*R + j)*256 + l] for n_out doubles (a multiple
 * of 256), R = reads_per_write in {1, 3, 7, 9}: reads R bytes per byte
 * written.  R = 9 is the reference's read9 roof (roofline/stream_omp.cpp:
 * 101-117); R = 7 matches BK5 in FP64 (56 B read, 8 B written per DoF). */
int bkx_probe_read_mix(int reads_per_write, const double* a, double* c, size_t n_out, bkx_stream_t stream);

/* Reads n doubles (n even) once, as 16-byte non-temporal loads, and writes
 * c[0 .. ceil(n / 16384) * 8 - 1] (one 16-byte sum per wavefront of 4096
 * values): the HBM read ceiling, which the read-heavy mixes above stay
 * below.  No reference counterpart. */
int bkx_probe_read(const double* a, size_t n, double* c, bkx_stream_t stream);

/* FP64 (and FP32) arithmetic issue rates (no reference counterpart: the reference's
 * roofs are bandwidth only).  n / 256 workgroups of 256 threads each run
 * `iters` iterations of, per wavefront,
 *   kind 0: 64 v_fma_f64 (128 flops each),
 *   kind 1: 4 v_mfma_f64_16x16x4_f64 (2048 flops each),
 *   kind 2: both, in one instruction stream,
 *   kind 3: 64 non-packed v_fma_f32 (128 flops each),
 * and write one value per thread to out (n doubles, a multiple of 256).
 * Flops per launch: n / 64 * iters * (8192 kinds 0, 1 and 3, 16384 kind 2). */
int bkx_probe_fp64_rate(int kind, int iters, double* out, size_t n, bkx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* BKX_H */
