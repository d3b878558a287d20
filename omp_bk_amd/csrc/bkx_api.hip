// bkx_api.hip -- the C ABI of include/bkx.h: argument checks, then the
// launch of the operator's kernel on the caller's stream.
#include <hip/hip_runtime.h>

#include <atomic>

#include <cmath>

#include "bkx_eo.hpp"
#include "bkx.h"
#include "bkx_launch.hpp"

namespace bkx {
int bk1_apply_f64(int p, size_t nelmt, const double* B, const double* JxW, const double* u, double* v,
                  hipStream_t s);
int bk3_apply_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* u,
                  double* v, hipStream_t s);
int bk1t_apply_f64(int variant, int p, size_t nelmt, const double* B, const double* JxW, const double* u,
                   double* v, hipStream_t s);
int bk3t_apply_f64(int variant, int p, size_t nelmt, const double* B, const double* D, const double* G,
                   const double* u, double* v, hipStream_t s);
int bk3_apply_kappa_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                        const double* u, double* v, hipStream_t s);
int bk3_adjoint_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                    const double* u, const double* b, double* dldu, double* dldk, double* loss, hipStream_t s);
int bk3_vjp_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                const double* u, const double* w, double* gu, double* gk, hipStream_t s);
int bk3_apply_kappa_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                         const float* u, double* v, hipStream_t s);
int bk3_adjoint_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                     const float* u, const float* b, double* dldu, double* dldk, double* loss, hipStream_t s);
int bk3_vjp_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                 const float* u, const float* w, double* gu, double* gk, hipStream_t s);
int bk3_apply_kappa_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                         const float* u, double* v, hipStream_t s);
int bk3_adjoint_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                     const float* u, const float* b, double* dldu, double* dldk, double* loss, hipStream_t s);
int bk3_vjp_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                 const float* u, const float* w, double* gu, double* gk, hipStream_t s);
int sum_f64(const double* x, size_t n, double* out, hipStream_t s);
int bk3_adjoint_loss_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                         const double* u, const double* b, double* dldu, double* dldk, double* L, hipStream_t s);
int geo_bk3_f64(int p, size_t nelmt, const double* B, const double* D, const double* wq, const double* coords,
                double* G, double* JxW, hipStream_t s);
int geo_bk5_f64(int p, size_t nelmt, const double* D, const double* wq, const double* coords, double* G,
                hipStream_t s);
int bk5_apply_bf16s(int p, size_t nelmt, const double* D, const float* G, const float* u, float* v, hipStream_t s);
int bk5_apply_oz(int p, size_t nelmt, const double* D, const float* G, const float* u, double* v, hipStream_t s);
int bk5_variants();
int bk5_apply_f64(int variant, int p, size_t nelmt, const double* D, const double* G, const double* u, double* v,
                  hipStream_t s);
int bk5_apply_f32c(int p, size_t nelmt, const double* D, const float* G, const float* u, double* v,
                   hipStream_t s);
int bk5_apply_f32w(int variant, int p, size_t nelmt, const double* D, const float* G, const float* u, double* v,
                   hipStream_t s);
int bk13_red_variants(int form);
int bk1_apply_red(int form, int variant, int p, size_t nelmt, const double* B, const void* JxW, const void* u,
                  double* v, hipStream_t s);
int bk3_apply_red(int form, int variant, int p, size_t nelmt, const double* B, const double* D, const void* G,
                  const void* u, double* v, hipStream_t s);
int bk3t_apply_eo_f64(int variant, int p, size_t nelmt, const double* P, const double* G, const double* u, double* v,
                      hipStream_t s);
int bk3_adjoint_eo_f64(int p, size_t nelmt, const double* P, const double* G, const double* kappa, const double* u,
                       const double* b, double* dldu, double* dldk, double* loss, hipStream_t s);
int bk3_adjoint_eo_f32w(int p, size_t nelmt, const double* P, const float* G, const float* kappa, const float* u,
                        const float* b, double* dldu, double* dldk, double* loss, hipStream_t s);
int bk3_apply_f32w_eo(int variant, int p, size_t nelmt, const double* P, const float* G, const float* u, double* v,
                      hipStream_t s);
}  // namespace bkx

namespace {

bool order_ok(int op, int p) { return p >= bkx_min_order(op) && p <= bkx_max_order(op); }

hipStream_t as_stream(bkx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Launch variant of each (operator, form, order).  The defaults are the
// fastest variants measured on an MI355X: the FP64 ones by the variant
// harnesses (`python -m omp_bk_amd.study pick`: median of three runs at
// 5e7 DoFs for BK5, 2e7 for BK1 and BK3), the BK1 and BK3 f32w ones at 2e7
// DoFs too, the others by omp_bk_amd/tune.py (BK5 over E = 65536, BK1 and
// BK3 over 32768 elements); the autotuner overwrites them at run time.
constexpr int kOps = 6;
constexpr int kForms = 5;
constexpr int kOrders = 9;

struct VariantTable {
    std::atomic<int> v[kOps][kForms][kOrders];
    VariantTable()
    {
        for (auto& op : v) {
            for (auto& form : op) {
                for (auto& x : form) {
                    x.store(0, std::memory_order_relaxed);
                }
            }
        }
        //                            p:   0   1   2   3   4   5   6   7   8
        static const int bk5_f64[kOrders] = {0, 27, 18, 35, 34, 36, 38, 36, 0};
        static const int bk5_f32w[kOrders] = {0, 10, 16, 35, 27, 39, 22, 34, 0};
        static const int bk3_f64[kOrders] = {0, 16, 14, 16, 17, 16, 13, 16, 18};
        static const int bk1_f64[kOrders] = {0, 5, 1, 19, 18, 19, 18, 19, 18};
        static const int bk3_f32w[kOrders] = {0, 9, 6, 8, 8, 10, 4, 8, 8};
        static const int bk3_f32c[kOrders] = {0, 2, 2, 2, 0, 2, 1, 0, 0};
        static const int bk3_ff[kOrders] = {0, 2, 2, 2, 2, 2, 0, 0, 1};
        static const int bk1_f32w[kOrders] = {0, 9, 4, 9, 4, 4, 4, 4, 4};
        static const int bk1_f32c[kOrders] = {0, 2, 2, 0, 0, 0, 1, 0, 1};
        static const int bk1_ff[kOrders] = {0, 2, 2, 0, 0, 0, 1, 0, 0};
        for (int p = 0; p < kOrders; ++p) {
            v[5][BKX_FORM_F64][p].store(bk5_f64[p], std::memory_order_relaxed);
            v[5][BKX_FORM_F32W][p].store(bk5_f32w[p], std::memory_order_relaxed);
            v[3][BKX_FORM_F64][p].store(bk3_f64[p], std::memory_order_relaxed);
            v[1][BKX_FORM_F64][p].store(bk1_f64[p], std::memory_order_relaxed);
            v[3][BKX_FORM_F32W][p].store(bk3_f32w[p], std::memory_order_relaxed);
            v[3][BKX_FORM_F32C][p].store(bk3_f32c[p], std::memory_order_relaxed);
            v[3][BKX_FORM_FF][p].store(bk3_ff[p], std::memory_order_relaxed);
            v[1][BKX_FORM_F32W][p].store(bk1_f32w[p], std::memory_order_relaxed);
            v[1][BKX_FORM_F32C][p].store(bk1_f32c[p], std::memory_order_relaxed);
            v[1][BKX_FORM_FF][p].store(bk1_ff[p], std::memory_order_relaxed);
        }
    }
};
VariantTable g_table;
auto& g_variant = g_table.v;

bool tune_index_ok(int op, int form, int p)
{
    return (op == 1 || op == 3 || op == 5) && form >= 0 && form < kForms && p >= 0 && p < kOrders;
}

int variant_of(int op, int form, int p) { return g_variant[op][form][p].load(std::memory_order_relaxed); }

// The even-odd form (bkx::EoMat) of the row-major R x C matrix M,
// after checking M[R-1-a][C-1-b] = S M[a][b] to 1e-14 of max |M|.
template <int R, int C, int S>
bool eo_form(const double* M, double* out)
{
    using E = bkx::EoMat<R, C, S>;
    double big = 0.0;
    for (int x = 0; x < R * C; ++x) {
        big = std::fmax(big, std::fabs(M[x]));
    }
    for (int a = 0; a < R; ++a) {
        for (int b = 0; b < C; ++b) {
            if (!(std::fabs(M[(R - 1 - a) * C + (C - 1 - b)] - S * M[a * C + b]) <= 1e-14 * big)) {
                return false;
            }
        }
    }
    for (int m = 0; m < E::hr; ++m) {
        for (int c = 0; c < E::cc; ++c) {
            out[E::cols_p + m * E::cc + c] = 0.5 * (M[m * C + c] + M[(R - 1 - m) * C + c]);
            out[E::cols_q + m * E::cc + c] = 0.5 * (M[m * C + c] - M[(R - 1 - m) * C + c]);
        }
    }
    if (R % 2 == 1) {
        for (int c = 0; c < E::cc; ++c) {
            out[E::cols_mid + c] = M[(R / 2) * C + c];
        }
    }
    for (int m = 0; m < E::hc; ++m) {
        for (int c = 0; c < E::cr; ++c) {
            out[E::rows_p + m * E::cr + c] = 0.5 * (M[c * C + m] + M[c * C + (C - 1 - m)]);
            out[E::rows_q + m * E::cr + c] = 0.5 * (M[c * C + m] - M[c * C + (C - 1 - m)]);
        }
    }
    if (C % 2 == 1) {
        for (int c = 0; c < E::cr; ++c) {
            out[E::rows_mid + c] = M[c * C + C / 2];
        }
    }
    return true;
}

}  // namespace

extern "C" {

const char* bkx_status_string(int status)
{
    switch (status) {
    case BKX_SUCCESS:
        return "success";
    case BKX_ERROR_ORDER:
        return "unsupported polynomial order p";
    case BKX_ERROR_NULL:
        return "null pointer argument";
    case BKX_ERROR_LAUNCH:
        return "kernel launch failed";
    case BKX_ERROR_SIZE:
        return "too many elements for one launch";
    case BKX_ERROR_ARG:
        return "invalid argument";
    case BKX_ERROR_DEVICE:
        return "HIP device error";
    default:
        return "unknown status";
    }
}

int bkx_version(void) { return 100; }

int bkx_tune_count(int op, int form)
{
    if (op == 5 && (form == BKX_FORM_F64 || form == BKX_FORM_F32W)) {
        return bkx::bk5_variants();
    }
    if ((op == 1 || op == 3) && (form == BKX_FORM_F32W || form == BKX_FORM_F32C || form == BKX_FORM_FF)) {
        return bkx::bk13_red_variants(form);   // bk13t_red.hip
    }
    if ((op == 1 || op == 3) && form == BKX_FORM_F64) {
        // transposing mapping (bk13t.hip): 0 all elements per workgroup,
        // 1 half, 2 all with a scheduling barrier per row (three LDS boxes
        // per element), 4 all, 5 half (two LDS boxes), 6 all, 7 half, 8 all
        // (two boxes, BK3 with three, three and two rows of G loaded ahead;
        // BK1: as 4, 5, 4); 3: LDS-contraction mapping (bk13_f64.hip);
        // 9-12: 6, 8, 7, 4 with non-temporal streams; 13-15: more rows of
        // G loaded ahead; 16-19: 9-12 with G read by ordinary loads;
        // 20-23: persistent pipelined BK3
        return 24;
    }
    return (op == 1 || op == 3 || op == 5) && form >= 0 && form < kForms ? 1 : -1;
}

int bkx_tune_set(int op, int form, int p, int variant)
{
    if (!tune_index_ok(op, form, p)) {
        return BKX_ERROR_ARG;
    }
    if (variant < 0 || variant >= bkx_tune_count(op, form)) {
        return BKX_ERROR_ARG;
    }
    g_variant[op][form][p].store(variant, std::memory_order_relaxed);
    return BKX_SUCCESS;
}

int bkx_tune_get(int op, int form, int p)
{
    return tune_index_ok(op, form, p) ? variant_of(op, form, p) : -1;
}

int bkx_min_order(int op) { return (op == 1 || op == 3 || op == 5) ? 1 : -1; }

int bkx_max_order(int op)
{
    if (op == 1 || op == 3) {
        return 8;
    }
    return op == 5 ? 7 : -1;
}

int bkx_bk1_apply_f64(int p, size_t nelmt, const double* basis, const double* JxW, const double* in, double* out,
                      bkx_stream_t stream)
{
    if (!order_ok(1, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !JxW || !in || !out) {
        return BKX_ERROR_NULL;
    }
    const int var = variant_of(1, BKX_FORM_F64, p);
    if (var == 3) {
        return bkx::bk1_apply_f64(p, nelmt, basis, JxW, in, out, as_stream(stream));
    }
    return bkx::bk1t_apply_f64(var, p, nelmt, basis, JxW, in, out, as_stream(stream));
}

int bkx_bk3_apply_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                      const double* in, double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    const int var = variant_of(3, BKX_FORM_F64, p);
    if (var == 3) {
        return bkx::bk3_apply_f64(p, nelmt, basis, dbasis, G, in, out, as_stream(stream));
    }
    return bkx::bk3t_apply_f64(var, p, nelmt, basis, dbasis, G, in, out, as_stream(stream));
}

}  // extern "C"

namespace {

int bk1_red(int form, int p, size_t nelmt, const double* basis, const void* JxW, const void* in, double* out,
            bkx_stream_t stream)
{
    if (!order_ok(1, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !JxW || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk1_apply_red(form, variant_of(1, form, p), p, nelmt, basis, JxW, in, out, as_stream(stream));
}

int bk3_red(int form, int p, size_t nelmt, const double* basis, const double* dbasis, const void* G, const void* in,
            double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_red(form, variant_of(3, form, p), p, nelmt, basis, dbasis, G, in, out, as_stream(stream));
}

}  // namespace

extern "C" {

int bkx_bk1_apply_f32w(int p, size_t nelmt, const double* basis, const float* JxW, const float* in, double* out,
                       bkx_stream_t stream)
{
    return bk1_red(BKX_FORM_F32W, p, nelmt, basis, JxW, in, out, stream);
}

int bkx_bk1_apply_f32c(int p, size_t nelmt, const double* basis, const float* JxW, const float* in, double* out,
                       bkx_stream_t stream)
{
    return bk1_red(BKX_FORM_F32C, p, nelmt, basis, JxW, in, out, stream);
}

int bkx_bk1_apply_ff(int p, size_t nelmt, const double* basis, const double* JxW, const double* in, double* out,
                     bkx_stream_t stream)
{
    return bk1_red(BKX_FORM_FF, p, nelmt, basis, JxW, in, out, stream);
}

int bkx_bk3_apply_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                       const float* in, double* out, bkx_stream_t stream)
{
    return bk3_red(BKX_FORM_F32W, p, nelmt, basis, dbasis, G, in, out, stream);
}

size_t bkx_bk3_basis_size(int p)
{
    if (!order_ok(3, p)) {
        return 0;
    }
    return size_t(bkx::with_order<1, 8>(p, [](auto O) {
        constexpr int NM = decltype(O)::value + 1;
        return bkx::EoMat<NM, NM + 1, 1>::size + bkx::EoMat<NM + 1, NM + 1, -1>::size;
    }));
}

int bkx_bk3_basis_prepare(int p, const double* basis, const double* dbasis, double* prepared)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (!basis || !dbasis || !prepared) {
        return BKX_ERROR_NULL;
    }
    return bkx::with_order<1, 8>(p, [&](auto O) {
        constexpr int NM = decltype(O)::value + 1;
        constexpr int NQ = NM + 1;
        const bool ok = eo_form<NM, NQ, 1>(basis, prepared) &&
                        eo_form<NQ, NQ, -1>(dbasis, prepared + bkx::EoMat<NM, NQ, 1>::size);
        return ok ? int(BKX_SUCCESS) : int(BKX_ERROR_ARG);
    });
}

int bkx_bk3_apply_prepared_f64(int p, size_t nelmt, const double* prepared, const double* G, const double* in,
                               double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!prepared || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3t_apply_eo_f64(variant_of(3, BKX_FORM_F64, p), p, nelmt, prepared, G, in, out, as_stream(stream));
}

int bkx_bk3_apply_prepared_f32w(int p, size_t nelmt, const double* prepared, const float* G, const float* in,
                                double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!prepared || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_f32w_eo(variant_of(3, BKX_FORM_F32W, p), p, nelmt, prepared, G, in, out, as_stream(stream));
}

int bkx_bk3_apply_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                       const float* in, double* out, bkx_stream_t stream)
{
    return bk3_red(BKX_FORM_F32C, p, nelmt, basis, dbasis, G, in, out, stream);
}

int bkx_bk3_apply_ff(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                     const double* in, double* out, bkx_stream_t stream)
{
    return bk3_red(BKX_FORM_FF, p, nelmt, basis, dbasis, G, in, out, stream);
}

int bkx_bk5_apply_f64(int p, size_t nelmt, const double* dbasis, const double* G, const double* in, double* out,
                      bkx_stream_t stream)
{
    if (!order_ok(5, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_f64(variant_of(5, BKX_FORM_F64, p), p, nelmt, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk5_apply_f32c(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                       bkx_stream_t stream)
{
    if (!order_ok(5, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_f32c(p, nelmt, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk5_apply_f32w(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                       bkx_stream_t stream)
{
    if (!order_ok(5, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_f32w(variant_of(5, BKX_FORM_F32W, p), p, nelmt, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk5_apply_bf16s(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, float* out,
                        bkx_stream_t stream)
{
    if (p != 7) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_bf16s(p, nelmt, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk5_apply_oz(int p, size_t nelmt, const double* dbasis, const float* G, const float* in, double* out,
                     bkx_stream_t stream)
{
    if (p < 5 || p > 7) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_oz(p, nelmt, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk3_apply_kappa_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                            const double* kappa, const double* in, double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_kappa_f64(p, nelmt, basis, dbasis, G, kappa, in, out, as_stream(stream));
}

int bkx_bk3_adjoint_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                        const double* kappa, const double* u, const double* b, double* dldu, double* dldkappa,
                        double* loss, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !b || !dldu || !dldkappa || !loss) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_adjoint_f64(p, nelmt, basis, dbasis, G, kappa, u, b, dldu, dldkappa, loss, as_stream(stream));
}

int bkx_bk3_adjoint_prepared_f64(int p, size_t nelmt, const double* prepared, const double* G, const double* kappa,
                                 const double* u, const double* b, double* dldu, double* dldkappa, double* loss,
                                 bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!prepared || !G || !kappa || !u || !b || !dldu || !dldkappa || !loss) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_adjoint_eo_f64(p, nelmt, prepared, G, kappa, u, b, dldu, dldkappa, loss, as_stream(stream));
}

int bkx_bk3_adjoint_prepared_f32w(int p, size_t nelmt, const double* prepared, const float* G, const float* kappa,
                                  const float* u, const float* b, double* dldu, double* dldkappa, double* loss,
                                  bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!prepared || !G || !kappa || !u || !b || !dldu || !dldkappa || !loss) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_adjoint_eo_f32w(p, nelmt, prepared, G, kappa, u, b, dldu, dldkappa, loss, as_stream(stream));
}

int bkx_bk3_vjp_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                    const double* kappa, const double* u, const double* w, double* grad_u, double* grad_kappa,
                    bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !w || !grad_u || !grad_kappa) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_vjp_f64(p, nelmt, basis, dbasis, G, kappa, u, w, grad_u, grad_kappa, as_stream(stream));
}

int bkx_bk3_apply_kappa_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                             const float* kappa, const float* in, double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_kappa_f32w(p, nelmt, basis, dbasis, G, kappa, in, out, as_stream(stream));
}

int bkx_bk3_adjoint_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                         const float* kappa, const float* u, const float* b, double* dldu, double* dldkappa,
                         double* loss, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !b || !dldu || !dldkappa || !loss) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_adjoint_f32w(p, nelmt, basis, dbasis, G, kappa, u, b, dldu, dldkappa, loss, as_stream(stream));
}

int bkx_bk3_vjp_f32w(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                     const float* kappa, const float* u, const float* w, double* grad_u, double* grad_kappa,
                     bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !w || !grad_u || !grad_kappa) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_vjp_f32w(p, nelmt, basis, dbasis, G, kappa, u, w, grad_u, grad_kappa, as_stream(stream));
}

int bkx_bk3_apply_kappa_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                             const float* kappa, const float* in, double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_kappa_f32c(p, nelmt, basis, dbasis, G, kappa, in, out, as_stream(stream));
}

int bkx_bk3_adjoint_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                         const float* kappa, const float* u, const float* b, double* dldu, double* dldkappa,
                         double* loss, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !b || !dldu || !dldkappa || !loss) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_adjoint_f32c(p, nelmt, basis, dbasis, G, kappa, u, b, dldu, dldkappa, loss, as_stream(stream));
}

int bkx_bk3_vjp_f32c(int p, size_t nelmt, const double* basis, const double* dbasis, const float* G,
                     const float* kappa, const float* u, const float* w, double* grad_u, double* grad_kappa,
                     bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !kappa || !u || !w || !grad_u || !grad_kappa) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_vjp_f32c(p, nelmt, basis, dbasis, G, kappa, u, w, grad_u, grad_kappa, as_stream(stream));
}

int bkx_geo_bk3_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* qweights,
                    const double* coords, double* G, double* JxW, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !qweights || !coords || (!G && !JxW)) {
        return BKX_ERROR_NULL;
    }
    return bkx::geo_bk3_f64(p, nelmt, basis, dbasis, qweights, coords, G, JxW, as_stream(stream));
}

int bkx_geo_bk5_f64(int p, size_t nelmt, const double* dbasis, const double* qweights, const double* coords,
                    double* G, bkx_stream_t stream)
{
    if (!order_ok(5, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !qweights || !coords || !G) {
        return BKX_ERROR_NULL;
    }
    return bkx::geo_bk5_f64(p, nelmt, dbasis, qweights, coords, G, as_stream(stream));
}

int bkx_bk3_adjoint_loss_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                             const double* kappa, const double* u, const double* b, double* dldu, double* dldkappa,
                             double* L, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (!L || (nelmt > 0 && (!basis || !dbasis || !G || !kappa || !u || !b || !dldu || !dldkappa))) {
        return BKX_ERROR_NULL;
    }
    if (nelmt > BKX_ADJOINT_LOSS_MAX) {
        return BKX_ERROR_SIZE;
    }
    return bkx::bk3_adjoint_loss_f64(p, nelmt, basis, dbasis, G, kappa, u, b, dldu, dldkappa, L, as_stream(stream));
}

int bkx_sum_f64(const double* x, size_t n, double* out, bkx_stream_t stream)
{
    if (!out || (n > 0 && !x)) {
        return BKX_ERROR_NULL;
    }
    return bkx::sum_f64(x, n, out, as_stream(stream));
}

}  // extern "C"
