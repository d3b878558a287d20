// restrict.hip -- element restriction around the element-local operators:
// the L-vector (one value per global node, continuous across element
// faces) to the E-vector the operators read (gather, u_E = R u_L) and back
// (scatter-add, v_L = R^T v_E), so that R^T A_E R is the assembled operator
// of a continuous Galerkin discretisation.  For BK5 the gather can also run
// inside the operator's loads (bkx_bk5_apply_restricted_f64).  No reference counterpart: the
// reference applies its operators to E-vectors only.
//
// offsets[x], x = e * elemsize + t, is the global node of element e's local
// node t (t = (i n + j) n + k), as libCEED's element restriction offsets.
// The plan holds them on the device together with their transpose: for each
// global node g, the E-vector positions x with offsets[x] = g in ascending
// order (a CSR structure).  The scatter sums each node's contributions in
// that order, one thread per node: deterministic, with no atomics, and
// bit-identical to the sequential loop v_L[offsets[x]] += v_E[x] over x
// ascending.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "bkx.h"
#include "bkx_launch.hpp"

struct bkx_restrict {
    size_t nelmt = 0;
    int elemsize = 0;
    size_t lsize = 0;
    int32_t* offsets = nullptr;   // nelmt * elemsize
    int32_t* row = nullptr;       // lsize + 1
    int32_t* col = nullptr;       // nelmt * elemsize, E-vector positions by node
};

namespace bkx {

int bk5_apply_gather_f64(int variant, int p, size_t nelmt, const double* D, const double* G, const int32_t* uidx,
                         const double* uL, double* v, hipStream_t s);   // bk5.hip

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void gather_kernel(size_t n, const int32_t* __restrict__ offsets,
                                                          const double* __restrict__ uL, double* __restrict__ uE)
{
    const size_t x = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (x < n) {
        uE[x] = uL[offsets[x]];
    }
}

// Two contributions per step: their positions, then their values, are
// loaded together (unconditionally, the positions clamped to the node's
// last one), and the values are added one at a time in ascending order, so
// that a node's sum takes one dependent pair of loads per two contributions
// instead of one per contribution and keeps the sequential loop's bits.
// Measured on the 64 x 32 x 32 box mesh at p = 7: 0.227 ms with one
// contribution per step, 0.186 with two, 0.191 with four, 0.275 with eight.
constexpr int kScatterStep = 2;

__global__ __launch_bounds__(kThreads) void scatter_kernel(size_t lsize, const int32_t* __restrict__ row,
                                                           const int32_t* __restrict__ col,
                                                           const double* __restrict__ vE, double* __restrict__ vL)
{
    const size_t g = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (g < lsize) {
        const int32_t t0 = row[g];
        const int32_t t1 = row[g + 1];
        double s = 0.0;
        for (int32_t t = t0; t < t1; t += kScatterStep) {
            int32_t c[kScatterStep];
            double x[kScatterStep];
#pragma unroll
            for (int k = 0; k < kScatterStep; ++k) {
                c[k] = col[t + k < t1 ? t + k : t1 - 1];
            }
#pragma unroll
            for (int k = 0; k < kScatterStep; ++k) {
                x[k] = vE[c[k]];
            }
#pragma unroll
            for (int k = 0; k < kScatterStep; ++k) {
                if (t + k < t1) {
                    s += x[k];
                }
            }
        }
        vL[g] = s;
    }
}

template <typename T>
int upload(T** dst, const std::vector<T>& src)
{
    if (hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(1, src.size()) * sizeof(T)) != hipSuccess) {
        return BKX_ERROR_DEVICE;
    }
    if (!src.empty() && hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        return BKX_ERROR_DEVICE;
    }
    return BKX_SUCCESS;
}

}  // namespace
}  // namespace bkx

extern "C" {

int bkx_restrict_create(size_t nelmt, int elemsize, const int32_t* offsets, size_t lsize, bkx_restrict** plan)
{
    if (!plan || (nelmt > 0 && !offsets)) {
        return BKX_ERROR_NULL;
    }
    *plan = nullptr;
    if (elemsize < 1) {
        return BKX_ERROR_ARG;
    }
    const size_t n = nelmt * size_t(elemsize);
    if (n > size_t(INT32_MAX) || lsize > size_t(INT32_MAX) - 1) {
        return BKX_ERROR_SIZE;
    }
    // counting sort of the E-vector positions by node (ascending within a node)
    std::vector<int32_t> row(lsize + 1, 0);
    for (size_t x = 0; x < n; ++x) {
        const int32_t g = offsets[x];
        if (g < 0 || size_t(g) >= lsize) {
            return BKX_ERROR_ARG;
        }
        ++row[size_t(g) + 1];
    }
    for (size_t g = 0; g < lsize; ++g) {
        row[g + 1] += row[g];
    }
    std::vector<int32_t> col(n);
    std::vector<int32_t> next(row.begin(), row.end() - 1);
    for (size_t x = 0; x < n; ++x) {
        col[size_t(next[size_t(offsets[x])]++)] = int32_t(x);
    }
    auto* pl = new (std::nothrow) bkx_restrict;
    if (!pl) {
        return BKX_ERROR_DEVICE;
    }
    pl->nelmt = nelmt;
    pl->elemsize = elemsize;
    pl->lsize = lsize;
    const std::vector<int32_t> off(offsets, offsets + n);
    int st = bkx::upload(&pl->offsets, off);
    if (st == BKX_SUCCESS) {
        st = bkx::upload(&pl->row, row);
    }
    if (st == BKX_SUCCESS) {
        st = bkx::upload(&pl->col, col);
    }
    if (st != BKX_SUCCESS) {
        bkx_restrict_destroy(pl);
        return st;
    }
    *plan = pl;
    return BKX_SUCCESS;
}

int bkx_restrict_destroy(bkx_restrict* plan)
{
    if (!plan) {
        return BKX_SUCCESS;
    }
    for (void* ptr : {static_cast<void*>(plan->offsets), static_cast<void*>(plan->row), static_cast<void*>(plan->col)}) {
        if (ptr) {
            (void)hipFree(ptr);
        }
    }
    delete plan;
    return BKX_SUCCESS;
}

int bkx_restrict_info(const bkx_restrict* plan, size_t* nelmt, int* elemsize, size_t* lsize)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (nelmt) {
        *nelmt = plan->nelmt;
    }
    if (elemsize) {
        *elemsize = plan->elemsize;
    }
    if (lsize) {
        *lsize = plan->lsize;
    }
    return BKX_SUCCESS;
}

int bkx_restrict_gather_f64(const bkx_restrict* plan, const double* uL, double* uE, bkx_stream_t stream)
{
    if (!plan || !uL || !uE) {
        return BKX_ERROR_NULL;
    }
    const size_t n = plan->nelmt * size_t(plan->elemsize);
    if (n == 0) {
        return BKX_SUCCESS;
    }
    const unsigned grid = bkx::grid_for(n, bkx::kThreads);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL(bkx::gather_kernel, dim3(grid), dim3(bkx::kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                       n, plan->offsets, uL, uE);
    return bkx::launch_status();
}

int bkx_restrict_scatter_f64(const bkx_restrict* plan, const double* vE, double* vL, bkx_stream_t stream)
{
    if (!plan || !vE || !vL) {
        return BKX_ERROR_NULL;
    }
    if (plan->lsize == 0) {
        return BKX_SUCCESS;
    }
    const unsigned grid = bkx::grid_for(plan->lsize, bkx::kThreads);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL(bkx::scatter_kernel, dim3(grid), dim3(bkx::kThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), plan->lsize, plan->row, plan->col, vE, vL);
    return bkx::launch_status();
}

int bkx_bk5_apply_restricted_f64(const bkx_restrict* plan, int p, const double* dbasis, const double* G,
                                 const double* uL, double* vE, bkx_stream_t stream)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (p < bkx_min_order(5) || p > bkx_max_order(5)) {
        return BKX_ERROR_ORDER;
    }
    if (plan->elemsize != (p + 1) * (p + 1) * (p + 1)) {
        return BKX_ERROR_ARG;
    }
    if (plan->nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !uL || !vE) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_gather_f64(bkx_tune_get(5, BKX_FORM_F64, p), p, plan->nelmt, dbasis, G, plan->offsets, uL,
                                     vE, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
