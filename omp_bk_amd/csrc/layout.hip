// layout.hip -- the reference harness's layout helpers on the device:
//
//   make_bk5_G (include/bk_harness.h:217-244): BK5's interleaved six factors
//     G6i(e, i, j, s, k) -> F = 6 (symmetric) or 9 (full tensor, row-major)
//     factors per point in layout 0 (factor-major G(f, i, j, k)), 1
//     (interleaved G(i, j, f, k)) or 2 (column G(f, k, i, j));
//   its inverse, any of those layouts -> G6i, the layout the BK5 operators
//     read (for F = 9 the entries on and above the diagonal: the tensor is
//     symmetric);
//   transpose_jxw (include/bk_harness.h:87-102): BK1's JxW(e, p, q, r) ->
//     JxW_t(e, r, q, p).
//
// Pure data movement, bound by HBM: every thread writes consecutive output
// values (coalesced stores) and gathers its input from the same element,
// whose factors (at most 18 KB in FP64) stay in the L2 cache between the
// lanes that read them.  A grid-stride loop over the flat output index; the
// extents are compile-time, so the index decoding is multiply-shift.
#include <hip/hip_runtime.h>

#include "bkx_launch.hpp"

namespace bkx {
namespace {

constexpr int kLayoutThreads = 256;
constexpr int kSym9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
constexpr int kUpper6[6] = {0, 1, 2, 4, 5, 8};   // s -> row-major (a, b), a <= b

// index of point (i, j, k), factor f of F, within one element
template <int NQ, int F, int LAYOUT>
__device__ __forceinline__ int g_index(int i, int j, int k, int f)
{
    if constexpr (LAYOUT == 1) {
        return ((i * NQ + j) * F + f) * NQ + k;
    } else if constexpr (LAYOUT == 2) {
        return ((f * NQ + k) * NQ + i) * NQ + j;
    } else {
        return ((f * NQ + i) * NQ + j) * NQ + k;
    }
}

// TO_G6I false: out = F factors in LAYOUT from in = G6i; true: the inverse.
template <typename T, int NQ, int F, int LAYOUT, bool TO_G6I>
__global__ __launch_bounds__(kLayoutThreads) void bk5_layout_kernel(size_t nelmt, const T* __restrict__ in,
                                                                    T* __restrict__ out)
{
    constexpr int nq3 = NQ * NQ * NQ;
    constexpr int nout = (TO_G6I ? 6 : F) * nq3;
    constexpr int nin = (TO_G6I ? F : 6) * nq3;
    const size_t total = nelmt * nout;
    for (size_t o = size_t(blockIdx.x) * kLayoutThreads + threadIdx.x; o < total;
         o += size_t(gridDim.x) * kLayoutThreads) {
        const size_t e = o / nout;
        const int x = int(o - e * nout);
        int src;
        if constexpr (TO_G6I) {
            // x = ((i * NQ + j) * 6 + s) * NQ + k
            const int k = x % NQ;
            const int s = (x / NQ) % 6;
            const int ij = x / (6 * NQ);
            src = g_index<NQ, F, LAYOUT>(ij / NQ, ij % NQ, k, F == 6 ? s : kUpper6[s]);
        } else {
            int i, j, k, f;
            if constexpr (LAYOUT == 1) {
                k = x % NQ;
                f = (x / NQ) % F;
                j = (x / (NQ * F)) % NQ;
                i = x / (NQ * F * NQ);
            } else if constexpr (LAYOUT == 2) {
                j = x % NQ;
                i = (x / NQ) % NQ;
                k = (x / (NQ * NQ)) % NQ;
                f = x / nq3;
            } else {
                k = x % NQ;
                j = (x / NQ) % NQ;
                i = (x / (NQ * NQ)) % NQ;
                f = x / nq3;
            }
            src = ((i * NQ + j) * 6 + (F == 6 ? f : kSym9[f])) * NQ + k;
        }
        out[o] = in[e * nin + src];
    }
}

template <typename T, int NQ>
__global__ __launch_bounds__(kLayoutThreads) void jxw_transpose_kernel(size_t nelmt, const T* __restrict__ in,
                                                                       T* __restrict__ out)
{
    constexpr int nq3 = NQ * NQ * NQ;
    const size_t total = nelmt * nq3;
    for (size_t o = size_t(blockIdx.x) * kLayoutThreads + threadIdx.x; o < total;
         o += size_t(gridDim.x) * kLayoutThreads) {
        const size_t e = o / nq3;
        const int x = int(o - e * nq3);   // (r, q, p) of the output
        const int p = x % NQ, q = (x / NQ) % NQ, r = x / (NQ * NQ);
        out[o] = in[e * nq3 + (p * NQ + q) * NQ + r];
    }
}

// enough workgroups to fill the device several times over; the loop covers
// the rest
inline unsigned layout_grid(size_t total)
{
    const size_t g = (total + kLayoutThreads - 1) / kLayoutThreads;
    return unsigned(g < 65536 ? g : 65536);
}

template <typename T, bool TO_G6I>
int bk5_layout(int p, size_t nelmt, int F, int layout, const T* in, T* out, hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) {
        constexpr int NQ = decltype(P)::value + 1;
        const size_t total = nelmt * size_t(TO_G6I ? 6 : F) * NQ * NQ * NQ;
        const unsigned grid = layout_grid(total);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(kLayoutThreads), 0, s, nelmt, in, out);
            return launch_status();
        };
        if (F == 6) {
            if (layout == 0) return go(bk5_layout_kernel<T, NQ, 6, 0, TO_G6I>);
            if (layout == 1) return go(bk5_layout_kernel<T, NQ, 6, 1, TO_G6I>);
            return go(bk5_layout_kernel<T, NQ, 6, 2, TO_G6I>);
        }
        if (layout == 0) return go(bk5_layout_kernel<T, NQ, 9, 0, TO_G6I>);
        if (layout == 1) return go(bk5_layout_kernel<T, NQ, 9, 1, TO_G6I>);
        return go(bk5_layout_kernel<T, NQ, 9, 2, TO_G6I>);
    });
}

template <typename T>
int jxw_transpose(int p, size_t nelmt, const T* in, T* out, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        constexpr int NQ = decltype(P)::value + 2;
        hipLaunchKernelGGL((jxw_transpose_kernel<T, NQ>), dim3(layout_grid(nelmt * NQ * NQ * NQ)),
                           dim3(kLayoutThreads), 0, s, nelmt, in, out);
        return launch_status();
    });
}

hipStream_t as_stream(bkx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

template <typename T, bool TO_G6I>
int bk5_layout_entry(int p, size_t nelmt, int F, int layout, const T* in, T* out, bkx_stream_t stream)
{
    if (p < 1 || p > 7) {
        return BKX_ERROR_ORDER;
    }
    if ((F != 6 && F != 9) || layout < 0 || layout > 2) {
        return BKX_ERROR_ARG;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!in || !out) {
        return BKX_ERROR_NULL;
    }
    return bk5_layout<T, TO_G6I>(p, nelmt, F, layout, in, out, as_stream(stream));
}

template <typename T>
int jxw_entry(int p, size_t nelmt, const T* in, T* out, bkx_stream_t stream)
{
    if (p < 1 || p > 8) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!in || !out) {
        return BKX_ERROR_NULL;
    }
    return jxw_transpose<T>(p, nelmt, in, out, as_stream(stream));
}

}  // namespace
}  // namespace bkx

extern "C" {

int bkx_bk5_make_G_f64(int p, size_t nelmt, const double* G6i, int F, int layout, double* G, bkx_stream_t stream)
{
    return bkx::bk5_layout_entry<double, false>(p, nelmt, F, layout, G6i, G, stream);
}

int bkx_bk5_make_G_f32(int p, size_t nelmt, const float* G6i, int F, int layout, float* G, bkx_stream_t stream)
{
    return bkx::bk5_layout_entry<float, false>(p, nelmt, F, layout, G6i, G, stream);
}

int bkx_bk5_interleave_G_f64(int p, size_t nelmt, int F, int layout, const double* G, double* G6i,
                             bkx_stream_t stream)
{
    return bkx::bk5_layout_entry<double, true>(p, nelmt, F, layout, G, G6i, stream);
}

int bkx_bk5_interleave_G_f32(int p, size_t nelmt, int F, int layout, const float* G, float* G6i, bkx_stream_t stream)
{
    return bkx::bk5_layout_entry<float, true>(p, nelmt, F, layout, G, G6i, stream);
}

int bkx_bk1_transpose_jxw_f64(int p, size_t nelmt, const double* JxW, double* JxW_t, bkx_stream_t stream)
{
    return bkx::jxw_entry<double>(p, nelmt, JxW, JxW_t, stream);
}

int bkx_bk1_transpose_jxw_f32(int p, size_t nelmt, const float* JxW, float* JxW_t, bkx_stream_t stream)
{
    return bkx::jxw_entry<float>(p, nelmt, JxW, JxW_t, stream);
}

}  // extern "C"
