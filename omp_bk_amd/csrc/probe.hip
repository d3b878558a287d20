// probe.hip -- bandwidth probes for the roofline of the streaming operators,
// and an FP64 issue-rate probe (vector FMAs against FP64 matrix-core
// products).  The bandwidth probes are the HIP counterparts of the reference's STREAM-like roofs
// (roofline/stream_omp.cpp:86-117): a copy, and a read-heavy mix that reads
// R blocks for every block it writes (R = 9 is the reference's read9, BK3's
// share of reads; R = 7 is BK5's 56 B read : 8 B written per DoF in FP64).
#include "bkx.h"
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

constexpr int kProbeThreads = 256;

// c = a: each workgroup copies kProbeUnroll consecutive 16-byte vectors per
// lane (lanes adjacent within each), one workgroup per 256 * kProbeUnroll
// vectors; the tail is guarded.
constexpr int kProbeUnroll = 4;

__global__ __launch_bounds__(kProbeThreads) void copy_kernel(const f64x2* __restrict__ a, f64x2* __restrict__ c,
                                                             size_t n2)
{
    const size_t base = size_t(blockIdx.x) * kProbeThreads * kProbeUnroll + threadIdx.x;
    f64x2 x[kProbeUnroll];
#pragma unroll
    for (int t = 0; t < kProbeUnroll; ++t) {
        const size_t i = base + size_t(t) * kProbeThreads;
        x[t] = i < n2 ? ld_stream(a + i) : f64x2{0.0, 0.0};
    }
#pragma unroll
    for (int t = 0; t < kProbeUnroll; ++t) {
        const size_t i = base + size_t(t) * kProbeThreads;
        if (i < n2) {
            st_stream(c + i, x[t]);
        }
    }
}

// c[blk*256 + l] = sum_j a[(blk*R + j)*256 + l] over 256-double blocks
// (two doubles per lane)
template <int R>
__global__ __launch_bounds__(kProbeThreads) void read_mix_kernel(const f64x2* __restrict__ a,
                                                                 f64x2* __restrict__ c, size_t nblk)
{
    constexpr int half = 128;   // f64x2 per 256-double block
    const int l = threadIdx.x % half;
    const size_t blk = size_t(blockIdx.x) * (kProbeThreads / half) + threadIdx.x / half;
    if (blk < nblk) {
        f64x2 x[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            x[j] = ld_stream(a + (blk * R + j) * half + l);
        }
        f64x2 s = f64x2{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < R; ++j) {
            s.x += x[j].x;
            s.y += x[j].y;
        }
        st_stream(c + blk * half + l, s);
    }
}

// a read alone: every 16-byte vector of a read once with the non-temporal
// hint, kReadUnroll per lane in flight; each wavefront writes one vector (its
// lanes' sums), so that the loads are not dead code and the writes are
// 1 / (64 kReadUnroll) of the reads (32 loads per lane amortise the
// wavefront's sum: 7.15 TB/s against 6.9 with 16).  The HBM read ceiling the
// mixes above stay below.
constexpr int kReadUnroll = 32;

__global__ __launch_bounds__(kProbeThreads) void read_kernel(const f64x2* __restrict__ a, f64x2* __restrict__ c,
                                                             size_t n2)
{
    const size_t base = size_t(blockIdx.x) * kProbeThreads * kReadUnroll + threadIdx.x;
    f64x2 x[kReadUnroll];
#pragma unroll
    for (int t = 0; t < kReadUnroll; ++t) {
        const size_t i = base + size_t(t) * kProbeThreads;
        x[t] = i < n2 ? ld_stream(a + i) : f64x2{0.0, 0.0};
    }
    f64x2 s = x[0];
#pragma unroll
    for (int t = 1; t < kReadUnroll; ++t) {
        s.x += x[t].x;
        s.y += x[t].y;
    }
#pragma unroll
    for (int m = kWave / 2; m > 0; m /= 2) {
        s.x += __shfl_xor(s.x, m);
        s.y += __shfl_xor(s.y, m);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        st_stream(c + size_t(blockIdx.x) * (kProbeThreads / kWave) + threadIdx.x / kWave, s);
    }
}

// FP64 arithmetic issue rates, the compute side of the roofline and the
// facts behind keeping the BK3 contractions on the vector ALU:
//   KIND 0: kRateFma vector FMAs per iteration (v_fma_f64, 128 flops per
//           wavefront instruction), in kFmaChains independent chains;
//   KIND 1: kRateMfma matrix-core products per iteration
//           (v_mfma_f64_16x16x4_f64, 2048 flops each), in as many chains;
//   KIND 2: both in one instruction stream (whether the two pipes overlap);
//   KIND 3: kRateFma non-packed FP32 FMAs per iteration (v_fma_f32, 128
//           flops per wavefront instruction; packed FP32 is disabled for
//           every kernel here), the rate float-float arithmetic runs at.
// Every lane writes one sum of its chains, so nothing is dead code.
constexpr int kRateThreads = 256;
constexpr int kFmaChains = 16;
constexpr int kRateFma = 64;    // 4 per chain
constexpr int kRateMfma = 4;
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int KIND>
__global__ __launch_bounds__(kRateThreads) void fp64_rate_kernel(int iters, double* __restrict__ out)
{
    const double a = 1.0 + 1e-6 * threadIdx.x;
    const double b = 1.0 - 1e-6 * (threadIdx.x & 7);
    double f[kFmaChains];
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) {
        f[k] = 1e-3 * (k + 1);
    }
    f64x4 c[kRateMfma];
#pragma unroll
    for (int k = 0; k < kRateMfma; ++k) {
        c[k] = f64x4{0.0, 0.0, 0.0, 1e-3 * k};
    }
    float h[kFmaChains];
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) {
        h[k] = 1e-3f * float(k + 1);
    }
    const float af = float(a), bf = float(b);
    for (int it = 0; it < iters; ++it) {
        if constexpr (KIND == 3) {
#pragma unroll
            for (int r = 0; r < kRateFma / kFmaChains; ++r) {
#pragma unroll
                for (int k = 0; k < kFmaChains; ++k) {
                    h[k] = __builtin_fmaf(h[k], bf, af);
                }
            }
        }
        if constexpr (KIND == 0 || KIND == 2) {
#pragma unroll
            for (int r = 0; r < kRateFma / kFmaChains; ++r) {
#pragma unroll
                for (int k = 0; k < kFmaChains; ++k) {
                    f[k] = __builtin_fma(f[k], b, a);
                }
            }
        }
        if constexpr (KIND == 1 || KIND == 2) {
#pragma unroll
            for (int k = 0; k < kRateMfma; ++k) {
                c[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[k], 0, 0, 0);
            }
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) {
        s += f[k];
    }
#pragma unroll
    for (int k = 0; k < kRateMfma; ++k) {
        s += c[k][0] + c[k][1] + c[k][2] + c[k][3];
    }
#pragma unroll
    for (int k = 0; k < kFmaChains; ++k) {
        s += double(h[k]);
    }
    out[size_t(blockIdx.x) * kRateThreads + threadIdx.x] = s;
}

}  // namespace
}  // namespace bkx

extern "C" {

int bkx_probe_read(const double* a, size_t n, double* c, bkx_stream_t stream)
{
    if (n == 0) {
        return BKX_SUCCESS;
    }
    if (!a || !c) {
        return BKX_ERROR_NULL;
    }
    if (n % 2) {
        return BKX_ERROR_ARG;
    }
    const size_t per_block = size_t(bkx::kProbeThreads) * bkx::kReadUnroll;
    const unsigned grid = bkx::grid_for(n / 2, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL(bkx::read_kernel, dim3(grid), dim3(bkx::kProbeThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const bkx::f64x2*>(a),
                       reinterpret_cast<bkx::f64x2*>(c), n / 2);
    return bkx::launch_status();
}

int bkx_probe_fp64_rate(int kind, int iters, double* out, size_t n, bkx_stream_t stream)
{
    if (n == 0) {
        return BKX_SUCCESS;
    }
    if (!out) {
        return BKX_ERROR_NULL;
    }
    if (n % bkx::kRateThreads || iters < 0) {
        return BKX_ERROR_ARG;
    }
    const unsigned grid = bkx::grid_for(n, bkx::kRateThreads);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (kind) {
    case 0:
        hipLaunchKernelGGL(bkx::fp64_rate_kernel<0>, dim3(grid), dim3(bkx::kRateThreads), 0, s, iters, out);
        break;
    case 1:
        hipLaunchKernelGGL(bkx::fp64_rate_kernel<1>, dim3(grid), dim3(bkx::kRateThreads), 0, s, iters, out);
        break;
    case 2:
        hipLaunchKernelGGL(bkx::fp64_rate_kernel<2>, dim3(grid), dim3(bkx::kRateThreads), 0, s, iters, out);
        break;
    case 3:
        hipLaunchKernelGGL(bkx::fp64_rate_kernel<3>, dim3(grid), dim3(bkx::kRateThreads), 0, s, iters, out);
        break;
    default:
        return BKX_ERROR_ARG;
    }
    return bkx::launch_status();
}

int bkx_probe_copy(const double* a, double* c, size_t n, bkx_stream_t stream)
{
    if (n == 0) {
        return BKX_SUCCESS;
    }
    if (!a || !c) {
        return BKX_ERROR_NULL;
    }
    if (n % 2) {
        return BKX_ERROR_ARG;
    }
    const size_t per_block = size_t(bkx::kProbeThreads) * bkx::kProbeUnroll;
    const unsigned grid = bkx::grid_for(n / 2, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL(bkx::copy_kernel, dim3(grid), dim3(bkx::kProbeThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const bkx::f64x2*>(a),
                       reinterpret_cast<bkx::f64x2*>(c), n / 2);
    return bkx::launch_status();
}

int bkx_probe_read_mix(int reads_per_write, const double* a, double* c, size_t n_out, bkx_stream_t stream)
{
    if (n_out == 0) {
        return BKX_SUCCESS;
    }
    if (!a || !c) {
        return BKX_ERROR_NULL;
    }
    if (n_out % 256) {
        return BKX_ERROR_ARG;
    }
    const size_t nblk = n_out / 256;
    const unsigned grid = bkx::grid_for(nblk, bkx::kProbeThreads / 128);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bkx::f64x2* a2 = reinterpret_cast<const bkx::f64x2*>(a);
    bkx::f64x2* c2 = reinterpret_cast<bkx::f64x2*>(c);
    switch (reads_per_write) {
    case 1:
        hipLaunchKernelGGL(bkx::read_mix_kernel<1>, dim3(grid), dim3(bkx::kProbeThreads), 0, s, a2, c2, nblk);
        break;
    case 3:
        hipLaunchKernelGGL(bkx::read_mix_kernel<3>, dim3(grid), dim3(bkx::kProbeThreads), 0, s, a2, c2, nblk);
        break;
    case 7:
        hipLaunchKernelGGL(bkx::read_mix_kernel<7>, dim3(grid), dim3(bkx::kProbeThreads), 0, s, a2, c2, nblk);
        break;
    case 9:
        hipLaunchKernelGGL(bkx::read_mix_kernel<9>, dim3(grid), dim3(bkx::kProbeThreads), 0, s, a2, c2, nblk);
        break;
    default:
        return BKX_ERROR_ARG;
    }
    return bkx::launch_status();
}

}  // extern "C"
