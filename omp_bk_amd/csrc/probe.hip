// probe.hip -- bandwidth probes for the roofline of the streaming operators,
// the HIP counterparts of the reference's STREAM-like roofs
// (roofline/stream_omp.cpp:86-117): a copy, and a read-heavy mix that reads
// R blocks for every block it writes (R = 9 is the reference's read9, BK3's
// share of reads; R = 7 is BK5's 56 B read : 8 B written per DoF in FP64).
#include "bkx.h"
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

constexpr int kProbeThreads = 256;

// c = a, 16 B per lane per access, grid-stride
__global__ __launch_bounds__(kProbeThreads) void copy_kernel(const f64x2* __restrict__ a, f64x2* __restrict__ c,
                                                             size_t n2)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += stride) {
        st_stream(c + i, ld_stream(a + i));
    }
}

// c[blk*256 + l] = sum_j a[(blk*R + j)*256 + l] over 256-double blocks
// (two doubles per lane)
template <int R>
__global__ __launch_bounds__(kProbeThreads) void read_mix_kernel(const f64x2* __restrict__ a,
                                                                 f64x2* __restrict__ c, size_t nblk)
{
    constexpr int half = 128;   // f64x2 per 256-double block
    const size_t stride = size_t(gridDim.x) * (blockDim.x / half);
    const int l = threadIdx.x % half;
    for (size_t blk = size_t(blockIdx.x) * (blockDim.x / half) + threadIdx.x / half; blk < nblk; blk += stride) {
        f64x2 s = f64x2{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const f64x2 x = ld_stream(a + (blk * R + j) * half + l);
            s.x += x.x;
            s.y += x.y;
        }
        st_stream(c + blk * half + l, s);
    }
}

unsigned probe_grid(size_t work_items)
{
    int dev = 0;
    int cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    }
    const size_t want = (work_items + kProbeThreads - 1) / kProbeThreads;
    return unsigned(std::min<size_t>(want, size_t(cus) * 16));
}

}  // namespace
}  // namespace bkx

extern "C" {

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
    const unsigned grid = bkx::probe_grid(n / 2);
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
    const unsigned grid = bkx::probe_grid(n_out / 2);
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
