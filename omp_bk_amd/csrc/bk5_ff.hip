// bk5_ff.hip -- BK5 collocated Laplacian with FP32 storage and float-float
// compensated FP32 arithmetic on gfx950 (the "FP32-compensated" form).
//
// u and the six geometric factors are stored in FP32 and are exact FP32
// values; D is split once into float-float (hi + lo); every contraction is
// a Dot2 compensated sum (TwoProd + TwoSum on the high parts, errors summed
// separately), the fluxes (rr, rs, rt) = G (qr, qs, qt) are float-float, and
// v is joined to FP64 (one double add) when it is stored.  The arithmetic
// follows the float-float scheme of the reference's BK1_ff
// (openmp/BK1_ff.cpp:73-241, openmp/float_float.h:85-134), here with Dot2
// accumulation instead of a renormalisation per multiply-add.
//
// The mapping is that of the FP64 kernel in bk5.hip: one wavefront per
// element (or per group of small elements), a lane owning the (j, k) column,
// u staged in LDS (as exact floats), one double-buffered LDS plane per
// i-layer for rs and rt (as float pairs), the i-direction contractions in
// registers with a wavefront-uniform D(i, n).
//
// HBM traffic per element: u 4 nq^3 B, G 24 nq^3 B, v 8 nq^3 B.
#include "bkx_device.hpp"
#include "bkx_ff.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

template <int NQ>
struct Bk5FfShape {
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int epw = kWave / nq2;
    static constexpr int row = NQ | 1;            // odd row stride in 8-byte words (bank spread)
    static constexpr int plane = NQ * row;
    static constexpr int box = NQ * plane;
};

__device__ __forceinline__ void pin(dot2& x)
{
    asm volatile("" : "+v"(x.s), "+v"(x.c));
}

template <int NQ, int WAVES>
__global__ __launch_bounds__(kWave * WAVES) void bk5_ff_kernel(size_t nelmt, const double* __restrict__ D,
                                                              const float* __restrict__ G,
                                                              const float* __restrict__ u,
                                                              double* __restrict__ v)
{
    using S = Bk5FfShape<NQ>;
    constexpr int nq2 = S::nq2;
    constexpr int nq3 = S::nq3;
    constexpr int epw = S::epw;

    __shared__ float s_u[WAVES][epw][S::box];
    __shared__ ff s_x[WAVES][2][2][epw][S::plane];
    __shared__ ff s_D[NQ * NQ];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = lane / nq2;
    const bool in_elem = slot < epw;
    const int sl = in_elem ? slot : epw - 1;
    const int jk = lane - slot * nq2;
    const int j = in_elem ? jk / NQ : 0;
    const int k = in_elem ? jk - j * NQ : 0;
    const size_t e = (size_t(blockIdx.x) * WAVES + wave) * epw + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    const int lx = j * S::row + k;

    // D split into float-float once per workgroup
    for (int x = threadIdx.x; x < NQ * NQ; x += kWave * WAVES) {
        s_D[x] = ff_split(D[x]);
    }
    __syncthreads();
    ff Dj[NQ], Dk[NQ], DTj[NQ], DTk[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
        Dj[n] = s_D[j * NQ + n];
        Dk[n] = s_D[k * NQ + n];
        DTj[n] = s_D[n * NQ + j];
        DTk[n] = s_D[n * NQ + k];
    }

    const float* ue = u + ee * nq3 + jk;
    float* su = s_u[wave][sl];
    float ur[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        ur[i] = active ? load<false>(ue + i * nq2) : 0.0f;
    }
    if (in_elem) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            su[i * S::plane + lx] = ur[i];
        }
    }
    wave_sync();

    const float* ge = G + ee * 6 * nq3 + j * 6 * NQ + k;
    float gn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        gn[f] = active ? load<false>(ge + f * NQ) : 0.0f;
    }
    dot2 acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        float g[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            g[f] = gn[f];
        }
        if (i + 1 < NQ) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                gn[f] = active ? load<false>(ge + (i + 1) * 6 * nq2 + f * NQ) : 0.0f;
            }
        }
        dot2 dr, ds, dt;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            dr.add(s_D[i * NQ + n], ur[n]);
        }
        const float* sp = su + i * S::plane;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            ds.add(Dj[n], sp[n * S::row + k]);
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            dt.add(Dk[n], sp[j * S::row + n]);
        }
        const ff qr = dr.value(), qs = ds.value(), qt = dt.value();
        // fluxes: G exact floats times float-float gradients
        dot2 fr, fs, ft;
        fr.add(qr, g[0]);
        fr.add(qs, g[1]);
        fr.add(qt, g[2]);
        fs.add(qr, g[1]);
        fs.add(qs, g[3]);
        fs.add(qt, g[4]);
        ft.add(qr, g[2]);
        ft.add(qs, g[4]);
        ft.add(qt, g[5]);
        const ff rr = fr.value();

        ff* xs = s_x[wave][i & 1][0][sl];
        ff* xt = s_x[wave][i & 1][1][sl];
        if (in_elem) {
            xs[lx] = fs.value();
            xt[lx] = ft.value();
        }
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            acc[m].add(s_D[i * NQ + m], rr);
        }
        wave_sync();
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            acc[i].add(DTj[n], xs[n * S::row + k]);
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            acc[i].add(DTk[n], xt[j * S::row + n]);
        }
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            pin(acc[m]);
        }
    }

    if (active) {
        double* ve = v + e * nq3 + jk;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            st_stream(ve + i * nq2, ff_join(acc[i].value()));
        }
    }
}

template <int NQ>
int launch_bk5_ff(size_t nelmt, const double* D, const float* G, const float* u, double* v, hipStream_t s)
{
    constexpr int waves = 4;
    constexpr size_t per_block = size_t(waves) * Bk5FfShape<NQ>::epw;
    const unsigned grid = grid_for(nelmt, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk5_ff_kernel<NQ, waves>), dim3(grid), dim3(kWave * waves), 0, s, nelmt, D, G, u, v);
    return launch_status();
}

}  // namespace

int bk5_apply_f32c(int p, size_t nelmt, const double* D, const float* G, const float* u, double* v, hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) { return launch_bk5_ff<decltype(P)::value + 1>(nelmt, D, G, u, v, s); });
}

}  // namespace bkx
