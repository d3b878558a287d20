// bk5_f64.hip -- BK5 collocated Laplacian in FP64 on gfx950.
//
// Operator (openmp/BK5.cpp:18-103): per element, with D(i, n) = dbasis[i*nq+n]
// and the six symmetric factors G(e, i, j, f, k),
//   qr = sum_n D(i,n) u(n,j,k),  qs = sum_n D(j,n) u(i,n,k),  qt = sum_n D(k,n) u(i,j,n)
//   (rr, rs, rt) = G(i,j,k) (qr, qs, qt)
//   v(i,j,k) = sum_n D(n,i) rr(n,j,k) + D(n,j) rs(i,n,k) + D(n,k) rt(i,j,n).
//
// Mapping.  One wavefront per element (nq >= 6) or per group of
// floor(64 / nq^2) elements (nq <= 5).  A lane owns the (j, k) column of its
// element and keeps the whole i-line of u, of rr and of the result in
// registers, so the i-direction contractions (qr and the rr transpose) are
// register-only with D(i, n) uniform across the wavefront.  The element's u
// is staged once in LDS for the j- and k-derivatives; rs and rt of one
// i-layer at a time go through a small double-buffered LDS plane for the
// transposed j- and k-contractions.  All LDS traffic stays inside one
// wavefront, so there is no workgroup barrier; layers are processed in i
// order so that only one layer of G (6 values per lane) is live at a time.
//
// HBM traffic per element: u and v once (8 nq^3 B each) and G once
// (48 nq^3 B), all with unit-stride lanes: u/v planes are 8 nq^2 contiguous
// bytes, G(i, j, f, :) rows are nq contiguous doubles.
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

template <int NQ>
struct Bk5Shape {
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int epw = kWave / nq2;       // elements per wavefront
    static constexpr int row = NQ + 1;            // padded LDS row (bank spread)
    static constexpr int plane = NQ * row;        // one padded (j, k) plane
    static constexpr int box = NQ * plane;        // one padded element box
    static_assert(epw >= 1, "one element must fit a wavefront");
};

template <int NQ, int WAVES>
__global__ __launch_bounds__(kWave * WAVES) void bk5_f64_kernel(size_t nelmt, const double* __restrict__ D,
                                                               const double* __restrict__ G,
                                                               const double* __restrict__ u,
                                                               double* __restrict__ v)
{
    using S = Bk5Shape<NQ>;
    constexpr int nq2 = S::nq2;
    constexpr int nq3 = S::nq3;
    constexpr int epw = S::epw;

    __shared__ double s_u[WAVES][epw][S::box];
    __shared__ double s_x[WAVES][2][2][epw][S::plane];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = lane / nq2;
    const bool in_elem = slot < epw;           // idle lanes: nq^2 does not divide 64
    const int sl = in_elem ? slot : epw - 1;   // idle lanes read (never write) a real slot
    const int jk = lane - slot * nq2;
    const int j = in_elem ? jk / NQ : 0;
    const int k = in_elem ? jk - j * NQ : 0;
    const size_t e = (size_t(blockIdx.x) * WAVES + wave) * epw + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    const int lx = j * S::row + k;             // this lane's slot in a padded plane

    // Per-lane rows and columns of D for the j and k contractions.
    double Dj[NQ], Dk[NQ], DTj[NQ], DTk[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
        Dj[n] = D[j * NQ + n];
        Dk[n] = D[k * NQ + n];
        DTj[n] = D[n * NQ + j];
        DTk[n] = D[n * NQ + k];
    }

    // u column along i -> registers and the padded LDS box.
    const double* ue = u + ee * nq3 + jk;
    double* su = s_u[wave][sl];
    double ur[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        ur[i] = active ? ld_stream(ue + i * nq2) : 0.0;
    }
    if (in_elem) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            su[i * S::plane + lx] = ur[i];
        }
    }
    wave_sync();

    // G(e, i, j, f, k): the lane's six factors of layer i are at
    // ge + i * 6 nq^2 + f * nq.  One layer is prefetched ahead.
    const double* ge = G + ee * 6 * nq3 + j * 6 * NQ + k;
    double gn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        gn[f] = active ? ld_stream(ge + f * NQ) : 0.0;
    }
    // acc(m) = v(m, j, k): the i-direction transpose sum_i D(i, m) rr(i) is
    // accumulated as each layer's rr appears (i ascending), the j/k
    // transposes of layer m are added at layer m.
    double acc[NQ];
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
        acc[m] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        double g[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            g[f] = gn[f];
        }
        if (i + 1 < NQ) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                gn[f] = active ? ld_stream(ge + (i + 1) * 6 * nq2 + f * NQ) : 0.0;
            }
        }
        double qr = 0.0, qs = 0.0, qt = 0.0;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qr += D[i * NQ + n] * ur[n];
        }
        const double* sp = su + i * S::plane;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qs += Dj[n] * sp[n * S::row + k];
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qt += Dk[n] * sp[j * S::row + n];
        }
        const double rr = g[0] * qr + g[1] * qs + g[2] * qt;
        const double rs = g[1] * qr + g[3] * qs + g[4] * qt;
        const double rt = g[2] * qr + g[4] * qs + g[5] * qt;

        double* xs = s_x[wave][i & 1][0][sl];
        double* xt = s_x[wave][i & 1][1][sl];
        if (in_elem) {
            xs[lx] = rs;
            xt[lx] = rt;
        }
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            acc[m] += D[i * NQ + m] * rr;
        }
        wave_sync();
        double a = acc[i];
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            a += DTj[n] * xs[n * S::row + k];
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            a += DTk[n] * xt[j * S::row + n];
        }
        acc[i] = a;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            pin(acc[m]);
        }
    }

    if (active) {
        double* ve = v + e * nq3 + jk;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            st_stream(ve + i * nq2, acc[i]);
        }
    }
}

template <int NQ>
int launch_bk5_f64(size_t nelmt, const double* D, const double* G, const double* u, double* v, hipStream_t s)
{
    constexpr int waves = 4;
    constexpr size_t per_block = size_t(waves) * Bk5Shape<NQ>::epw;
    const unsigned grid = grid_for(nelmt, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk5_f64_kernel<NQ, waves>), dim3(grid), dim3(kWave * waves), 0, s, nelmt, D, G, u, v);
    return launch_status();
}

}  // namespace

int bk5_apply_f64(int p, size_t nelmt, const double* D, const double* G, const double* u, double* v, hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) { return launch_bk5_f64<decltype(P)::value + 1>(nelmt, D, G, u, v, s); });
}

}  // namespace bkx
