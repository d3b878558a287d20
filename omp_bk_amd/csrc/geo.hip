// geo.hip -- geometric factors from element nodal coordinates, in FP64 on
// gfx950 (the step before the operators: the reference feeds its operators
// synthetic factors, include/bk_common.h:92-107, so there is no reference
// kernel; the factors are those its harness describes, w detJ J^-1 J^-T and
// w detJ).
//
// coords(e, c, i, j, k): the element's coordinates x, y, z (c = 0, 1, 2) at
// its n^3 GLL nodes, n = p + 1.  J(c, d) = d x_c / d xi_d of the degree-p
// interpolant of the coordinates, with xi_1, xi_2, xi_3 the directions of
// the first, second and third tensor index.  At each point, with weight
// w = w_a w_b w_c:
//   JxW = w detJ,   G = w detJ J^-1 J^-T = w adj(J) adj(J)^T / detJ
// (adj the adjugate; factors rr rs rt ss st tt).
//
// BK3 / BK1 (Gauss points, q = p + 2): the reference gradient of each
// coordinate component, computed by the phases of the transposing BK3
// mapping (bk3_phases.hpp: interpolate with B, differentiate with the Gauss
// collocation derivative D), geo_epb(p) elements per workgroup, J kept in
// registers;
// G is written factor-major, G(e, f, p, q, r), JxW as JxW(e, p, q, r).
// BK5 (GLL nodes, collocated): J by the collocated derivative D(i, n) along
// each direction from the coordinates staged in the LDS, as many elements
// per wavefront as fit; G is written interleaved, G(e, i, j, f, k).
//
// The coordinates are differentiated relative to the element's first node,
// x - x(0, 0, 0): the derivative matrices annihilate constants, and this
// removes the cancellation of differentiating absolute coordinates (up to
// |x| / h times larger than the derivatives), which cost two to three digits
// of J at p >= 7.  The subtraction is exact wherever x(0, 0, 0) is more
// than the element size from the origin (Sterbenz), and its rounding is
// relative to the element size elsewhere.
#include "bk3_phases.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

// G = w adj(J) adj(J)^T / detJ and w detJ from J (row c = d x_c / d xi)
__device__ __forceinline__ void factors_of(const double (&J)[3][3], double w, double (&g)[6], double& wdet)
{
    // adjugate A = cofactor^T: A(d, c) = cofactor(c, d), so that J A = det I
    double A[3][3];
    A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double det = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
    wdet = w * det;
    // J^-1 = A / det: G(a, b) = w det sum_c J^-1(a, c) J^-1(b, c) = (w / det) sum_c A(a, c) A(b, c)
    const double s = w / det;
    constexpr int fa[6] = {0, 0, 0, 1, 1, 2};
    constexpr int fb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int a = fa[f], b = fb[f];
        g[f] = s * (A[a][0] * A[b][0] + A[a][1] * A[b][1] + A[a][2] * A[b][2]);
    }
}

template <int NM, int NQ, int EPB = 1>
struct GeoShape {
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int threads = ((EPB * nq2 + kWave - 1) / kWave) * kWave;
};

// Elements per workgroup of the Gauss-point kernel, the fastest measured on
// MI355X at about 33 M DoF (ms for 1, 2, 3, 4, 8 and 16 elements, capped at
// floor(256 / q^2)):
//   p = 1: 4.94 3.26 2.69 2.38 2.28 2.21   p = 5: 0.80 0.84 0.84 0.85 0.80 0.80
//   p = 2: 1.90 1.27 1.06 0.96 0.95 1.01   p = 6: 0.57 0.61 0.68 0.63 0.62 0.63
//   p = 3: 1.44 1.13 1.27 1.16 1.12 1.15   p = 7: 1.04 1.07 0.92 0.94 0.93 0.91
//   p = 4: 1.11 1.13 0.97 1.05 0.92 0.94   p = 8: 0.71 0.71 0.71 0.73 0.71 0.72
// At p = 5, 6 and 8 the kernel's registers (J of three components for a
// column of points) already limit the CU to a few elements.
constexpr int geo_epb(int p)
{
    constexpr int epb[9] = {1, 16, 4, 8, 7, 1, 1, 3, 1};
    return epb[p];
}

// BK3 / BK1 points.  EPB elements per workgroup; lane l < q^2 of slot s is
// the (q, r) lane of element s's Gauss-point plane, holding column p.
template <int NM, int NQ, int EPB>
__global__ __launch_bounds__((GeoShape<NM, NQ, EPB>::threads)) void geo_gauss_kernel(
    size_t nelmt, const double* __restrict__ Bm, const double* __restrict__ Dm, const double* __restrict__ wq,
    const double* __restrict__ coords, double* __restrict__ G, double* __restrict__ JxW)
{
    using E = bk3p::Elem<NM, NQ>;
    using S = GeoShape<NM, NQ, EPB>;
    __shared__ double s_box[EPB][3 * E::box];

    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const bool in_elem = slot < EPB;
    const size_t e = size_t(blockIdx.x) * EPB + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    E el;
    el.init(s_box[in_elem ? slot : 0], l, in_elem);
    const bool modal = in_elem && l < E::nm2;
    // J(c, d) at the lane's column of points (p, lane), p < NQ, in registers:
    // the LDS then holds only the gradient's three boxes, so that several
    // workgroups share a CU
    double Jc[3][3][NQ];
    // the lane's column of coordinate c, loaded unconditionally (column 0
    // outside the modal plane, whose values no phase reads), one coordinate
    // ahead of its gradient so that the next column's loads are in flight
    // during the current gradient
    const int lm = modal ? l : 0;
    double xn[NM], x0n;
    auto load_col = [&](int c) {
        const double* xc = coords + (ee * 3 + c) * E::nm3;
        x0n = xc[0];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            xn[i] = xc[lm + i * E::nm2];
        }
    };
    load_col(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double x[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            x[i] = xn[i] - x0n;
        }
        if (c < 2) {
            load_col(c + 1);
        }
        double qr[NQ];
        el.template grad<0>(Bm, Dm, x, qr);   // ends with the workgroup synchronised
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            const int xx = E::X(p, el.lq, el.lq2);
            Jc[c][0][p] = qr[p];
            Jc[c][1][p] = el.B[xx];
            Jc[c][2][p] = el.C[xx];
        }
        // the next gradient writes box A, then box B only after a barrier,
        // so this lane's reads of B and C above come first
    }
    if (!active) {
        return;
    }
    const double wqr = wq[el.lq] * wq[el.lq2];
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
        double J[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                J[c][d] = Jc[c][d][p];
            }
        }
        double g[6], wdet;
        factors_of(J, wq[p] * wqr, g, wdet);
        if (G) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                G[(e * 6 + f) * S::nq3 + p * S::nq2 + l] = g[f];
            }
        }
        if (JxW) {
            JxW[e * S::nq3 + p * S::nq2 + l] = wdet;
        }
    }
}

// BK5 points (GLL nodes, collocated).  gll_epb(n) elements per workgroup
// of one wavefront (n^2 lanes each, so that the low orders fill it); lane
// l = (j, k) < n^2 of slot s loops over i.
constexpr int gll_epb(int n) { return kWave / (n * n) > 0 ? kWave / (n * n) : 1; }

template <int N>
__global__ __launch_bounds__((GeoShape<N, N, gll_epb(N)>::threads)) void geo_gll_kernel(
    size_t nelmt, const double* __restrict__ D, const double* __restrict__ wq, const double* __restrict__ coords,
    double* __restrict__ G)
{
    constexpr int n2 = N * N;
    constexpr int n3 = n2 * N;
    constexpr int EPB = gll_epb(N);
    constexpr int T = GeoShape<N, N, EPB>::threads;
    constexpr int per = (EPB * 3 * n3 + T - 1) / T;   // coordinates staged per lane
    __shared__ double sx[EPB * 3 * n3];
    const size_t e0 = size_t(blockIdx.x) * EPB;
    const int l0 = threadIdx.x;
    // the workgroup's elements are consecutive: all of a lane's loads in
    // flight at once (clamped to the last element), then relative to each
    // element's node 0
    const double* xe = coords + e0 * 3 * n3;
    double xv[per];
    if constexpr (EPB == 1) {
#pragma unroll
        for (int r = 0; r < per; ++r) {
            const int t = l0 + r * T;
            xv[r] = xe[t < 3 * n3 ? t : 0];
        }
        const double x0[3] = {xe[0], xe[n3], xe[2 * n3]};
#pragma unroll
        for (int r = 0; r < per; ++r) {
            const int t = l0 + r * T;
            if (t < 3 * n3) {
                sx[t] = xv[r] - x0[t / n3];
            }
        }
    } else {
        const size_t nv = (nelmt - e0 < size_t(EPB) ? nelmt - e0 : size_t(EPB)) * 3 * n3;
        double xz[per];
#pragma unroll
        for (int r = 0; r < per; ++r) {
            const int t = l0 + r * T;
            const int tc = size_t(t) < nv ? t : 0;
            xv[r] = xe[tc];
            xz[r] = xe[(tc / (3 * n3)) * 3 * n3 + ((tc / n3) % 3) * n3];
        }
#pragma unroll
        for (int r = 0; r < per; ++r) {
            const int t = l0 + r * T;
            if (t < EPB * 3 * n3) {
                sx[t] = xv[r] - xz[r];
            }
        }
    }
    __syncthreads();
    const int slot = l0 / n2;
    const int l = l0 - slot * n2;
    const size_t e = e0 + slot;
    if (slot >= EPB || e >= nelmt) {
        return;
    }
    const int j = l / N;
    const int k = l - j * N;
    const auto Dr = launder<1>(D);
    const double* sxe = sx + slot * 3 * n3;
#pragma unroll 1
    for (int i = 0; i < N; ++i) {
        double J[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double* xc = sxe + c * n3;
            double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
            for (int m = 0; m < N; ++m) {
                d0 += Dr[i * N + m] * xc[(m * N + j) * N + k];
                d1 += Dr[j * N + m] * xc[(i * N + m) * N + k];
                d2 += Dr[k * N + m] * xc[(i * N + j) * N + m];
            }
            J[c][0] = d0;
            J[c][1] = d1;
            J[c][2] = d2;
        }
        double g[6], wdet;
        factors_of(J, wq[i] * wq[j] * wq[k], g, wdet);
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            st_stream(G + e * 6 * n3 + ((i * N + j) * 6 + f) * N + k, g[f]);
        }
    }
}

}  // namespace

int geo_bk3_f64(int p, size_t nelmt, const double* B, const double* D, const double* wq, const double* coords,
                double* G, double* JxW, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        constexpr int NM = decltype(P)::value + 1;
        constexpr int NQ = NM + 1;
        constexpr int EPB = geo_epb(NM - 1);
        static_assert(EPB * NQ * NQ <= 1024, "one workgroup");
        const unsigned grid = grid_for(nelmt, EPB);
        if (grid == 0) {
            return int(BKX_ERROR_SIZE);
        }
        hipLaunchKernelGGL((geo_gauss_kernel<NM, NQ, EPB>), dim3(grid), dim3(GeoShape<NM, NQ, EPB>::threads), 0, s,
                           nelmt, B, D, wq, coords, G, JxW);
        return launch_status();
    });
}

int geo_bk5_f64(int p, size_t nelmt, const double* D, const double* wq, const double* coords, double* G,
                hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) {
        constexpr int N = decltype(P)::value + 1;
        constexpr int EPB = gll_epb(N);
        const unsigned grid = grid_for(nelmt, EPB);
        if (grid == 0) {
            return int(BKX_ERROR_SIZE);
        }
        hipLaunchKernelGGL((geo_gll_kernel<N>), dim3(grid), dim3(GeoShape<N, N, EPB>::threads), 0, s, nelmt, D, wq,
                           coords, G);
        return launch_status();
    });
}

}  // namespace bkx
