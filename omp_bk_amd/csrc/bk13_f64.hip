// bk13_f64.hip -- BK1 (mass) and BK3 (Poisson) in FP64 on gfx950.
//
// Operators (openmp/bk1_sumfact.h:16-159, openmp/bk3_sumfact.h:17-187), with
// nm = p + 1 modes and nq = p + 2 Gauss points per direction,
// B(i, p) = basis[i*nq+p] and (BK3) D(n, p) = dbasis[n*nq+p]:
//   U(p,q,r) = sum_ijk B(i,p) B(j,q) B(k,r) u(i,j,k)           (interpolate)
//   BK1: W = JxW * U
//   BK3: qr = sum_n D(n,p) U(n,q,r), qs = .. D(n,q) U(p,n,r), qt = .. D(n,r) U(p,q,n)
//        (rr, rs, rt) = G(p,q,r) (qr, qs, qt)    (six factors, factor-major G)
//        W(p,q,r) = sum_n D(p,n) rr(n,q,r) + D(q,n) rs(p,n,r) + D(r,n) rt(p,q,n)
//   v(i,j,k) = sum_pqr B(i,p) B(j,q) B(k,r) W(p,q,r)           (interpolate back)
// The contractions run in the reference's order (i, then j, then k going
// in; r, then q, then p coming back), each summed in ascending index.
//
// Mapping.  EPB elements per workgroup, each owning nq^2 consecutive lanes
// (a plane of the element's box).  A lane holds one column of the current
// tensor along the first index in registers, so the first-index
// contractions (i <-> p and the p-derivative) are register-only with a
// wavefront-uniform basis; the other two directions go through padded LDS
// boxes, one workgroup barrier per phase:
//   A  lane (j,k): u(:,j,k) from HBM -> a(p,j,k)         -> box0
//   B  lane (q,k): box0 -> b(p,q,k)                       -> box1
//   C  lane (q,r): box1 -> U(p,q,r) in registers
//      BK1: W = JxW U;  BK3: U -> box0, gradient and fluxes with G from HBM,
//      rs -> box1, rt -> box2, then the transposed gradient -> W in registers
//   D  W -> box0; lane (q,k): box0 -> y(p,q,k)            -> box1
//   E  lane (j,k): box1 -> z(p,j,k), then p -> i in registers -> v to HBM
// HBM traffic per element: u and v once (8 nm^3 B each), JxW (8 nq^3 B) or
// G (48 nq^3 B) once, with unit-stride lanes along each plane.
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

template <int NM, int NQ>
struct Bk13Shape {
    static constexpr int nm2 = NM * NM;
    static constexpr int nm3 = nm2 * NM;
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int epb = (256 / nq2) > 0 ? (256 / nq2) : 1;   // elements per workgroup
    static constexpr int threads = ((epb * nq2 + kWave - 1) / kWave) * kWave;
    static constexpr int row = NQ + 1;                                // padded row
    static constexpr int box = NQ * NQ * row;                         // padded nq^3 box
};

template <int NM, int NQ, bool BK3>
__global__ __launch_bounds__((Bk13Shape<NM, NQ>::threads)) void bk13_f64_kernel(
    size_t nelmt, const double* __restrict__ B, const double* __restrict__ D, const double* __restrict__ Q,
    const double* __restrict__ u, double* __restrict__ v)
{
    using S = Bk13Shape<NM, NQ>;
    constexpr int nboxes = BK3 ? 3 : 2;
    constexpr int R = S::row;
    // Padded box index (a, b, c) -> (a * NQ + b) * R + c.
    auto X = [](int a, int b, int c) { return (a * NQ + b) * R + c; };

    __shared__ double s_box[S::epb][nboxes][S::box];

    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const size_t e = size_t(blockIdx.x) * S::epb + slot;
    const bool in_elem = slot < S::epb;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    double* box0 = s_box[in_elem ? slot : 0][0];
    double* box1 = s_box[in_elem ? slot : 0][1];

    // ---- A: lane (j,k), contract i -> p --------------------------------
    if (in_elem && l < S::nm2) {
        const int j = l / NM;
        const int k = l - j * NM;
        double ui[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            ui[i] = active ? load<false>(u + ee * S::nm3 + i * S::nm2 + l) : 0.0;
        }
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                t += ui[i] * B[i * NQ + p];
            }
            box0[X(p, j, k)] = t;
        }
    }
    __syncthreads();

    // ---- B: lane (q,k), contract j -> q --------------------------------
    if (in_elem && l < NQ * NM) {
        const int q = l / NM;
        const int k = l - q * NM;
        double Bq[NM];
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            Bq[j] = B[j * NQ + q];
        }
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                t += box0[X(p, j, k)] * Bq[j];
            }
            box1[X(p, q, k)] = t;
        }
    }
    __syncthreads();

    // ---- C: lane (q,r), contract k -> r; the quadrature-point operator ---
    double W[NQ];
    double U[NQ];
    const int q = l / NQ;
    const int r = l - q * NQ;
    if (in_elem) {
        double Br[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            Br[k] = B[k * NQ + r];
        }
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                t += box1[X(p, q, k)] * Br[k];
            }
            U[p] = t;
        }
        if constexpr (!BK3) {
            // BK1: W = JxW * U (Q = JxW(e, p, q, r))
            const double* je = Q + ee * S::nq3 + l;
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                W[p] = U[p] * (active ? load<false>(je + p * S::nq2) : 0.0);
            }
        } else {
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                box0[X(p, q, r)] = U[p];
            }
            W[0] = 0.0;
        }
    }
    if constexpr (BK3) {
        double* box2 = s_box[in_elem ? slot : 0][2];
        double rr[NQ];
        __syncthreads();
        if (in_elem) {
            // gradient along p (registers), q and r (box0); fluxes with G
            double Dq[NQ], Dr[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                Dq[n] = D[n * NQ + q];
                Dr[n] = D[n * NQ + r];
            }
            const double* ge = Q + ee * 6 * S::nq3 + l;   // G(e, f, p, q, r)
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                double qr = 0.0, qs = 0.0, qt = 0.0;
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    qr += U[n] * D[n * NQ + p];
                }
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    qs += box0[X(p, n, r)] * Dq[n];
                }
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    qt += box0[X(p, q, n)] * Dr[n];
                }
                double g[6];
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    g[f] = active ? load<false>(ge + f * S::nq3 + p * S::nq2) : 0.0;
                }
                rr[p] = g[0] * qr + g[1] * qs + g[2] * qt;
                box1[X(p, q, r)] = g[1] * qr + g[3] * qs + g[4] * qt;
                box2[X(p, q, r)] = g[2] * qr + g[4] * qs + g[5] * qt;
            }
        }
        __syncthreads();
        if (in_elem) {
            double DTq[NQ], DTr[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                DTq[n] = D[q * NQ + n];
                DTr[n] = D[r * NQ + n];
            }
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                double t = 0.0;
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    t += rr[n] * D[p * NQ + n];
                }
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    t += box1[X(p, n, r)] * DTq[n];
                }
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    t += box2[X(p, q, n)] * DTr[n];
                }
                W[p] = t;
            }
        }
    }
    __syncthreads();
    if (in_elem) {
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            box0[X(p, q, r)] = W[p];
        }
    }
    __syncthreads();

    // ---- D: lane (q,k), contract r -> k --------------------------------
    if (in_elem && l < NQ * NM) {
        const int qq = l / NM;
        const int k = l - qq * NM;
        double Bk[NQ];
#pragma unroll
        for (int rr2 = 0; rr2 < NQ; ++rr2) {
            Bk[rr2] = B[k * NQ + rr2];
        }
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t = 0.0;
#pragma unroll
            for (int rr2 = 0; rr2 < NQ; ++rr2) {
                t += box0[X(p, qq, rr2)] * Bk[rr2];
            }
            box1[X(p, qq, k)] = t;
        }
    }
    __syncthreads();

    // ---- E: lane (j,k), contract q -> j, then p -> i -------------------
    if (active && l < S::nm2) {
        const int j = l / NM;
        const int k = l - j * NM;
        double Bj[NQ];
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
            Bj[qq] = B[j * NQ + qq];
        }
        double z[NQ];
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t = 0.0;
#pragma unroll
            for (int qq = 0; qq < NQ; ++qq) {
                t += box1[X(p, qq, k)] * Bj[qq];
            }
            z[p] = t;
        }
        double* ve = v + e * S::nm3 + l;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double t = 0.0;
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                t += z[p] * B[i * NQ + p];
            }
            st_stream(ve + i * S::nm2, t);
        }
    }
}

template <int NM, int NQ, bool BK3>
int launch_bk13_f64(size_t nelmt, const double* B, const double* D, const double* Q, const double* u, double* v,
                    hipStream_t s)
{
    using S = Bk13Shape<NM, NQ>;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk13_f64_kernel<NM, NQ, BK3>), dim3(grid), dim3(S::threads), 0, s, nelmt, B, D, Q, u, v);
    return launch_status();
}

}  // namespace

int bk1_apply_f64(int p, size_t nelmt, const double* B, const double* JxW, const double* u, double* v, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        constexpr int nm = decltype(P)::value + 1;
        return launch_bk13_f64<nm, nm + 1, false>(nelmt, B, nullptr, JxW, u, v, s);
    });
}

int bk3_apply_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* u, double* v,
                  hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        constexpr int nm = decltype(P)::value + 1;
        return launch_bk13_f64<nm, nm + 1, true>(nelmt, B, D, G, u, v, s);
    });
}

}  // namespace bkx
