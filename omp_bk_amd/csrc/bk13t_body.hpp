// bk13t_body.hpp -- the transposing BK1/BK3 mapping as a device function
// (one workgroup's pass over up to epb elements), shared by the contiguous
// kernels of bk13t.hip and the mixed-order kernels of bk3_mixed.hip.
// The mapping is described in bk13t.hip.
#pragma once

#include "bkx_device.hpp"

namespace bkx {
namespace bk13t {

// DIV divides the default number of elements per workgroup, floor(256 / nq^2)
// (fewer elements per workgroup: less LDS per workgroup, more workgroups
// per CU, fewer lanes idle at the barriers).
template <int NM, int NQ, int DIV = 1>
struct Bk13tShape {
    static constexpr int nm2 = NM * NM;
    static constexpr int nm3 = nm2 * NM;
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int full = 256 / nq2 > 0 ? 256 / nq2 : 1;
    static constexpr int epb = full / DIV > 0 ? full / DIV : 1;
    static constexpr int threads = ((epb * nq2 + kWave - 1) / kWave) * kWave;
    static constexpr int row = NQ + 1;
    static constexpr int box = NQ * NQ * row;
};

// With SB, a scheduling barrier after each row of M keeps one row of scalar
// operands live at a time (lower register use, the row's load latency
// exposed); without it the compiler schedules the row loads freely.
template <bool SB>
__device__ __forceinline__ void row_fence()
{
    if constexpr (SB) {
        __builtin_amdgcn_sched_barrier(0);
    }
}

// out[c] = sum_m M[m * NQ + c] in[m] (c < NOUT): row m of M times in[m],
// m ascending.
template <int NQ, bool SB, int NIN, int NOUT, typename P>
__device__ __forceinline__ void contract_cols(P M, const double (&in)[NIN], double (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out[c] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            out[c] += M[m * NQ + c] * in[m];
        }
        row_fence<SB>();
    }
}

// the same for two inputs sharing the rows of M
template <int NQ, bool SB, int NIN, int NOUT, typename P>
__device__ __forceinline__ void contract_cols2(P M, const double (&in0)[NIN], double (&out0)[NOUT],
                                               const double (&in1)[NIN], double (&out1)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out0[c] = 0.0;
        out1[c] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            out0[c] += M[m * NQ + c] * in0[m];
            out1[c] += M[m * NQ + c] * in1[m];
        }
        row_fence<SB>();
    }
}

// out[c] = sum_m M[c * NQ + m] in[m] (c < NOUT): row c of M dotted with in.
template <int NQ, bool SB, int NIN, int NOUT, typename P>
__device__ __forceinline__ void contract_rows(P M, const double (&in)[NIN], double (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t += M[c * NQ + m] * in[m];
        }
        out[c] = t;
        row_fence<SB>();
    }
}

// the same for three inputs sharing the rows of M
template <int NQ, bool SB, int NIN, int NOUT, typename P>
__device__ __forceinline__ void contract_rows3(P M, const double (&in0)[NIN], double (&out0)[NOUT],
                                               const double (&in1)[NIN], double (&out1)[NOUT],
                                               const double (&in2)[NIN], double (&out2)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t0 += M[c * NQ + m] * in0[m];
            t1 += M[c * NQ + m] * in1[m];
            t2 += M[c * NQ + m] * in2[m];
        }
        out0[c] = t0;
        out1[c] = t1;
        out2[c] = t2;
        row_fence<SB>();
    }
}

// One workgroup's pass over up to epb elements.  ue, Qe and ve are this
// lane's element's u, JxW/G and v (any readable address when !active);
// smem holds epb * 3 boxes.
template <int NM, int NQ, bool BK3, int DIV, bool SB>
__device__ __forceinline__ void bk13t_body(const double* __restrict__ B, const double* __restrict__ D,
                                           const double* __restrict__ Qe, const double* __restrict__ ue,
                                           double* __restrict__ ve, bool active, int slot, int l, double* smem,
                                           int dep = 0)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr int nq2 = S::nq2;
    constexpr int R = S::row;
    const bool in_elem = slot < S::epb;
    const int sl = in_elem ? slot : 0;
    double* bA = smem + (sl * 3 + 0) * S::box;
    double* bB = smem + (sl * 3 + 1) * S::box;
    double* bC = smem + (sl * 3 + 2) * S::box;
    auto X = [](int a, int b, int c) { return (a * NQ + b) * R + c; };
    // lane (a, b) of an NA x NB plane, when l < NA * NB
    const int l_m = l / NM, l_m2 = l - (l / NM) * NM;    // (l / NM, l % NM)
    const int l_q = l / NQ, l_q2 = l - (l / NQ) * NQ;    // (l / NQ, l % NQ)

    // ---- P1: lane (j,k), contract i -> p ---------------------------------
    if (in_elem && l < S::nm2) {
        double x[NM], a[NQ];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            x[i] = active ? ue[i * S::nm2 + l] : 0.0;
        }
        contract_cols<NQ, SB>(launder<1>(B, dep), x, a);
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            bA[X(p, l_m, l_m2)] = a[p];
        }
    }
    if constexpr (!BK3) {
        // JxW(e, p, q, r) staged into box C with unit-stride loads
        if (in_elem) {
#pragma unroll
            for (int t = 0; t < NQ; ++t) {
                const int x = t * nq2 + l;
                bC[x + x / NQ] = active ? Qe[x] : 0.0;
            }
        }
    }
    __syncthreads();

    // ---- P2: lane (p,k), contract j -> q ---------------------------------
    if (in_elem && l < NQ * NM) {
        double y[NM], b[NQ];
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            y[j] = bA[X(l_m, j, l_m2)];
        }
        contract_cols<NQ, SB>(launder<2>(B, dep), y, b);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            bB[X(l_m, q, l_m2)] = b[q];
        }
    }
    __syncthreads();

    // ---- P3: lane (p,q), contract k -> r ---------------------------------
    double W[NQ];
    if (in_elem) {
        double y[NM], U[NQ];
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            y[k] = bB[X(l_q, l_q2, k)];
        }
        contract_cols<NQ, SB>(launder<3>(B, dep), y, U);
        if constexpr (!BK3) {
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                W[r] = U[r] * bC[X(l_q, l_q2, r)];
            }
        } else {
            double qt[NQ];
            contract_cols<NQ, SB>(launder<4>(D, dep), U, qt);
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                bA[X(l_q, l_q2, r)] = U[r];
                bC[X(l_q, l_q2, r)] = qt[r];
            }
        }
    }
    if constexpr (BK3) {
        __syncthreads();
        // ---- P4: qs in lane (p,r) col q; qr in lane (q,r) col p ----------
        double qr[NQ];
        if (in_elem) {
            double y[NQ], z[NQ], qs[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                y[q] = bA[X(l_q, q, l_q2)];      // lane (p, r), col q
                z[q] = bA[X(q, l_q, l_q2)];      // lane (q, r), col p
            }
            contract_cols2<NQ, SB>(launder<5>(D, dep), y, qs, z, qr);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                bB[X(l_q, q, l_q2)] = qs[q];
            }
        }
        __syncthreads();
        // ---- P5: fluxes in lane (q,r) col p ---------------------------------
        double rr[NQ];
        if (in_elem) {
            const double* ge = Qe + l;   // G(e, f, p, q, r) at f = p = 0
            double gn[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                gn[f] = active ? ge[f * S::nq3] : 0.0;
            }
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                double g[6];
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    g[f] = gn[f];
                }
                if (p + 1 < NQ) {
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        gn[f] = active ? ge[f * S::nq3 + (p + 1) * nq2] : 0.0;
                    }
                }
                const int x = X(p, l_q, l_q2);
                const double qs = bB[x];
                const double qt = bC[x];
                rr[p] = g[0] * qr[p] + g[1] * qs + g[2] * qt;
                bB[x] = g[1] * qr[p] + g[3] * qs + g[4] * qt;
                bC[x] = g[2] * qr[p] + g[4] * qs + g[5] * qt;
            }
        }
        __syncthreads();
        // ---- P6: transposed gradient ---------------------------------------
        double Ws[NQ], Wt[NQ];
        if (in_elem) {
            double Wr[NQ], y[NQ], z[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                y[q] = bB[X(l_q, q, l_q2)];      // rs, lane (p, r), col q
                z[q] = bC[X(l_q, l_q2, q)];      // rt, lane (p, q), col r
            }
            // Wr: lane (q, r), col p;  Ws: lane (p, r), col q;  Wt: lane (p, q), col r
            contract_rows3<NQ, SB>(launder<7>(D, dep), rr, Wr, y, Ws, z, Wt);
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                bA[X(p, l_q, l_q2)] = Wr[p];
            }
        }
        __syncthreads();
        if (in_elem) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                bB[X(l_q, q, l_q2)] = Ws[q];
            }
        }
        __syncthreads();
        if (in_elem) {
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                const int x = X(l_q, l_q2, r);
                W[r] = Wt[r] + bA[x] + bB[x];
            }
        }
    }

    // ---- P7: lane (p,q), contract r -> k ---------------------------------
    if (in_elem) {
        double Y[NM];
        contract_rows<NQ, SB>(launder<10>(B, dep), W, Y);
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            bA[X(l_q, l_q2, k)] = Y[k];
        }
    }
    __syncthreads();

    // ---- P8: lane (p,k), contract q -> j ---------------------------------
    if (in_elem && l < NQ * NM) {
        double y[NQ], Z[NM];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            y[q] = bA[X(l_m, q, l_m2)];
        }
        contract_rows<NQ, SB>(launder<11>(B, dep), y, Z);
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            bB[X(l_m, j, l_m2)] = Z[j];
        }
    }
    __syncthreads();

    // ---- P9: lane (j,k), contract p -> i, store --------------------------
    if (active && l < S::nm2) {
        double z[NQ], o[NM];
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            z[p] = bB[X(p, l_m, l_m2)];
        }
        contract_rows<NQ, SB>(launder<12>(B, dep), z, o);
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            ve[i * S::nm2 + l] = o[i];
        }
    }
}

}  // namespace bk13t
}  // namespace bkx
