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

// What a lane loads from HBM before an element's phases start: its u
// column u(e, :, j, k) (lanes (j, k) < nm^2), and BK1's JxW(e, t, lane) for
// t < nq (staged into LDS in P1/P2), or BK3's first GPF rows G(e, :, p, q, r)
// for p < GPF (used in P5; with GIN false they are loaded at the start of P3
// instead).  bk13t_load issues the loads; the registers are read when the
// phases get to them, so the loads are in flight meanwhile.
template <int NM, int NQ, bool BK3, int GPF>
struct Bk13tIn {
    static constexpr int nqv = BK3 ? (GPF > 0 ? GPF * 6 : 1) : NQ;
    double x[NM];
    double q[nqv];
};

// ue, Qe: this lane's element's u and JxW/G.  When the lane's element is not
// active they must still point at a whole readable element of the same
// order: the loads are unconditional (no branch around each load).
template <int NM, int NQ, bool BK3, int GPF>
__device__ __forceinline__ void bk13t_load_g(Bk13tIn<NM, NQ, BK3, GPF>& in, const double* __restrict__ Qe,
                                             bool in_elem, int l)
{
    constexpr int nq2 = NQ * NQ;
    constexpr int nq3 = nq2 * NQ;
    if constexpr (BK3 && GPF > 0) {
        if (in_elem) {
#pragma unroll
            for (int p = 0; p < GPF; ++p) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    in.q[p * 6 + f] = Qe[l + f * nq3 + p * nq2];
                }
            }
        }
    }
}

template <int NM, int NQ, bool BK3, int GPF, bool GIN = true>
__device__ __forceinline__ void bk13t_load(Bk13tIn<NM, NQ, BK3, GPF>& in, const double* __restrict__ ue,
                                           const double* __restrict__ Qe, bool in_elem, int l)
{
    constexpr int nm2 = NM * NM;
    constexpr int nq2 = NQ * NQ;
    constexpr int nq3 = nq2 * NQ;
    if (in_elem && l < nm2) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            in.x[i] = ue[i * nm2 + l];
        }
    }
    if constexpr (!BK3) {
        if (in_elem) {
#pragma unroll
            for (int t = 0; t < NQ; ++t) {
                in.q[t] = Qe[t * nq2 + l];
            }
        }
    } else if constexpr (GIN) {
        bk13t_load_g(in, Qe, in_elem, l);
    }
    (void)nq3;
}

// One workgroup's pass over up to epb elements, from the loads in `in`.  Qe
// and ve are this lane's element's JxW/G and v (Qe readable as in
// bk13t_load; v is stored only when active).  With `next`, the loads of the
// lane's next element (next_ue, next_Qe) are issued into `in` at the start
// of P7, when the registers of `in` are free, so that they overlap P7-P9 of
// this element.  smem holds epb * NB boxes.
//
// NB = 3: boxes A, B, C as in the phase table of bk13t.hip.
// NB = 2: boxes A and B only (two thirds of the LDS, so more workgroups per
// CU), at the cost of two more barriers:
//   BK1: JxW is staged into box A after every lane has read box A in P2.
//   BK3: P3 stores U to box A, then (after a barrier) qt to box B, which it
//        has just read; P4 stores qs to box A after every lane has read U.
//        P5 then works in place on qs (box A) and qt (box B), and P6 writes
//        Wr and Ws after a barrier.
// GPF > 0 (BK3): the first GPF rows of G come with `in`, and P5 keeps GPF
// rows in flight (row p + GPF is loaded when row p is used).  With GPF = 0,
// P5 loads one row ahead.
template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF, bool GIN = true>
__device__ __forceinline__ void bk13t_run(const double* __restrict__ B, const double* __restrict__ D,
                                          Bk13tIn<NM, NQ, BK3, GPF>& in, const double* __restrict__ Qe,
                                          double* __restrict__ ve, bool active, int slot, int l, double* smem,
                                          bool next, const double* __restrict__ next_ue,
                                          const double* __restrict__ next_Qe, int dep = 0)
{
    static_assert(NB == 2 || NB == 3, "two or three LDS boxes per element");
    static_assert(GPF >= 0 && GPF <= NQ, "G prefetch depth");
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr int nq2 = S::nq2;
    constexpr int R = S::row;
    const bool in_elem = slot < S::epb;
    const int sl = in_elem ? slot : 0;
    double* bA = smem + (sl * NB + 0) * S::box;
    double* bB = smem + (sl * NB + 1) * S::box;
    // box C: the third box, or (NB = 2) the box holding qt and rt, which is B
    double* bC = smem + (sl * NB + (NB == 3 ? 2 : 1)) * S::box;
    // box holding qs and rs: B (NB = 3) or A (NB = 2)
    double* bS = NB == 3 ? bB : bA;
    auto X = [](int a, int b, int c) { return (a * NQ + b) * R + c; };
    // lane (a, b) of an NA x NB plane, when l < NA * NB
    const int l_m = l / NM, l_m2 = l - (l / NM) * NM;    // (l / NM, l % NM)
    const int l_q = l / NQ, l_q2 = l - (l / NQ) * NQ;    // (l / NQ, l % NQ)

    // ---- P1: lane (j,k), contract i -> p ---------------------------------
    if (in_elem && l < S::nm2) {
        double a[NQ];
        contract_cols<NQ, SB>(launder<1>(B, dep), in.x, a);
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            bA[X(p, l_m, l_m2)] = a[p];
        }
    }
    // BK1: JxW(e, p, q, r), loaded with unit stride, into box C now (NB = 3)
    // or into box A after P2 has read it (NB = 2)
    if constexpr (!BK3 && NB == 3) {
        if (in_elem) {
#pragma unroll
            for (int t = 0; t < NQ; ++t) {
                const int x = t * nq2 + l;
                bC[x + x / NQ] = in.q[t];
            }
        }
    }
    __syncthreads();

    // ---- P2: lane (p,k), contract j -> q ---------------------------------
    {
        double b[NQ];
        const bool p2 = in_elem && l < NQ * NM;
        if (p2) {
            double y[NM];
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                y[j] = bA[X(l_m, j, l_m2)];
            }
            contract_cols<NQ, SB>(launder<2>(B, dep), y, b);
        }
        if constexpr (!BK3 && NB == 2) {
            __syncthreads();
            if (in_elem) {
#pragma unroll
                for (int t = 0; t < NQ; ++t) {
                    const int x = t * nq2 + l;
                    bA[x + x / NQ] = in.q[t];
                }
            }
        }
        if (p2) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                bB[X(l_m, q, l_m2)] = b[q];
            }
        }
    }
    __syncthreads();
    // BK1: box holding JxW in P3
    const double* bJ = NB == 3 ? bC : bA;

    // ---- P3: lane (p,q), contract k -> r ---------------------------------
    double W[NQ];
    const double* ge = Qe + l;   // BK3: G(e, f, p, q, r) at f = p = 0, (q, r) = lane
    if constexpr (!GIN) {
        bk13t_load_g(in, Qe, in_elem, l);
    }
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
                W[r] = U[r] * bJ[X(l_q, l_q2, r)];
            }
        } else {
            contract_cols<NQ, SB>(launder<4>(D, dep), U, W);   // qt, held in W
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                bA[X(l_q, l_q2, r)] = U[r];
                if constexpr (NB == 3) {
                    bC[X(l_q, l_q2, r)] = W[r];
                }
            }
        }
    }
    if constexpr (BK3) {
        __syncthreads();
        if constexpr (NB == 2) {
            // every lane has read box B (b): qt -> box B (= C)
            if (in_elem) {
#pragma unroll
                for (int r = 0; r < NQ; ++r) {
                    bC[X(l_q, l_q2, r)] = W[r];
                }
            }
        }
        // ---- P4: qs in lane (p,r) col q; qr in lane (q,r) col p ----------
        double qr[NQ];
        {
            double qs[NQ];
            if (in_elem) {
                double y[NQ], z[NQ];
                if constexpr (NB == 3) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bA[X(l_q, q, l_q2)];      // lane (p, r), col q
                        z[q] = bA[X(q, l_q, l_q2)];      // lane (q, r), col p
                    }
                    contract_cols2<NQ, SB>(launder<5>(D, dep), y, qs, z, qr);
                } else {
                    // one input at a time (fewer registers live; D's rows
                    // are read twice from the scalar cache)
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bA[X(l_q, q, l_q2)];
                    }
                    contract_cols<NQ, SB>(launder<5>(D, dep), y, qs);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        z[q] = bA[X(q, l_q, l_q2)];
                    }
                    contract_cols<NQ, SB>(launder<6>(D, dep), z, qr);
                }
            }
            if constexpr (NB == 2) {
                __syncthreads();   // every lane has read U from box A
            }
            if (in_elem) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    bS[X(l_q, q, l_q2)] = qs[q];
                }
            }
        }
        __syncthreads();
        // ---- P5: fluxes in lane (q,r) col p ---------------------------------
        double rr[NQ];
        if (in_elem) {
            double gn[6];
            if constexpr (GPF == 0) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    gn[f] = ge[f * S::nq3];
                }
            }
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                double g[6];
                if constexpr (GPF == 0) {
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        g[f] = gn[f];
                    }
                    if (p + 1 < NQ) {
#pragma unroll
                        for (int f = 0; f < 6; ++f) {
                            gn[f] = ge[f * S::nq3 + (p + 1) * nq2];
                        }
                    }
                } else {
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        g[f] = in.q[(p % GPF) * 6 + f];
                    }
                    if (p + GPF < NQ) {
#pragma unroll
                        for (int f = 0; f < 6; ++f) {
                            in.q[(p % GPF) * 6 + f] = ge[f * S::nq3 + (p + GPF) * nq2];
                        }
                    }
                }
                const int x = X(p, l_q, l_q2);
                const double qs = bS[x];
                const double qt = bC[x];
                rr[p] = g[0] * qr[p] + g[1] * qs + g[2] * qt;
                bS[x] = g[1] * qr[p] + g[3] * qs + g[4] * qt;
                bC[x] = g[2] * qr[p] + g[4] * qs + g[5] * qt;
                if constexpr (GPF > 0) {
                    // keep the rows' loads where they are written: GPF rows
                    // in flight, not all of G hoisted to the top of P5
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        __syncthreads();
        // ---- P6: transposed gradient ---------------------------------------
        double Ws[NQ], Wt[NQ];
        {
            double Wr[NQ];
            if (in_elem) {
                double y[NQ], z[NQ];
                // Wr: lane (q, r), col p;  Ws: lane (p, r), col q;  Wt: lane (p, q), col r
                if constexpr (NB == 3) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bS[X(l_q, q, l_q2)];      // rs, lane (p, r), col q
                        z[q] = bC[X(l_q, l_q2, q)];      // rt, lane (p, q), col r
                    }
                    contract_rows3<NQ, SB>(launder<7>(D, dep), rr, Wr, y, Ws, z, Wt);
                } else {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        z[q] = bC[X(l_q, l_q2, q)];
                    }
                    contract_rows<NQ, SB>(launder<7>(D, dep), z, Wt);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bS[X(l_q, q, l_q2)];
                    }
                    contract_rows<NQ, SB>(launder<8>(D, dep), y, Ws);
                    contract_rows<NQ, SB>(launder<9>(D, dep), rr, Wr);
                }
            }
            if constexpr (NB == 2) {
                __syncthreads();   // every lane has read rs and rt
            }
            if (in_elem) {
#pragma unroll
                for (int p = 0; p < NQ; ++p) {
                    bA[X(p, l_q, l_q2)] = Wr[p];
                }
            }
        }
        if constexpr (NB == 3) {
            __syncthreads();
        }
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
    if (next) {
        bk13t_load<NM, NQ, BK3, GPF, GIN>(in, next_ue, next_Qe, in_elem, l);
    }
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

// One workgroup's pass over up to epb elements: bk13t_load, then bk13t_run.
// ue, Qe and ve are this lane's element's u, JxW/G and v (ue and Qe
// readable as in bk13t_load when !active).
template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB = 3, int GPF = 0>
__device__ __forceinline__ void bk13t_body(const double* __restrict__ B, const double* __restrict__ D,
                                           const double* __restrict__ Qe, const double* __restrict__ ue,
                                           double* __restrict__ ve, bool active, int slot, int l, double* smem,
                                           int dep = 0)
{
    Bk13tIn<NM, NQ, BK3, GPF> in;
    bk13t_load(in, ue, Qe, slot < Bk13tShape<NM, NQ, DIV>::epb, l);
    bk13t_run<NM, NQ, BK3, DIV, SB, NB, GPF>(B, D, in, Qe, ve, active, slot, l, smem, false, nullptr, nullptr, dep);
}

}  // namespace bk13t
}  // namespace bkx
