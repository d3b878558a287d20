// bk13t_body.hpp -- the transposing BK1/BK3 mapping as a device function
// (one workgroup's pass over up to epb elements), shared by the contiguous
// kernels of bk13t.hip and the mixed-order kernels of bk3_mixed.hip.
// The mapping is described in bk13t.hip.
#pragma once

#include "bkx.h"
#include "bkx_device.hpp"
#include "bkx_eo.hpp"
#include "bkx_ff.hpp"

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
    // an odd row stride (in 8-byte words) spreads every lane plane's column
    // accesses over the LDS banks: NQ + 1 when NQ is even, NQ when it is odd
    // (NQ + 1 even doubles the bank conflicts of the transposes at odd NQ)
    static constexpr int row = NQ | 1;
    static constexpr int box = NQ * NQ * row;
};

// ---- arithmetic forms ------------------------------------------------------
// FORM (enum bkx_form): storage type TS of u and JxW/G, value type V of the
// registers and LDS boxes, and the type L a loaded value is kept in:
//   F64   TS double, V double, L double   (the reference's SumFactorization<double>)
//   F32W  TS float,  V double, L float    (FP32 storage, FP64 arithmetic)
//   F32C  TS float,  V ff,     L float    (FP32 storage, float-float arithmetic)
//   FF    TS double, V ff,     L ff       (FP64 storage split into float-float
//                                          when loaded, float-float arithmetic:
//                                          the reference's BK1_ff kernels)
// v is stored in FP64 by every form (a float-float value joined with one add).
// B and D are FP64 for the double forms (wavefront-uniform scalar loads) and
// float-float for the others (a per-workgroup LDS copy, split once).
template <int FORM>
struct Form;
template <>
struct Form<BKX_FORM_F64> {
    using TS = double;
    using V = double;
    using L = double;
};
template <>
struct Form<BKX_FORM_F32W> {
    using TS = float;
    using V = double;
    using L = float;
};
template <>
struct Form<BKX_FORM_F32C> {
    using TS = float;
    using V = ff;
    using L = float;
};
template <>
struct Form<BKX_FORM_FF> {
    using TS = double;
    using V = ff;
    using L = ff;
};

// loads of the streamed inputs (u, JxW, G), non-temporal when NT is set
template <bool NT>
__device__ __forceinline__ double load_as(const double* p, double*) { return load<NT>(p); }
template <bool NT>
__device__ __forceinline__ float load_as(const float* p, float*) { return load<NT>(p); }
template <bool NT>
__device__ __forceinline__ ff load_as(const double* p, ff*) { return ff_split(load<NT>(p)); }

// Sum of products, in ascending order of the terms.  FP64: a running FMA
// chain (the first product rounded alone: -0.0 + x is x for every x);
// float-float: a Dot2 accumulator.
template <typename V>
struct Acc;
template <>
struct Acc<double> {
    double a = -0.0;
    __device__ __forceinline__ void mac(double m, double x) { a += m * x; }
    __device__ __forceinline__ void mac(double m, float x) { a += m * double(x); }
    __device__ __forceinline__ double get() const { return a; }
};
template <>
struct Acc<ff> {
    dot2 d;
    __device__ __forceinline__ void mac(ff m, ff x) { d.add(m, x); }
    __device__ __forceinline__ void mac(ff m, float x) { d.add(m, x); }
    __device__ __forceinline__ ff get() const { return d.value(); }
};

__device__ __forceinline__ double sum3(double a, double b, double c) { return a + b + c; }
__device__ __forceinline__ ff sum3(ff a, ff b, ff c)
{
    dot2 d;
    d.add(a);
    d.add(b);
    d.add(c);
    return d.value();
}

__device__ __forceinline__ double mul(double a, double b) { return a * b; }
__device__ __forceinline__ ff mul(ff a, ff b)
{
    dot2 d;
    d.add(a, b);
    return d.value();
}

// a loaded value as a value of the form
__device__ __forceinline__ double to_v(double x, double*) { return x; }
__device__ __forceinline__ double to_v(float x, double*) { return double(x); }
__device__ __forceinline__ ff to_v(float x, ff*) { return ff{x, 0.0f}; }
__device__ __forceinline__ ff to_v(ff x, ff*) { return x; }

template <bool NT>
__device__ __forceinline__ void store_v(double* p, double x)
{
    if constexpr (NT) {
        st_stream(p, x);
    } else {
        *p = x;
    }
}
template <bool NT>
__device__ __forceinline__ void store_v(double* p, ff x)
{
    store_v<NT>(p, ff_join(x));
}

// the rows of B or D: FP64 through a scalar-load pointer, or the LDS copy
template <int TAG>
__device__ __forceinline__ global_ptr<double> mat(const double* p, int dep)
{
    return launder<TAG>(p, dep);
}
template <int TAG>
__device__ __forceinline__ const ff* mat(const ff* p, int)
{
    return p;
}

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
template <int NQ, bool SB, int NIN, int NOUT, typename P, typename X, typename V>
__device__ __forceinline__ void contract_cols(P M, const X (&in)[NIN], V (&out)[NOUT])
{
    Acc<V> acc[NOUT];
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            acc[c].mac(M[m * NQ + c], in[m]);
        }
        row_fence<SB>();
    }
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out[c] = acc[c].get();
    }
}

// the same for two inputs sharing the rows of M
template <int NQ, bool SB, int NIN, int NOUT, typename P, typename V>
__device__ __forceinline__ void contract_cols2(P M, const V (&in0)[NIN], V (&out0)[NOUT], const V (&in1)[NIN],
                                               V (&out1)[NOUT])
{
    Acc<V> acc0[NOUT], acc1[NOUT];
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            acc0[c].mac(M[m * NQ + c], in0[m]);
            acc1[c].mac(M[m * NQ + c], in1[m]);
        }
        row_fence<SB>();
    }
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out0[c] = acc0[c].get();
        out1[c] = acc1[c].get();
    }
}

// out[c] = sum_m M[c * NQ + m] in[m] (c < NOUT): row c of M dotted with in.
template <int NQ, bool SB, int NIN, int NOUT, typename P, typename V>
__device__ __forceinline__ void contract_rows(P M, const V (&in)[NIN], V (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        Acc<V> t;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t.mac(M[c * NQ + m], in[m]);
        }
        out[c] = t.get();
        row_fence<SB>();
    }
}

// the same for three inputs sharing the rows of M
template <int NQ, bool SB, int NIN, int NOUT, typename P, typename V>
__device__ __forceinline__ void contract_rows3(P M, const V (&in0)[NIN], V (&out0)[NOUT], const V (&in1)[NIN],
                                               V (&out1)[NOUT], const V (&in2)[NIN], V (&out2)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        Acc<V> t0, t1, t2;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t0.mac(M[c * NQ + m], in0[m]);
            t1.mac(M[c * NQ + m], in1[m]);
            t2.mac(M[c * NQ + m], in2[m]);
        }
        out0[c] = t0.get();
        out1[c] = t1.get();
        out2[c] = t2.get();
        row_fence<SB>();
    }
}

// ---- even-odd contractions (EoMat, bkx_eo.hpp) -----------------------------
template <int NQ, bool SB, int NIN, int NOUT, int R, int C, int S, typename X, typename V>
__device__ __forceinline__ void contract_cols(EoMat<R, C, S> M, const X (&in)[NIN], V (&out)[NOUT])
{
    static_assert(NIN == R && NOUT == C, "the even-odd form contracts a whole matrix");
    using E = EoMat<R, C, S>;
    V e[E::hr > 0 ? E::hr : 1], o[E::hr > 0 ? E::hr : 1];
#pragma unroll
    for (int m = 0; m < E::hr; ++m) {
        e[m] = V(in[m]) + V(in[R - 1 - m]);
        o[m] = V(in[m]) - V(in[R - 1 - m]);
    }
#pragma unroll
    for (int c = 0; c < E::cc; ++c) {
        const bool mid = (C % 2 == 1) && c == C / 2;
        Acc<V> a, b;
        if (!mid || S == 1) {
#pragma unroll
            for (int m = 0; m < E::hr; ++m) {
                a.mac(M.p[E::cols_p + m * E::cc + c], e[m]);
            }
            if constexpr (R % 2 == 1) {
                a.mac(M.p[E::cols_mid + c], V(in[R / 2]));
            }
        }
        if (!mid || S == -1) {
#pragma unroll
            for (int m = 0; m < E::hr; ++m) {
                b.mac(M.p[E::cols_q + m * E::cc + c], o[m]);
            }
        }
        if (mid) {
            out[c] = S == 1 ? a.get() : b.get();
        } else {
            const V A = a.get(), B = b.get();
            out[c] = A + B;
            out[C - 1 - c] = S == 1 ? A - B : B - A;
        }
    }
}

template <int NQ, bool SB, int NIN, int NOUT, int R, int C, int S, typename V>
__device__ __forceinline__ void contract_rows(EoMat<R, C, S> M, const V (&in)[NIN], V (&out)[NOUT])
{
    static_assert(NIN == C && NOUT == R, "the even-odd form contracts a whole matrix");
    using E = EoMat<R, C, S>;
    V e[E::hc > 0 ? E::hc : 1], o[E::hc > 0 ? E::hc : 1];
#pragma unroll
    for (int m = 0; m < E::hc; ++m) {
        e[m] = in[m] + in[C - 1 - m];
        o[m] = in[m] - in[C - 1 - m];
    }
#pragma unroll
    for (int c = 0; c < E::cr; ++c) {
        const bool mid = (R % 2 == 1) && c == R / 2;
        Acc<V> a, b;
        if (!mid || S == 1) {
#pragma unroll
            for (int m = 0; m < E::hc; ++m) {
                a.mac(M.p[E::rows_p + m * E::cr + c], e[m]);
            }
            if constexpr (C % 2 == 1) {
                a.mac(M.p[E::rows_mid + c], in[C / 2]);
            }
        }
        if (!mid || S == -1) {
#pragma unroll
            for (int m = 0; m < E::hc; ++m) {
                b.mac(M.p[E::rows_q + m * E::cr + c], o[m]);
            }
        }
        if (mid) {
            out[c] = S == 1 ? a.get() : b.get();
        } else {
            const V A = a.get(), B = b.get();
            out[c] = A + B;
            out[R - 1 - c] = S == 1 ? A - B : B - A;
        }
    }
}

template <int NQ, bool SB, int NIN, int NOUT, int R, int C, int S, typename V>
__device__ __forceinline__ void contract_cols2(EoMat<R, C, S> M, const V (&in0)[NIN], V (&out0)[NOUT],
                                               const V (&in1)[NIN], V (&out1)[NOUT])
{
    contract_cols<NQ, SB>(M, in0, out0);
    contract_cols<NQ, SB>(M, in1, out1);
}

template <int NQ, bool SB, int NIN, int NOUT, int R, int C, int S, typename V>
__device__ __forceinline__ void contract_rows3(EoMat<R, C, S> M, const V (&in0)[NIN], V (&out0)[NOUT],
                                               const V (&in1)[NIN], V (&out1)[NOUT], const V (&in2)[NIN],
                                               V (&out2)[NOUT])
{
    contract_rows<NQ, SB>(M, in0, out0);
    contract_rows<NQ, SB>(M, in1, out1);
    contract_rows<NQ, SB>(M, in2, out2);
}

// the matrix operand of a contraction with B (nm x nq, centro-symmetric) or
// D (nq x nq, centro-antisymmetric): rows through the scalar cache, or
// (EO) the even-odd form at the same pointer
template <int TAG, bool EO, int NM, int NQ, typename MT>
__device__ __forceinline__ auto mat_b(const MT* p, int dep)
{
    if constexpr (EO) {
        return EoMat<NM, NQ, 1>{launder<TAG>(p, dep)};
    } else {
        return mat<TAG>(p, dep);
    }
}
template <int TAG, bool EO, int NQ, typename MT>
__device__ __forceinline__ auto mat_d(const MT* p, int dep)
{
    if constexpr (EO) {
        return EoMat<NQ, NQ, -1>{launder<TAG>(p, dep)};
    } else {
        return mat<TAG>(p, dep);
    }
}

// B (nm x nq) and D (nq x nq) split into float-float in LDS by the whole
// workgroup (the float-float forms); the caller synchronises afterwards.
template <int NM, int NQ, bool BK3>
__device__ __forceinline__ void split_bases(const double* __restrict__ B, const double* __restrict__ D, ff* sB,
                                            ff* sD, int tid, int nthreads)
{
    for (int x = tid; x < NM * NQ; x += nthreads) {
        sB[x] = ff_split(B[x]);
    }
    if constexpr (BK3) {
        for (int x = tid; x < NQ * NQ; x += nthreads) {
            sD[x] = ff_split(D[x]);
        }
    }
}

// Values of JxW / G a lane holds ahead of use: BK1 all nq of its JxW(e, :, lane),
// BK3 GPF rows of six factors (one when GPF = 0).
constexpr int bk13t_nqv(bool bk3, int gpf, int nq) { return bk3 ? (gpf > 0 ? gpf * 6 : 1) : nq; }

// One workgroup's pass over up to epb elements.  ue, Qe and ve are this
// lane's element's u, JxW/G and v.  When !active they must still point at a
// whole readable element of the same order: the loads are unconditional
// (no branch around each load) and only the store is guarded.  MB, MD are
// the rows of B and D (FP64 device pointers, or the float-float LDS copy of
// split_bases for the float-float forms); smem holds epb * NB boxes.
//
// The lane's u column u(e, :, j, k) (lanes (j, k) < nm^2), and BK1's
// JxW(e, t, lane) for t < nq or BK3's first GPF rows G(e, :, p, q, r),
// p < GPF, are loaded at the start, so that they are in flight through the
// first phases.
//
// NB = 3: boxes A, B, C as in the phase table of bk13t.hip.
// NB = 2: boxes A and B only (two thirds of the LDS, so more workgroups per
// CU), at the cost of two more barriers:
//   BK1: JxW is staged into box A after every lane has read box A in P2.
//   BK3: P3 stores U to box A, then (after a barrier) qt to box B, which it
//        has just read; P4 stores qs to box A after every lane has read U.
//        P5 then works in place on qs (box A) and qt (box B), and P6 writes
//        Wr and Ws after a barrier.
// GPF > 0 (BK3): P5 keeps GPF rows of G in flight (row p + GPF is loaded
// when row p is used).  With GPF = 0, P5 loads one row ahead.
// NT: bit 0 makes the loads of u and the stores of v non-temporal, bit 1
// the loads of JxW / G.  A G plane G(e, f, p, :, :) is nq^2 values, which
// for most orders begins and ends inside a 128-byte line shared with the
// next plane; the next row reads that line again a little later, so
// streaming G can cost a second fetch of it.
//
// The body is bk13t_load (the loads above, into xin and qin) followed by
// bk13t_phases.  With PIPE (BK3, GPF > 0: the persistent kernel of
// bk13t.hip) the phases also load the workgroup's next element: the last
// GPF refills of the G ring in P5 fetch rows 0 .. GPF - 1 of the next
// element (Qn), and after P5 the lane's next u column (un) replaces xin, so
// that the next pass starts with its first loads already in flight.  In that
// mode the loads of u are unconditional (lanes outside the element read a
// valid column, clamped), the prologue loads G before u, as the phases do,
// and every lane stores in P9 (the kernel waits for its prologue's loads
// before the loop): each pass then ends with the same loads and stores in
// flight, and the compiler's wait for the next element's loads at the top of
// the loop counts the stores instead of waiting for them.
// With EO (FP64 arithmetic) B and D point at their even-odd forms (EoMat),
// and every contraction runs in that form; the phases are unchanged.
template <int NM, int NQ, bool BK3, int DIV, int GPF, int FORM, int NT, bool PIPE>
__device__ __forceinline__ void bk13t_load(const typename Form<FORM>::TS* __restrict__ Qe,
                                           const typename Form<FORM>::TS* __restrict__ ue, int slot, int l,
                                           typename Form<FORM>::L (&xin)[NM],
                                           typename Form<FORM>::L (&qin)[bk13t_nqv(BK3, GPF, NQ + 0)])
{
    using L = typename Form<FORM>::L;
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr int nq2 = S::nq2;
    const bool in_elem = slot < S::epb;
    L* const tagL = nullptr;   // selects load_as overloads
    const auto* ge = Qe + l;   // BK3: G(e, f, p, q, r) at f = p = 0, (q, r) = lane
    if constexpr (PIPE) {
        static_assert(BK3 && GPF > 0, "the pipelined body streams BK3's G ring");
#pragma unroll
        for (int p = 0; p < GPF; ++p) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                qin[p * 6 + f] = load_as<(NT & 2) != 0>(ge + f * S::nq3 + p * nq2, tagL);
            }
        }
        const int lu = l < S::nm2 ? l : 0;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            xin[i] = load_as<(NT & 1) != 0>(ue + i * S::nm2 + lu, tagL);
        }
        return;
    }
    if (in_elem && l < S::nm2) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            xin[i] = load_as<(NT & 1) != 0>(ue + i * S::nm2 + l, tagL);
        }
    }
    if (in_elem) {
        if constexpr (!BK3) {
#pragma unroll
            for (int t = 0; t < NQ; ++t) {
                qin[t] = load_as<(NT & 2) != 0>(Qe + t * nq2 + l, tagL);
            }
        } else if constexpr (GPF > 0) {
#pragma unroll
            for (int p = 0; p < GPF; ++p) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    qin[p * 6 + f] = load_as<(NT & 2) != 0>(ge + f * S::nq3 + p * nq2, tagL);
                }
            }
        }
    }
}

template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF, int FORM, typename MT, int NT, bool PIPE,
          bool EO = false>
__device__ __forceinline__ void bk13t_phases(const MT* __restrict__ B, const MT* __restrict__ D,
                                             const typename Form<FORM>::TS* __restrict__ Qe,
                                             double* __restrict__ ve, bool active, int slot, int l, double* smem,
                                             int dep, typename Form<FORM>::L (&xin)[NM],
                                             typename Form<FORM>::L (&qin)[bk13t_nqv(BK3, GPF, NQ + 0)],
                                             const typename Form<FORM>::TS* __restrict__ Qn,
                                             const typename Form<FORM>::TS* __restrict__ un)
{
    static_assert(NB == 2 || NB == 3, "two or three LDS boxes per element");
    static_assert(GPF >= 0 && GPF <= NQ, "G prefetch depth");
    using V = typename Form<FORM>::V;
    using L = typename Form<FORM>::L;
    static_assert(sizeof(V) == sizeof(double), "an LDS box holds 8-byte values");
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr int nq2 = S::nq2;
    constexpr int R = S::row;
    const bool in_elem = slot < S::epb;
    const int sl = in_elem ? slot : 0;
    V* bA = reinterpret_cast<V*>(smem) + (sl * NB + 0) * S::box;
    V* bB = reinterpret_cast<V*>(smem) + (sl * NB + 1) * S::box;
    // box C: the third box, or (NB = 2) the box holding qt and rt, which is B
    V* bC = reinterpret_cast<V*>(smem) + (sl * NB + (NB == 3 ? 2 : 1)) * S::box;
    // box holding qs and rs: B (NB = 3) or A (NB = 2)
    V* bS = NB == 3 ? bB : bA;
    auto X = [](int a, int b, int c) { return (a * NQ + b) * R + c; };
    // lane (a, b) of an NA x NB plane, when l < NA * NB
    const int l_m = l / NM, l_m2 = l - (l / NM) * NM;    // (l / NM, l % NM)
    const int l_q = l / NQ, l_q2 = l - (l / NQ) * NQ;    // (l / NQ, l % NQ)
    L* const tagL = nullptr;   // selects load_as overloads
    V* const tagV = nullptr;

    const auto* ge = Qe + l;   // BK3: G(e, f, p, q, r) at f = p = 0, (q, r) = lane

    // ---- P1: lane (j,k), contract i -> p ---------------------------------
    if (in_elem && l < S::nm2) {
        V a[NQ];
        contract_cols<NQ, SB>(mat_b<1, EO, NM, NQ>(B, dep), xin, a);
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
                bC[x + (x / NQ) * (R - NQ)] = to_v(qin[t], tagV);
            }
        }
    }
    __syncthreads();

    // ---- P2: lane (p,k), contract j -> q ---------------------------------
    {
        V b[NQ];
        const bool p2 = in_elem && l < NQ * NM;
        if (p2) {
            V y[NM];
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                y[j] = bA[X(l_m, j, l_m2)];
            }
            contract_cols<NQ, SB>(mat_b<2, EO, NM, NQ>(B, dep), y, b);
        }
        if constexpr (!BK3 && NB == 2) {
            __syncthreads();
            if (in_elem) {
#pragma unroll
                for (int t = 0; t < NQ; ++t) {
                    const int x = t * nq2 + l;
                    bA[x + (x / NQ) * (R - NQ)] = to_v(qin[t], tagV);
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
    const V* bJ = NB == 3 ? bC : bA;

    // ---- P3: lane (p,q), contract k -> r ---------------------------------
    V W[NQ];
    if (in_elem) {
        V y[NM], U[NQ];
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            y[k] = bB[X(l_q, l_q2, k)];
        }
        contract_cols<NQ, SB>(mat_b<3, EO, NM, NQ>(B, dep), y, U);
        if constexpr (!BK3) {
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                W[r] = mul(U[r], bJ[X(l_q, l_q2, r)]);
            }
        } else {
            contract_cols<NQ, SB>(mat_d<4, EO, NQ>(D, dep), U, W);   // qt, held in W
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
        V qr[NQ];
        {
            V qs[NQ];
            if (in_elem) {
                V y[NQ], z[NQ];
                if constexpr (NB == 3) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bA[X(l_q, q, l_q2)];      // lane (p, r), col q
                        z[q] = bA[X(q, l_q, l_q2)];      // lane (q, r), col p
                    }
                    contract_cols2<NQ, SB>(mat_d<5, EO, NQ>(D, dep), y, qs, z, qr);
                } else {
                    // one input at a time (fewer registers live; D's rows
                    // are read twice from the scalar cache)
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bA[X(l_q, q, l_q2)];
                    }
                    contract_cols<NQ, SB>(mat_d<5, EO, NQ>(D, dep), y, qs);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        z[q] = bA[X(q, l_q, l_q2)];
                    }
                    contract_cols<NQ, SB>(mat_d<6, EO, NQ>(D, dep), z, qr);
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
        V rr[NQ];
        if (in_elem) {
            L gn[6];
            if constexpr (GPF == 0) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    gn[f] = load_as<(NT & 2) != 0>(ge + f * S::nq3, tagL);
                }
            }
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                L g[6];
                if constexpr (GPF == 0) {
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        g[f] = gn[f];
                    }
                    if (p + 1 < NQ) {
#pragma unroll
                        for (int f = 0; f < 6; ++f) {
                            gn[f] = load_as<(NT & 2) != 0>(ge + f * S::nq3 + (p + 1) * nq2, tagL);
                        }
                    }
                } else {
#pragma unroll
                    for (int f = 0; f < 6; ++f) {
                        g[f] = qin[(p % GPF) * 6 + f];
                    }
                    if (p + GPF < NQ) {
#pragma unroll
                        for (int f = 0; f < 6; ++f) {
                            qin[(p % GPF) * 6 + f] = load_as<(NT & 2) != 0>(ge + f * S::nq3 + (p + GPF) * nq2, tagL);
                        }
                    } else if constexpr (PIPE) {
                        // the next element's row p % GPF: the slot just freed
                        // is the one that row is read from (rows 0 .. GPF - 1
                        // of the next element each take their own slot over
                        // the last GPF rows, whatever nq mod GPF is)
#pragma unroll
                        for (int f = 0; f < 6; ++f) {
                            qin[(p % GPF) * 6 + f] =
                                load_as<(NT & 2) != 0>(Qn + l + f * S::nq3 + (p % GPF) * nq2, tagL);
                        }
                    }
                }
                const int x = X(p, l_q, l_q2);
                const V qs = bS[x];
                const V qt = bC[x];
                // (rr, rs, rt) = G (qr, qs, qt), G symmetric: rr, rs, rt, ss, st, tt
                Acc<V> a0, a1, a2;
                a0.mac(qr[p], g[0]);
                a0.mac(qs, g[1]);
                a0.mac(qt, g[2]);
                a1.mac(qr[p], g[1]);
                a1.mac(qs, g[3]);
                a1.mac(qt, g[4]);
                a2.mac(qr[p], g[2]);
                a2.mac(qs, g[4]);
                a2.mac(qt, g[5]);
                rr[p] = a0.get();
                bS[x] = a1.get();
                bC[x] = a2.get();
                if constexpr (GPF > 0) {
                    // keep the rows' loads where they are written: GPF rows
                    // in flight, not all of G hoisted to the top of P5
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        if constexpr (PIPE) {
            // the lane's u column of the next element
            const int lu = l < S::nm2 ? l : 0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                xin[i] = load_as<(NT & 1) != 0>(un + i * S::nm2 + lu, tagL);
            }
        }
        __syncthreads();
        // ---- P6: transposed gradient ---------------------------------------
        V Ws[NQ], Wt[NQ];
        {
            V Wr[NQ];
            if (in_elem) {
                V y[NQ], z[NQ];
                // Wr: lane (q, r), col p;  Ws: lane (p, r), col q;  Wt: lane (p, q), col r
                if constexpr (NB == 3) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bS[X(l_q, q, l_q2)];      // rs, lane (p, r), col q
                        z[q] = bC[X(l_q, l_q2, q)];      // rt, lane (p, q), col r
                    }
                    contract_rows3<NQ, SB>(mat_d<7, EO, NQ>(D, dep), rr, Wr, y, Ws, z, Wt);
                } else {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        z[q] = bC[X(l_q, l_q2, q)];
                    }
                    contract_rows<NQ, SB>(mat_d<7, EO, NQ>(D, dep), z, Wt);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        y[q] = bS[X(l_q, q, l_q2)];
                    }
                    contract_rows<NQ, SB>(mat_d<8, EO, NQ>(D, dep), y, Ws);
                    contract_rows<NQ, SB>(mat_d<9, EO, NQ>(D, dep), rr, Wr);
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
                W[r] = sum3(Wt[r], bA[x], bB[x]);
            }
        }
    }

    // ---- P7: lane (p,q), contract r -> k ---------------------------------
    if (in_elem) {
        V Y[NM];
        contract_rows<NQ, SB>(mat_b<10, EO, NM, NQ>(B, dep), W, Y);
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            bA[X(l_q, l_q2, k)] = Y[k];
        }
    }
    __syncthreads();

    // ---- P8: lane (p,k), contract q -> j ---------------------------------
    if (in_elem && l < NQ * NM) {
        V y[NQ], Z[NM];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            y[q] = bA[X(l_m, q, l_m2)];
        }
        contract_rows<NQ, SB>(mat_b<11, EO, NM, NQ>(B, dep), y, Z);
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            bB[X(l_m, j, l_m2)] = Z[j];
        }
    }
    __syncthreads();

    // ---- P9: lane (j,k), contract p -> i, store --------------------------
    if constexpr (PIPE) {
        // Every lane stores, so that every pass issues the same NM stores
        // after its loads of the next element and the wait for those loads
        // at the top of the loop counts the stores instead of waiting for
        // them.  A lane outside the modal plane stores the column of plane
        // lane l mod nm^2 of its own box: the same bits as that lane, to the
        // same addresses.  Lanes past the element count hold the clamped
        // element's data and ve, and lanes past the slots slot 0's (the
        // caller points ve there), so theirs are that element's own values.
        const int lp = l < S::nm2 ? l : l % S::nm2;
        V z[NQ], o[NM];
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            z[p] = bB[X(p, lp / NM, lp % NM)];
        }
        contract_rows<NQ, SB>(mat_b<12, EO, NM, NQ>(B, dep), z, o);
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            store_v<(NT & 1) != 0>(ve + i * S::nm2 + lp, o[i]);
        }
    } else if (active && l < S::nm2) {
        V z[NQ], o[NM];
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            z[p] = bB[X(p, l_m, l_m2)];
        }
        contract_rows<NQ, SB>(mat_b<12, EO, NM, NQ>(B, dep), z, o);
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            store_v<(NT & 1) != 0>(ve + i * S::nm2 + l, o[i]);
        }
    }
}

template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB = 3, int GPF = 0, int FORM = BKX_FORM_F64,
          typename MT = double, int NT = 0, bool EO = false>
__device__ __forceinline__ void bk13t_body(const MT* __restrict__ B, const MT* __restrict__ D,
                                           const typename Form<FORM>::TS* __restrict__ Qe,
                                           const typename Form<FORM>::TS* __restrict__ ue,
                                           double* __restrict__ ve, bool active, int slot, int l, double* smem,
                                           int dep = 0)
{
    typename Form<FORM>::L xin[NM];
    typename Form<FORM>::L qin[bk13t_nqv(BK3, GPF, NQ)];
    bk13t_load<NM, NQ, BK3, DIV, GPF, FORM, NT, false>(Qe, ue, slot, l, xin, qin);
    bk13t_phases<NM, NQ, BK3, DIV, SB, NB, GPF, FORM, MT, NT, false, EO>(B, D, Qe, ve, active, slot, l, smem, dep, xin,
                                                                         qin, nullptr, nullptr);
}

}  // namespace bk13t
}  // namespace bkx
