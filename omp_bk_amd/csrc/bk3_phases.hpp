// bk3_phases.hpp -- the phases of the transposing BK3 mapping (bk13t.hip) as
// device functions, for kernels that chain several gradient / divergence
// passes per element (the adjoint in bk3_adj.hip).
//
// An element owns nq^2 consecutive lanes of the workgroup (lane l) and three
// padded LDS boxes A, B, C of nq^3 values indexed X(p|i, q|j, r|k).  Between
// calls the workgroup must be synchronised by the caller where noted.
//
// The value type V is double (FP64 arithmetic, B and D read as
// wavefront-uniform scalar operands) or ff (float-float arithmetic: every
// contraction a Dot2 compensated sum, bkx_ff.hpp, with B and D split into
// float-float in a per-workgroup LDS copy).
//
//   grad(x):  modal values x[i] of lane (j,k) (l < nm^2), column i
//             -> qr[p] in lane (q,r), column p; qs(p,q,r) in box B and
//                qt(p,q,r) in box C (the reference gradient at the Gauss
//                points, openmp/bk3_sumfact.h:63-119)
//   div(rr):  rr[p] in lane (q,r) column p, rs in box B, rt in box C
//             -> o[i] in lane (j,k) (l < nm^2), column i:
//                o = B^T-interpolated transposed gradient
//                (openmp/bk3_sumfact.h:126-177)
#pragma once

#include "bkx_device.hpp"
#include "bkx_eo.hpp"
#include "bkx_ff.hpp"

namespace bkx {
namespace bk3p {

template <bool SB>
__device__ __forceinline__ void row_fence()
{
    if constexpr (SB) {
        __builtin_amdgcn_sched_barrier(0);
    }
}

// out[c] = sum_m M[m * LD + c] in[m]
template <int LD, int NIN, int NOUT, typename P>
__device__ __forceinline__ void cols(P M, const double (&in)[NIN], double (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out[c] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            out[c] += M[m * LD + c] * in[m];
        }
    }
}

template <int LD, int NIN, int NOUT, typename P>
__device__ __forceinline__ void cols2(P M, const double (&in0)[NIN], double (&out0)[NOUT],
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
            out0[c] += M[m * LD + c] * in0[m];
            out1[c] += M[m * LD + c] * in1[m];
        }
    }
}

// out[c] = sum_m M[c * LD + m] in[m]
template <int LD, int NIN, int NOUT, typename P>
__device__ __forceinline__ void rows(P M, const double (&in)[NIN], double (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t += M[c * LD + m] * in[m];
        }
        out[c] = t;
    }
}

template <int LD, int NIN, int NOUT, typename P>
__device__ __forceinline__ void rows3(P M, const double (&in0)[NIN], double (&out0)[NOUT],
                                      const double (&in1)[NIN], double (&out1)[NOUT],
                                      const double (&in2)[NIN], double (&out2)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t0 += M[c * LD + m] * in0[m];
            t1 += M[c * LD + m] * in1[m];
            t2 += M[c * LD + m] * in2[m];
        }
        out0[c] = t0;
        out1[c] = t1;
        out2[c] = t2;
    }
}

// The float-float contractions: M is the float-float LDS copy of B or D,
// the inputs are exact floats (loaded values) or float-float values, and
// every output is a Dot2 sum over m ascending.
template <int LD, int NIN, int NOUT, typename X>
__device__ __forceinline__ void cols(const ff* M, const X (&in)[NIN], ff (&out)[NOUT])
{
    dot2 acc[NOUT];
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) {
            acc[c].add(M[m * LD + c], in[m]);
        }
    }
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        out[c] = acc[c].value();
    }
}

template <int LD, int NIN, int NOUT>
__device__ __forceinline__ void cols2(const ff* M, const ff (&in0)[NIN], ff (&out0)[NOUT], const ff (&in1)[NIN],
                                      ff (&out1)[NOUT])
{
    cols<LD>(M, in0, out0);
    cols<LD>(M, in1, out1);
}

template <int LD, int NIN, int NOUT>
__device__ __forceinline__ void rows(const ff* M, const ff (&in)[NIN], ff (&out)[NOUT])
{
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
        dot2 t;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            t.add(M[c * LD + m], in[m]);
        }
        out[c] = t.value();
    }
}

template <int LD, int NIN, int NOUT>
__device__ __forceinline__ void rows3(const ff* M, const ff (&in0)[NIN], ff (&out0)[NOUT], const ff (&in1)[NIN],
                                      ff (&out1)[NOUT], const ff (&in2)[NIN], ff (&out2)[NOUT])
{
    rows<LD>(M, in0, out0);
    rows<LD>(M, in1, out1);
    rows<LD>(M, in2, out2);
}

// The same contractions with the matrix in even-odd form (bkx_eo.hpp).
template <int LD, int NIN, int NOUT, int R, int C, int S>
__device__ __forceinline__ void cols(EoMat<R, C, S> M, const double (&in)[NIN], double (&out)[NOUT])
{
    static_assert(NIN == R && NOUT == C, "the even-odd form contracts a whole matrix");
    using E = EoMat<R, C, S>;
    double e[E::hr > 0 ? E::hr : 1], o[E::hr > 0 ? E::hr : 1];
#pragma unroll
    for (int m = 0; m < E::hr; ++m) {
        e[m] = in[m] + in[R - 1 - m];
        o[m] = in[m] - in[R - 1 - m];
    }
#pragma unroll
    for (int c = 0; c < E::cc; ++c) {
        const bool mid = (C % 2 == 1) && c == C / 2;
        double a = 0.0, b = 0.0;
        if (!mid || S == 1) {
#pragma unroll
            for (int m = 0; m < E::hr; ++m) {
                a += M.p[E::cols_p + m * E::cc + c] * e[m];
            }
            if constexpr (R % 2 == 1) {
                a += M.p[E::cols_mid + c] * in[R / 2];
            }
        }
        if (!mid || S == -1) {
#pragma unroll
            for (int m = 0; m < E::hr; ++m) {
                b += M.p[E::cols_q + m * E::cc + c] * o[m];
            }
        }
        if (mid) {
            out[c] = S == 1 ? a : b;
        } else {
            out[c] = a + b;
            out[C - 1 - c] = S == 1 ? a - b : b - a;
        }
    }
}

template <int LD, int NIN, int NOUT, int R, int C, int S>
__device__ __forceinline__ void cols2(EoMat<R, C, S> M, const double (&in0)[NIN], double (&out0)[NOUT],
                                      const double (&in1)[NIN], double (&out1)[NOUT])
{
    cols<LD>(M, in0, out0);
    cols<LD>(M, in1, out1);
}

template <int LD, int NIN, int NOUT, int R, int C, int S>
__device__ __forceinline__ void rows(EoMat<R, C, S> M, const double (&in)[NIN], double (&out)[NOUT])
{
    static_assert(NIN == C && NOUT == R, "the even-odd form contracts a whole matrix");
    using E = EoMat<R, C, S>;
    double e[E::hc > 0 ? E::hc : 1], o[E::hc > 0 ? E::hc : 1];
#pragma unroll
    for (int m = 0; m < E::hc; ++m) {
        e[m] = in[m] + in[C - 1 - m];
        o[m] = in[m] - in[C - 1 - m];
    }
#pragma unroll
    for (int c = 0; c < E::cr; ++c) {
        const bool mid = (R % 2 == 1) && c == R / 2;
        double a = 0.0, b = 0.0;
        if (!mid || S == 1) {
#pragma unroll
            for (int m = 0; m < E::hc; ++m) {
                a += M.p[E::rows_p + m * E::cr + c] * e[m];
            }
            if constexpr (C % 2 == 1) {
                a += M.p[E::rows_mid + c] * in[C / 2];
            }
        }
        if (!mid || S == -1) {
#pragma unroll
            for (int m = 0; m < E::hc; ++m) {
                b += M.p[E::rows_q + m * E::cr + c] * o[m];
            }
        }
        if (mid) {
            out[c] = S == 1 ? a : b;
        } else {
            out[c] = a + b;
            out[R - 1 - c] = S == 1 ? a - b : b - a;
        }
    }
}

template <int LD, int NIN, int NOUT, int R, int C, int S>
__device__ __forceinline__ void rows3(EoMat<R, C, S> M, const double (&in0)[NIN], double (&out0)[NOUT],
                                      const double (&in1)[NIN], double (&out1)[NOUT],
                                      const double (&in2)[NIN], double (&out2)[NOUT])
{
    rows<LD>(M, in0, out0);
    rows<LD>(M, in1, out1);
    rows<LD>(M, in2, out2);
}

// the operand of a contraction with B (EO: the even-odd form of the
// nm x nq B at that pointer) or D (of the nq x nq D)
template <int TAG, bool EO, int NM, int NQ>
__device__ __forceinline__ auto op_b(const double* p)
{
    if constexpr (EO) {
        return EoMat<NM, NQ, 1>{launder<TAG>(p)};
    } else {
        return launder<TAG>(p);
    }
}
template <int TAG, bool EO, int NQ>
__device__ __forceinline__ auto op_d(const double* p)
{
    if constexpr (EO) {
        return EoMat<NQ, NQ, -1>{launder<TAG>(p)};
    } else {
        return launder<TAG>(p);
    }
}
// the float-float LDS copies (no even-odd form)
template <int TAG, bool EO, int NM, int NQ>
__device__ __forceinline__ const ff* op_b(const ff* p)
{
    static_assert(!EO, "the float-float contractions use B directly");
    return p;
}
template <int TAG, bool EO, int NQ>
__device__ __forceinline__ const ff* op_d(const ff* p)
{
    static_assert(!EO, "the float-float contractions use D directly");
    return p;
}

// B (nm x nq) and D (nq x nq) split into float-float into the LDS arrays sB,
// sD by the workgroup's threads; the caller synchronises afterwards.
template <int NM, int NQ>
__device__ __forceinline__ void split_bases(const double* __restrict__ B, const double* __restrict__ D, ff* sB, ff* sD)
{
    for (int x = threadIdx.x; x < NM * NQ; x += blockDim.x) {
        sB[x] = ff_split(B[x]);
    }
    for (int x = threadIdx.x; x < NQ * NQ; x += blockDim.x) {
        sD[x] = ff_split(D[x]);
    }
}

// a + b + c, left to right (FP64) or as one compensated sum (float-float)
__device__ __forceinline__ double add3(double a, double b, double c) { return a + b + c; }
__device__ __forceinline__ ff add3(ff a, ff b, ff c)
{
    dot2 d;
    d.add(a);
    d.add(b);
    d.add(c);
    return d.value();
}

template <int NM, int NQ, typename V = double>
struct Elem {
    static constexpr int nm2 = NM * NM;
    static constexpr int nm3 = nm2 * NM;
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int row = NQ | 1;            // odd row stride in 8-byte words (bank spread)
    static constexpr int box = NQ * NQ * row;

    V* A;
    V* B;
    V* C;
    int l;            // lane within the element's plane, 0 .. nq^2 - 1
    bool in_elem;     // lane belongs to an element slot of the workgroup
    int lm, lm2;      // (l / NM, l % NM)
    int lq, lq2;      // (l / NQ, l % NQ)

    __device__ static int X(int a, int b, int c) { return (a * NQ + b) * row + c; }

    __device__ void init(V* boxes, int lane, bool in)
    {
        A = boxes;
        B = boxes + box;
        C = boxes + 2 * box;
        l = lane;
        in_elem = in;
        lm = l / NM;
        lm2 = l - lm * NM;
        lq = l / NQ;
        lq2 = l - lq * NQ;
    }

    // modal x (lane (j,k), l < nm^2) -> qr (lane (q,r)), qs -> B, qt -> C.
    // Ends with the workgroup synchronised.  MT: double (device rows of B
    // and D) or ff (their LDS copy); XT: V or, for ff, the exact float
    // values as loaded.
    template <int TAG, bool EO = false, typename MT, typename XT>
    __device__ void grad(const MT* Bm, const MT* Dm, const XT (&x)[NM], V (&qr)[NQ])
    {
        if (in_elem && l < nm2) {
            V a[NQ];
            cols<NQ>(op_b<TAG + 1, EO, NM, NQ>(Bm), x, a);
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                A[X(p, lm, lm2)] = a[p];
            }
        }
        __syncthreads();
        if (in_elem && l < NQ * NM) {
            V y[NM], b[NQ];
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                y[j] = A[X(lm, j, lm2)];
            }
            cols<NQ>(op_b<TAG + 2, EO, NM, NQ>(Bm), y, b);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                B[X(lm, q, lm2)] = b[q];
            }
        }
        __syncthreads();
        if (in_elem) {
            V y[NM], U[NQ], qt[NQ];
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                y[k] = B[X(lq, lq2, k)];
            }
            cols<NQ>(op_b<TAG + 3, EO, NM, NQ>(Bm), y, U);
            cols<NQ>(op_d<TAG + 4, EO, NQ>(Dm), U, qt);
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                A[X(lq, lq2, r)] = U[r];
                C[X(lq, lq2, r)] = qt[r];
            }
        }
        __syncthreads();
        if (in_elem) {
            V y[NQ], z[NQ], qs[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                y[q] = A[X(lq, q, lq2)];
                z[q] = A[X(q, lq, lq2)];
            }
            cols2<NQ>(op_d<TAG + 5, EO, NQ>(Dm), y, qs, z, qr);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                B[X(lq, q, lq2)] = qs[q];
            }
        }
        __syncthreads();
    }

    // rr (lane (q,r)), rs in B, rt in C -> o (lane (j,k), l < nm^2).
    // Starts after a workgroup synchronisation (the caller's writes of B, C);
    // ends with the workgroup synchronised.
    template <int TAG, bool EO = false, typename MT>
    __device__ void div(const MT* Bm, const MT* Dm, const V (&rr)[NQ], V (&o)[NM])
    {
        V Ws[NQ], Wt[NQ];
        if (in_elem) {
            V Wr[NQ], y[NQ], z[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                y[q] = B[X(lq, q, lq2)];
                z[q] = C[X(lq, lq2, q)];
            }
            rows3<NQ>(op_d<TAG + 1, EO, NQ>(Dm), rr, Wr, y, Ws, z, Wt);
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                A[X(p, lq, lq2)] = Wr[p];
            }
        }
        __syncthreads();
        if (in_elem) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                B[X(lq, q, lq2)] = Ws[q];
            }
        }
        __syncthreads();
        if (in_elem) {
            V W[NQ], Y[NM];
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                const int x = X(lq, lq2, r);
                W[r] = add3(Wt[r], A[x], B[x]);
            }
            rows<NQ>(op_b<TAG + 2, EO, NM, NQ>(Bm), W, Y);
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                A[X(lq, lq2, k)] = Y[k];
            }
        }
        __syncthreads();
        if (in_elem && l < NQ * NM) {
            V y[NQ], Z[NM];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                y[q] = A[X(lm, q, lm2)];
            }
            rows<NQ>(op_b<TAG + 3, EO, NM, NQ>(Bm), y, Z);
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                B[X(lm, j, lm2)] = Z[j];
            }
        }
        __syncthreads();
        if (in_elem && l < nm2) {
            V z[NQ];
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                z[p] = B[X(p, lm, lm2)];
            }
            rows<NQ>(op_b<TAG + 4, EO, NM, NQ>(Bm), z, o);
        }
        __syncthreads();
    }
};

}  // namespace bk3p
}  // namespace bkx
