// bk13t.hip -- BK1 (mass) and BK3 (Poisson) in FP64 on gfx950, transposing
// mapping.
//
// Operators as in bk13_f64.hip (openmp/bk1_sumfact.h:16-159,
// openmp/bk3_sumfact.h:17-187).  Here every contraction runs along the
// lane's register column, with the rows of B and D uniform across the
// wavefront (scalar operands), and the LDS only moves the element between
// lane planes.  A lane plane of an element is indexed (a, b) and the
// register column by c; the tensors keep their natural index order
// (p|i, q|j, r|k) in padded LDS boxes:
//
//   P1  lane (j,k)  col i: u from HBM;       a = B^T u along i -> col p  -> box A
//   P2  lane (p,k)  col j: box A;            b = B^T a along j -> col q  -> box B
//   P3  lane (p,q)  col k: box B;            U = B^T b along k -> col r
//       BK1: W = JxW U (JxW staged through box C by coalesced loads)
//       BK3: qt = D^T U along r;  U -> box A, qt -> box C
//   P4  lane (p,r)  col q: box A (U);        qs along q -> box B
//       lane (q,r)  col p: box A (U);        qr along p, kept in registers
//   P5  lane (q,r)  col p: qs (box B), qt (box C), G from HBM (a G(f, p, :, :)
//       plane is nq^2 contiguous values, one lane each): rr kept,
//       rs -> box B, rt -> box C (in place)
//   P6  (q,r): Wr = D rr along p -> box A;  (p,r): Ws = D rs along q;
//       (p,q): Wt = D rt along r;  Ws -> box B;  (p,q): W = Wt + A + B
//   P7  lane (p,q)  col r: Y = B W along r -> col k  -> box A
//   P8  lane (p,k)  col q: box A;  Z = B Y along q -> col j -> box B
//   P9  lane (j,k)  col p: box B;  v = B Z along p -> col i -> HBM
//
// LDS traffic per element is a few box writes and reads (about 16 nq^3
// accesses for BK3), where contracting through LDS reads nq values per
// output (about 60 nq^3).  Workgroups hold floor(256 / nq^2) elements, nq^2
// lanes each, and synchronise between phases.
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

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

template <int NM, int NQ, bool BK3, int DIV, bool SB>
__global__ __launch_bounds__((Bk13tShape<NM, NQ, DIV>::threads)) void bk13t_kernel(size_t nelmt, const double* __restrict__ B,
                                                    const double* __restrict__ D, const double* __restrict__ Q,
                                                    const double* __restrict__ u, double* __restrict__ v)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr int nq2 = S::nq2;
    constexpr int R = S::row;
    __shared__ double s_box[S::epb][3][S::box];

    const int slot = threadIdx.x / nq2;
    const int l = threadIdx.x - slot * nq2;
    const bool in_elem = slot < S::epb;
    const size_t e = size_t(blockIdx.x) * S::epb + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    const int sl = in_elem ? slot : 0;
    double* bA = s_box[sl][0];
    double* bB = s_box[sl][1];
    double* bC = s_box[sl][2];
    auto X = [](int a, int b, int c) { return (a * NQ + b) * R + c; };
    // lane (a, b) of an NA x NB plane, when l < NA * NB
    const int l_m = l / NM, l_m2 = l - (l / NM) * NM;    // (l / NM, l % NM)
    const int l_q = l / NQ, l_q2 = l - (l / NQ) * NQ;    // (l / NQ, l % NQ)

    // ---- P1: lane (j,k), contract i -> p ---------------------------------
    if (in_elem && l < S::nm2) {
        double x[NM], a[NQ];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            x[i] = active ? u[ee * S::nm3 + i * S::nm2 + l] : 0.0;
        }
        contract_cols<NQ, SB>(launder<1>(B), x, a);
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
                bC[x + x / NQ] = active ? Q[ee * S::nq3 + x] : 0.0;
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
        contract_cols<NQ, SB>(launder<2>(B), y, b);
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
        contract_cols<NQ, SB>(launder<3>(B), y, U);
        if constexpr (!BK3) {
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                W[r] = U[r] * bC[X(l_q, l_q2, r)];
            }
        } else {
            double qt[NQ];
            contract_cols<NQ, SB>(launder<4>(D), U, qt);
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
            contract_cols2<NQ, SB>(launder<5>(D), y, qs, z, qr);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                bB[X(l_q, q, l_q2)] = qs[q];
            }
        }
        __syncthreads();
        // ---- P5: fluxes in lane (q,r) col p ---------------------------------
        double rr[NQ];
        if (in_elem) {
            const double* ge = Q + ee * 6 * S::nq3 + l;   // G(e, f, p, q, r) at f = p = 0
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
            contract_rows3<NQ, SB>(launder<7>(D), rr, Wr, y, Ws, z, Wt);
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
        contract_rows<NQ, SB>(launder<10>(B), W, Y);
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
        contract_rows<NQ, SB>(launder<11>(B), y, Z);
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
        contract_rows<NQ, SB>(launder<12>(B), z, o);
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            v[e * S::nm3 + i * S::nm2 + l] = o[i];
        }
    }
}

}  // namespace

template <int NM, int NQ, bool BK3, int DIV, bool SB>
int launch_bk13t(size_t nelmt, const double* B, const double* D, const double* Q, const double* u, double* v,
                 hipStream_t s)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk13t_kernel<NM, NQ, BK3, DIV, SB>), dim3(grid), dim3(S::threads), 0, s, nelmt, B, D, Q, u, v);
    return launch_status();
}

template <int NM, bool BK3>
int launch_bk13t_variant(int div_variant, size_t nelmt, const double* B, const double* D, const double* Q,
                         const double* u, double* v, hipStream_t s)
{
    switch (div_variant) {
    case 0:
        return launch_bk13t<NM, NM + 1, BK3, 1, false>(nelmt, B, D, Q, u, v, s);
    case 1:
        return launch_bk13t<NM, NM + 1, BK3, 2, false>(nelmt, B, D, Q, u, v, s);
    case 2:
        return launch_bk13t<NM, NM + 1, BK3, 1, true>(nelmt, B, D, Q, u, v, s);
    default:
        return BKX_ERROR_ARG;
    }
}

int bk1t_apply_f64(int variant, int p, size_t nelmt, const double* B, const double* JxW, const double* u, double* v,
                   hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        return launch_bk13t_variant<decltype(P)::value + 1, false>(variant, nelmt, B, nullptr, JxW, u, v, s);
    });
}

int bk3t_apply_f64(int variant, int p, size_t nelmt, const double* B, const double* D, const double* G,
                   const double* u, double* v, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        return launch_bk13t_variant<decltype(P)::value + 1, true>(variant, nelmt, B, D, G, u, v, s);
    });
}

}  // namespace bkx
