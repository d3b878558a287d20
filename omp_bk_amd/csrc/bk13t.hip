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
#include "bk13t_body.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

using bk13t::Bk13tShape;
using bk13t::bk13t_body;

// Occupancy target (wavefronts per SIMD) for GPF rows of G held in
// registers: none without prefetch, 4 for up to three rows, 3 for up to six,
// else 2 (a whole element's G, 6 nq values per lane, is loaded with u and
// in flight through the first four phases).
constexpr int bk13t_waves(int gpf) { return gpf == 0 ? 1 : (gpf <= 3 ? 4 : (gpf <= 6 ? 3 : 2)); }

// EO: B and D point at their even-odd forms (EoMat in bkx_eo.hpp).
template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF, int NT, bool EO>
__global__ __launch_bounds__((Bk13tShape<NM, NQ, DIV>::threads))
__attribute__((amdgpu_waves_per_eu(bk13t_waves(GPF)))) void bk13t_kernel(
    size_t nelmt, const double* __restrict__ B, const double* __restrict__ D, const double* __restrict__ Q,
    const double* __restrict__ u, double* __restrict__ v)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    __shared__ double s_box[S::epb * NB * S::box];
    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const size_t e = size_t(blockIdx.x) * S::epb + slot;
    const bool active = slot < S::epb && e < nelmt;
    const size_t ee = active ? e : 0;
    constexpr size_t qper = BK3 ? 6 * S::nq3 : S::nq3;
    bk13t_body<NM, NQ, BK3, DIV, SB, NB, GPF, BKX_FORM_F64, double, NT, EO>(B, D, Q + ee * qper, u + ee * S::nm3,
                                                                           v + ee * S::nm3, active, slot, l, s_box);
}

// Persistent BK3 (PIPE): workgroup w processes the element groups
// w, w + W, w + 2W, ... (W = the number of resident workgroups), each
// group's first loads issued during the previous group's flux phase.  The
// loads past the last group re-read the current one (never used), so that
// every path issues the same memory operations.  Every workgroup leaves the
// loop after its last group; the barriers inside the phases are reached by
// all of its threads on every pass (the group loop is uniform).
template <int NM, int NQ, int DIV, int GPF, int NT, int MINW>
__global__ __launch_bounds__((Bk13tShape<NM, NQ, DIV>::threads))
__attribute__((amdgpu_waves_per_eu(MINW))) void bk3p_kernel(
    size_t nelmt, const double* __restrict__ B, const double* __restrict__ D, const double* __restrict__ Q,
    const double* __restrict__ u, double* __restrict__ v)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    __shared__ double s_box[S::epb * 2 * S::box];
    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const size_t ngroups = (nelmt + S::epb - 1) / S::epb;
    size_t g = blockIdx.x;
    if (g >= ngroups) {
        return;
    }
    constexpr size_t qper = 6 * S::nq3;
    // element of this thread in group grp: threads past the slots take slot
    // 0's (the box they read), slots past the count element 0
    auto elem = [&](size_t grp) {
        const size_t e = grp * S::epb + (slot < S::epb ? slot : 0);
        return e < nelmt ? e : size_t(0);
    };
    size_t ee = elem(g);
    double xin[NM];
    double qin[bk13t::bk13t_nqv(true, GPF, NQ)];
    bk13t::bk13t_load<NM, NQ, true, DIV, GPF, BKX_FORM_F64, NT, true>(Q + ee * qper, u + ee * S::nm3, slot, l, xin, qin);
    // The first element's loads complete before the loop (once per wave):
    // the path into the loop from here then has nothing in flight, and the
    // compiler's waits at the top of the loop are those of the back edge,
    // which counts the pass's stores (always NM, issued after its loads of
    // the next element) instead of waiting for them.
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0)
    for (;;) {
        const size_t e = g * S::epb + slot;
        const bool active = slot < S::epb && e < nelmt;
        const bool has_next = g + gridDim.x < ngroups;
        const size_t gn = has_next ? g + gridDim.x : g;
        const size_t en = elem(gn);
        bk13t::bk13t_phases<NM, NQ, true, DIV, false, 2, GPF, BKX_FORM_F64, double, NT, true>(
            B, D, Q + ee * qper, v + ee * S::nm3, active, slot, l, s_box, int(g), xin, qin, Q + en * qper,
            u + en * S::nm3);
        if (!has_next) {
            break;
        }
        g = gn;
        ee = en;
    }
}

}  // namespace

// Launches the persistent BK3 kernel with at most as many workgroups as are
// resident at once.
template <int NM, int NQ, int DIV, int GPF, int NT, int MINW = 3>
int launch_bk3p(size_t nelmt, const double* B, const double* D, const double* Q, const double* u, double* v,
                hipStream_t s)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    const unsigned resident = resident_blocks<bk3p_kernel<NM, NQ, DIV, GPF, NT, MINW>>(S::threads);
    if (resident == 0) {
        return BKX_ERROR_LAUNCH;
    }
    const size_t ngroups = (nelmt + S::epb - 1) / S::epb;
    const size_t grid = ngroups < resident ? ngroups : resident;
    hipLaunchKernelGGL((bk3p_kernel<NM, NQ, DIV, GPF, NT, MINW>), dim3(unsigned(grid)), dim3(S::threads), 0, s, nelmt, B, D,
                       Q, u, v);
    return launch_status();
}

template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF = 0, int NT = 0, bool EO = false>
int launch_bk13t(size_t nelmt, const double* B, const double* D, const double* Q, const double* u, double* v,
                 hipStream_t s)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk13t_kernel<NM, NQ, BK3, DIV, SB, NB, GPF, NT, EO>), dim3(grid), dim3(S::threads), 0, s, nelmt, B, D, Q, u, v);
    return launch_status();
}

// BK3 with the bases in even-odd form (P: B's form, then D's): the launch
// shapes of the default variants 13, 14, 16, 17 and 18; any other variant
// runs as 16.
template <int NM>
int launch_bk3_eo(int variant, size_t nelmt, const double* P, const double* G, const double* u, double* v,
                  hipStream_t s)
{
    constexpr int NQ = NM + 1;
    const double* B = P;
    const double* D = P + EoMat<NM, NQ, 1>::size;
    switch (variant) {
    case 13:
        return launch_bk13t<NM, NQ, true, 1, false, 2, (NQ < 6 ? NQ : 6), 3, true>(nelmt, B, D, G, u, v, s);
    case 14:
        return launch_bk13t<NM, NQ, true, 1, false, 2, NQ, 3, true>(nelmt, B, D, G, u, v, s);
    case 17:
        return launch_bk13t<NM, NQ, true, 1, false, 2, 2, 1, true>(nelmt, B, D, G, u, v, s);
    case 18:
        return launch_bk13t<NM, NQ, true, 2, false, 2, 3, 1, true>(nelmt, B, D, G, u, v, s);
    default:
        return launch_bk13t<NM, NQ, true, 1, false, 2, 3, 1, true>(nelmt, B, D, G, u, v, s);
    }
}

template <int NM, bool BK3>
int launch_bk13t_variant(int div_variant, size_t nelmt, const double* B, const double* D, const double* Q,
                         const double* u, double* v, hipStream_t s)
{
    // variant: 0 all elements per workgroup, 1 half of them, 2 all with a
    // scheduling barrier per row (three LDS boxes each); 4 all, 5 half (two
    // LDS boxes); BK3 only: 6 all, 7 half (two boxes, three rows of G loaded
    // with u), 8 all (two boxes, two rows); BK1 6, 7, 8 are 4, 5, 4.
    // Variant 3 is the LDS-contraction mapping of bk13_f64.hip.  9-12 are 6,
    // 8, 7 and 4 with non-temporal loads of u and JxW/G and stores of v
    // (each G plane is read by one instruction, so the streaming hint loses
    // no cache reuse).  BK3 13-15: non-temporal, min(6, nq) rows of G or all
    // nq rows loaded with u (all elements or half of them per workgroup);
    // BK1 13-15 are 12.  16-19 are 9-12 with G (JxW) read by ordinary loads
    // and only u and v streamed.  BK3 20-23: the persistent pipelined kernel
    // (three or two rows of G in the ring, all or half the elements per
    // workgroup; 23 without an occupancy target); BK1 20-23 are 19.
    switch (div_variant) {
    case 0:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 3>(nelmt, B, D, Q, u, v, s);
    case 1:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 3>(nelmt, B, D, Q, u, v, s);
    case 2:
        return launch_bk13t<NM, NM + 1, BK3, 1, true, 3>(nelmt, B, D, Q, u, v, s);
    case 4:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2>(nelmt, B, D, Q, u, v, s);
    case 5:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 2>(nelmt, B, D, Q, u, v, s);
    case 6:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 3 : 0>(nelmt, B, D, Q, u, v, s);
    case 7:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 2, BK3 ? 3 : 0>(nelmt, B, D, Q, u, v, s);
    case 8:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 2 : 0>(nelmt, B, D, Q, u, v, s);
    case 9:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 3 : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 10:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 2 : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 11:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 2, BK3 ? 3 : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 12:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, 0, 3>(nelmt, B, D, Q, u, v, s);
    case 13:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? (NM + 1 < 6 ? NM + 1 : 6) : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 14:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? NM + 1 : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 15:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 2, BK3 ? NM + 1 : 0, 3>(nelmt, B, D, Q, u, v, s);
    case 16:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 3 : 0, 1>(nelmt, B, D, Q, u, v, s);
    case 17:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, BK3 ? 2 : 0, 1>(nelmt, B, D, Q, u, v, s);
    case 18:
        return launch_bk13t<NM, NM + 1, BK3, 2, false, 2, BK3 ? 3 : 0, 1>(nelmt, B, D, Q, u, v, s);
    case 19:
        return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, 0, 1>(nelmt, B, D, Q, u, v, s);
    case 20:
    case 21:
    case 22:
    case 23:
        if constexpr (BK3) {
            switch (div_variant) {
            case 20:
                return launch_bk3p<NM, NM + 1, 1, 3, 1>(nelmt, B, D, Q, u, v, s);
            case 21:
                return launch_bk3p<NM, NM + 1, 1, 2, 1>(nelmt, B, D, Q, u, v, s);
            case 22:
                return launch_bk3p<NM, NM + 1, 2, 3, 1>(nelmt, B, D, Q, u, v, s);
            default:
                return launch_bk3p<NM, NM + 1, 1, 3, 1, 1>(nelmt, B, D, Q, u, v, s);
            }
        } else {
            return launch_bk13t<NM, NM + 1, BK3, 1, false, 2, 0, 1>(nelmt, B, D, Q, u, v, s);
        }
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

int bk3t_apply_eo_f64(int variant, int p, size_t nelmt, const double* P, const double* G, const double* u, double* v,
                      hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto O) { return launch_bk3_eo<decltype(O)::value + 1>(variant, nelmt, P, G, u, v, s); });
}

}  // namespace bkx
