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

template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF>
__global__ __launch_bounds__((Bk13tShape<NM, NQ, DIV>::threads))
__attribute__((amdgpu_waves_per_eu(GPF > 0 ? 4 : 1))) void bk13t_kernel(
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
    bk13t_body<NM, NQ, BK3, DIV, SB, NB, GPF>(B, D, Q + ee * qper, u + ee * S::nm3, v + ee * S::nm3, active, slot, l, s_box);
}

}  // namespace

template <int NM, int NQ, bool BK3, int DIV, bool SB, int NB, int GPF = 0>
int launch_bk13t(size_t nelmt, const double* B, const double* D, const double* Q, const double* u, double* v,
                 hipStream_t s)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk13t_kernel<NM, NQ, BK3, DIV, SB, NB, GPF>), dim3(grid), dim3(S::threads), 0, s, nelmt, B, D, Q, u, v);
    return launch_status();
}

template <int NM, bool BK3>
int launch_bk13t_variant(int div_variant, size_t nelmt, const double* B, const double* D, const double* Q,
                         const double* u, double* v, hipStream_t s)
{
    // variant: 0 all elements per workgroup, 1 half of them, 2 all with a
    // scheduling barrier per row (three LDS boxes each); 4 all, 5 half (two
    // LDS boxes); BK3 only: 6 all, 7 half (two boxes, three rows of G loaded
    // with u), 8 all (two boxes, two rows); BK1 6, 7, 8 are 4, 5, 4.
    // Variant 3 is the LDS-contraction mapping of bk13_f64.hip.
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
