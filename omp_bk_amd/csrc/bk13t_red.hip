// bk13t_red.hip -- BK1 (mass) and BK3 (Poisson) in the reduced-storage and
// float-float forms on gfx950, with the transposing mapping of bk13t.hip:
//
//   F32W  u and JxW/G stored in FP32, FP64 arithmetic, v FP64
//   F32C  u and JxW/G stored in FP32 (exact floats), float-float arithmetic
//         (every contraction a Dot2 compensated sum), v FP64
//   FF    u and JxW/G stored in FP64 and split into float-float when loaded,
//         float-float arithmetic, v FP64: the reference's BK1_ff kernels
//         (openmp/BK1_ff.cpp:73-723, float_float.h:73-134), here for BK3 too
//
// The operators are those of openmp/bk1_sumfact.h:16-159 and
// openmp/bk3_sumfact.h:17-187.  The float-float forms split B and D into
// float-float once per workgroup (an LDS copy, read as wavefront-uniform
// broadcasts).  HBM bytes per element: F32W and F32C read 4 (n^3 + q^3)
// (BK1) or 4 (n^3 + 6 q^3) (BK3) and write 8 n^3; FF moves what FP64 does.
#include <type_traits>

#include "bk13t_body.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

using bk13t::Bk13tShape;

// EO (FP64-arithmetic forms): B and D point at their even-odd forms.
template <int NM, int NQ, bool BK3, int DIV, int NB, int GPF, int FORM, int NT, bool EO>
__global__ __launch_bounds__((Bk13tShape<NM, NQ, DIV>::threads))
__attribute__((amdgpu_waves_per_eu(GPF > 0 ? 4 : 1))) void bk13t_red_kernel(
    size_t nelmt, const double* __restrict__ B, const double* __restrict__ D,
    const typename bk13t::Form<FORM>::TS* __restrict__ Q, const typename bk13t::Form<FORM>::TS* __restrict__ u,
    double* __restrict__ v)
{
    using S = Bk13tShape<NM, NQ, DIV>;
    constexpr bool kFF = std::is_same<typename bk13t::Form<FORM>::V, ff>::value;
    __shared__ double s_box[S::epb * NB * S::box];
    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const size_t e = size_t(blockIdx.x) * S::epb + slot;
    const bool active = slot < S::epb && e < nelmt;
    const size_t ee = active ? e : 0;
    constexpr size_t qper = BK3 ? 6 * S::nq3 : S::nq3;
    if constexpr (kFF) {
        __shared__ ff s_mat[NM * NQ + (BK3 ? NQ * NQ : 1)];
        bk13t::split_bases<NM, NQ, BK3>(B, D, s_mat, s_mat + NM * NQ, threadIdx.x, S::threads);
        __syncthreads();
        // one row of the LDS matrix live at a time (row_fence): otherwise a
        // contraction's broadcast reads are all hoisted, two registers each
        bk13t::bk13t_body<NM, NQ, BK3, DIV, true, NB, GPF, FORM, ff>(
            s_mat, s_mat + NM * NQ, Q + ee * qper, u + ee * S::nm3, v + ee * S::nm3, active, slot, l, s_box);
    } else {
        bk13t::bk13t_body<NM, NQ, BK3, DIV, false, NB, GPF, FORM, double, NT, EO>(
            B, D, Q + ee * qper, u + ee * S::nm3, v + ee * S::nm3, active, slot, l, s_box);
    }
}

template <int NM, bool BK3, int DIV, int NB, int GPF, int FORM, int NT = 0, bool EO = false>
int launch_red(size_t nelmt, const double* B, const double* D, const void* Q, const void* u, double* v,
               hipStream_t s)
{
    using S = Bk13tShape<NM, NM + 1, DIV>;
    using TS = typename bk13t::Form<FORM>::TS;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk13t_red_kernel<NM, NM + 1, BK3, DIV, NB, GPF, FORM, NT, EO>), dim3(grid), dim3(S::threads), 0, s,
                       nelmt, B, D, static_cast<const TS*>(Q), static_cast<const TS*>(u), v);
    return launch_status();
}

// Variants per form.  F32W: 0 all elements per workgroup, 1 half (two LDS
// boxes, BK3 with two rows of G loaded ahead), 2 all (two boxes, BK3 one row
// ahead in P5), 3 all (three boxes); 4-6 are 2, 3 and 0 with non-temporal
// loads of u and stores of v, 7 is 2 with JxW / G streamed as well, 8 and 9
// are 6 with three rows of G ahead (all elements or half of them), 10 and 11
// with five and all nq rows ahead.  F32C
// and FF (bound by the vector ALU): 0 all, 1 half (two boxes), 2 all (three
// boxes).
template <int FORM>
constexpr int red_variants()
{
    return FORM == BKX_FORM_F32W ? 12 : 3;
}

template <int NM, bool BK3, int FORM>
int launch_red_variant(int variant, size_t nelmt, const double* B, const double* D, const void* Q, const void* u,
                       double* v, hipStream_t s)
{
    if constexpr (FORM == BKX_FORM_F32W) {
        switch (variant) {
        case 0:
            return launch_red<NM, BK3, 1, 2, BK3 ? 2 : 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 1:
            return launch_red<NM, BK3, 2, 2, BK3 ? 2 : 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 2:
            return launch_red<NM, BK3, 1, 2, 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 3:
            return launch_red<NM, BK3, 1, 3, 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 4:
            return launch_red<NM, BK3, 1, 2, 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 5:
            return launch_red<NM, BK3, 1, 3, 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 6:
            return launch_red<NM, BK3, 1, 2, BK3 ? 2 : 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 7:
            return launch_red<NM, BK3, 1, 2, 0, FORM, 3>(nelmt, B, D, Q, u, v, s);
        case 8:
            return launch_red<NM, BK3, 1, 2, BK3 ? 3 : 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 9:
            return launch_red<NM, BK3, 2, 2, BK3 ? 3 : 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 10:
            return launch_red<NM, BK3, 1, 2, BK3 ? (NM + 1 < 5 ? NM + 1 : 5) : 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        case 11:
            return launch_red<NM, BK3, 1, 2, BK3 ? NM + 1 : 0, FORM, 1>(nelmt, B, D, Q, u, v, s);
        default:
            return BKX_ERROR_ARG;
        }
    } else {
        switch (variant) {
        case 0:
            return launch_red<NM, BK3, 1, 2, 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 1:
            return launch_red<NM, BK3, 2, 2, 0, FORM>(nelmt, B, D, Q, u, v, s);
        case 2:
            return launch_red<NM, BK3, 1, 3, 0, FORM>(nelmt, B, D, Q, u, v, s);
        default:
            return BKX_ERROR_ARG;
        }
    }
}

template <bool BK3>
int apply_red(int form, int variant, int p, size_t nelmt, const double* B, const double* D, const void* Q,
              const void* u, double* v, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        constexpr int NM = decltype(P)::value + 1;
        switch (form) {
        case BKX_FORM_F32W:
            return launch_red_variant<NM, BK3, BKX_FORM_F32W>(variant, nelmt, B, D, Q, u, v, s);
        case BKX_FORM_F32C:
            return launch_red_variant<NM, BK3, BKX_FORM_F32C>(variant, nelmt, B, D, Q, u, v, s);
        case BKX_FORM_FF:
            return launch_red_variant<NM, BK3, BKX_FORM_FF>(variant, nelmt, B, D, Q, u, v, s);
        default:
            return int(BKX_ERROR_ARG);
        }
    });
}

}  // namespace

int bk13_red_variants(int form)
{
    switch (form) {
    case BKX_FORM_F32W:
        return red_variants<BKX_FORM_F32W>();
    case BKX_FORM_F32C:
        return red_variants<BKX_FORM_F32C>();
    case BKX_FORM_FF:
        return red_variants<BKX_FORM_FF>();
    default:
        return -1;
    }
}

// BK3 f32w with the bases in even-odd form (P: B's form, then D's): the
// launch shapes of the variants 4, 6 and 8-11; any other variant runs as 8.
int bk3_apply_f32w_eo(int variant, int p, size_t nelmt, const double* P, const float* G, const float* u, double* v,
                      hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto O) {
        constexpr int NM = decltype(O)::value + 1;
        constexpr int F = BKX_FORM_F32W;
        const double* B = P;
        const double* D = P + EoMat<NM, NM + 1, 1>::size;
        switch (variant) {
        case 4:
            return launch_red<NM, true, 1, 2, 0, F, 1, true>(nelmt, B, D, G, u, v, s);
        case 6:
            return launch_red<NM, true, 1, 2, 2, F, 1, true>(nelmt, B, D, G, u, v, s);
        case 9:
            return launch_red<NM, true, 2, 2, 3, F, 1, true>(nelmt, B, D, G, u, v, s);
        case 10:
            return launch_red<NM, true, 1, 2, (NM + 1 < 5 ? NM + 1 : 5), F, 1, true>(nelmt, B, D, G, u, v, s);
        case 11:
            return launch_red<NM, true, 1, 2, NM + 1, F, 1, true>(nelmt, B, D, G, u, v, s);
        default:
            return launch_red<NM, true, 1, 2, 3, F, 1, true>(nelmt, B, D, G, u, v, s);
        }
    });
}

int bk1_apply_red(int form, int variant, int p, size_t nelmt, const double* B, const void* JxW, const void* u,
                  double* v, hipStream_t s)
{
    return apply_red<false>(form, variant, p, nelmt, B, nullptr, JxW, u, v, s);
}

int bk3_apply_red(int form, int variant, int p, size_t nelmt, const double* B, const double* D, const void* G,
                  const void* u, double* v, hipStream_t s)
{
    return apply_red<true>(form, variant, p, nelmt, B, D, G, u, v, s);
}

}  // namespace bkx
