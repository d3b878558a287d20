// bk5_oz.hip -- BK5 collocated Laplacian at p = 5, 6, 7 with FP32 storage and
// its contractions carried by BF16 MFMA products of exact integer digits
// (an Ozaki-style error-free splitting; no reference counterpart: the
// reference runs BK5 in float or double, openmp/BK5.cpp:18-103).
//
// Every contraction out(i, x) = sum_n M(i, n) w(n, x) (M = D for the
// gradient, D^T for the divergence; x runs over the 64 points of the cross
// plane) is evaluated from digits:
//   M(i, n) = 2^{e_M(i)} sum_b dM_b(i, n) 2^{-8(b+1)}   (per-row exponent)
//   w(n, x) = 2^{e_w}    sum_a dw_a(n, x) 2^{-8(a+1)}   (one exponent per
//                                                        element and field)
// with six digits each, dM_b and dw_a integers of at most 8 bits (|d| <= 256,
// exactly representable in BF16), found by the exact chain
// d = rint(r * 256), r <- r * 256 - d.  The products of digits of one level
// t = a + b,
//   C_t(i, x) = sum_{a + b = t} sum_n dM_b(i, n) dw_a(n, x),
// are integers below 2^20 (at most six pairs of eight products of two
// 8-bit digits), so their FP32 accumulation in the MFMA is exact in any
// order.  The output is recombined once, in FP64, from levels 0..5:
//   out(i, x) = 2^{e_M(i) + e_w - 16} sum_t C_t(i, x) 2^{-8t},
// whose truncation (digit pairs of level >= 6, digits beyond the sixth)
// is below 2^-44 of the operands' scale.  Measured against the FP64
// operator on the same FP32 inputs: ~2e-14 max-relative.
//
// MFMA mapping (v_mfma_f32_32x32x16_bf16, one wavefront per element): rows
// of A and C are (level group g, output i), 4 x 8; columns of B and C are 32
// cross points x (two tiles cover the n^2 <= 64 cross points); K = 16 is
// (digit parity s, contraction index n).  For n = 6, 7 (p = 5, 6) the rows,
// columns and K beyond n carry zero digits and their outputs are not
// stored.  Lane l (h = l / 32) holds C[8 (r / 4) + 4 h + r % 4][l % 32],
// r = 0..15, so that every lane holds all level groups of its four outputs:
// the recombination is local to the lane.  Its eight K slots of B carry
// digits a and a + 1 of its four values w(4 h + 0..3, l % 32), and its eight
// K slots of A the digits of M(row l % 32, 4 h + 0..3) that pair with them,
// so each MFMA adds the products of two digits a, a + 1 of w (the two halves
// of K):
//   A_a((g, i), n) = dM_{g - a}(i, n)           into levels 0..3 (a = 0..3),
//   A'_a((g, i), n) = dM_{4 + g - a}(i, n), g < 2, into levels 4, 5 (a = 0..5):
// two plus three MFMAs per direction and 32 cross points.  Which K slots the
// hardware gives a lane does not matter: A and B use the same ones.  The A
// operands (digits of D and D^T) are built once per wavefront.
//
// Flow per element: u streamed once into LDS; per direction, u read in that
// direction's layout (four values per lane), split into digits, gradient
// MFMAs, recombined gradients (FP64) to LDS; one pointwise pass forms the
// three flux fields (rr = g0 qr + g1 qs + g2 qt, rs, rt) in FP64 from the
// gradients and G (FP32, global, read once) in place of the gradients; per
// direction of the divergence, the field it contracts (rr for i, rs for j,
// rt for k) read at that direction's points, split into digits, divergence
// MFMAs, recombined and summed into v in LDS; v stored in FP64.  FP64 is
// used for the recombination of the levels and for the pointwise fluxes (9
// FMAs per point); every contraction product is an exact BF16 product.
//
// Bytes per DoF: u 4 + G 24 (FP32) + v 8 (FP64) = 36, as the other FP32-
// storage forms.
#include <climits>
#include <cmath>

#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

using bf16x4 = short __attribute__((ext_vector_type(4)));
using s16x8 = short __attribute__((ext_vector_type(8)));
using bf16x8 = __bf16 __attribute__((ext_vector_type(8)));
using fx16 = float __attribute__((ext_vector_type(16)));
using bf16x2 = __bf16 __attribute__((ext_vector_type(2)));
using fx2 = float __attribute__((ext_vector_type(2)));

constexpr int kK = 8;          // MFMA K and output rows: points per direction, padded
constexpr int kDig = 6;        // digits per operand
constexpr int kLoLevels = 4;   // levels 0..3 in the first accumulator
constexpr int kOzGrid = 3072;  // workgroups when the runtime cannot report the resident count

// frexp exponent: |x| / 2^e in [0.5, 1) (0 for x = 0)
__device__ __forceinline__ int exp_of(float m)
{
    int e = 0;
    (void)frexpf(m, &e);
    return e;
}
__device__ __forceinline__ int exp_of(double m)
{
    int e = 0;
    (void)frexp(m, &e);
    return e;
}

// the digits of X, |X| < 1: X = sum_a d[a] 2^{-8(a+1)} + (a residual below
// 2^{-49}); every step is exact (the product by 256 and the difference of
// a value and its nearest integer are representable)
template <typename T>
__device__ __forceinline__ void digits(T X, T (&d)[kDig])
{
#pragma clang fp contract(off)
    T r = X;
#pragma unroll
    for (int a = 0; a < kDig; ++a) {
        const T t = r * T(256);
        const T di = rint(t);
        d[a] = di;
        r = t - di;
    }
}

// The same digits of an FP64 X as floats: the first three by the FP64
// chain, the last three by the FP32 chain from the residual rounded to FP32
// (|r| <= 1/2 after three digits, so the rounding costs at most 2^-25 |r|,
// 2^-50 of X's scale, below the truncation of the sixth digit).
__device__ __forceinline__ void digits_mixed(double X, float (&d)[kDig])
{
#pragma clang fp contract(off)
    double r = X;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double t = r * 256.0;
        const double di = rint(t);
        d[a] = float(di);
        r = t - di;
    }
    float rf = float(r);
#pragma unroll
    for (int a = 3; a < kDig; ++a) {
        const float t = rf * 256.0f;
        const float di = rintf(t);
        d[a] = di;
        rf = t - di;
    }
}

// BF16 bit pattern of an integer-valued float of magnitude <= 256 (exact)
__device__ __forceinline__ short bf16_of(float d) { return short(__float_as_uint(d) >> 16); }

// two integer-valued floats (|d| <= 256) as packed BF16: v_cvt_pk_bf16_f32,
// exact for such values
__device__ __forceinline__ uint32_t bf16_pair(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((fx2){a, b}, bf16x2));
}

__device__ __forceinline__ bf16x4 bf16_quad(float a, float b, float c, float d)
{
    const uint32_t lo = bf16_pair(a, b), hi = bf16_pair(c, d);
    return __builtin_bit_cast(bf16x4, (uint64_t(hi) << 32) | lo);
}

// the eight K slots of a lane: the four values of x, then those of y
__device__ __forceinline__ bf16x8 k_pair(bf16x4 x, bf16x4 y)
{
    return __builtin_bit_cast(bf16x8, s16x8(__builtin_shufflevector(x, y, 0, 1, 2, 3, 4, 5, 6, 7)));
}

__device__ __forceinline__ float wave_max(float x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        x = fmaxf(x, __shfl_xor(x, o));
    }
    return x;
}

// Points of direction d, N points per direction: line index n and cross
// point x = (a, b) = (x / N, x % N) address point n sn + a sa + b sb, that is
// (n, a, b) for d = 0, (a, n, b) for d = 1, (a, b, n) for d = 2.  The strides
// are selected once per direction (uniform values, no branches per point).
template <int N>
struct OzDir {
    int sn, sa, sb;
    __device__ __forceinline__ explicit OzDir(int d)
        : sn(d == 0 ? N * N : (d == 1 ? N : 1)), sa(d == 0 ? N : N * N), sb(d == 2 ? N : 1)
    {
    }
    __device__ __forceinline__ int base(int x) const { return (x / N) * sa + (x % N) * sb; }
};

// max over the wavefront of frexp exponents (lanes whose values are all zero
// pass INT_MIN), three at a time: the frexp exponent is nondecreasing in
// |x|, so this is the exponent of the largest magnitude; 0 when every value
// is zero, as exp_of(0)
__device__ __forceinline__ void wave_max_exp3(int (&e)[3])
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            e[c] = max(e[c], __shfl_xor(e[c], o));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e[c] = e[c] == INT_MIN ? 0 : e[c];
    }
}

// x[k] for a lane-dependent k by selects (an index into a private array that
// varies across the lanes would put the array in scratch memory); 0 when k
// is outside 0 .. K - 1
template <int K, typename T>
__device__ __forceinline__ T pick(const T (&x)[K], int k)
{
    T r = T(0);
#pragma unroll
    for (int c = 0; c < K; ++c) {
        r = k == c ? x[c] : r;
    }
    return r;
}

// The A operands of M (D, or D^T with TR) for this lane, and the scales of
// its four output rows: lane l holds row (g, i) = l % 32 of A, columns
// n = 4 h + 0..3 (h = l / 32), and outputs i = 4 h + 0..3.
struct OzMat {
    bf16x8 lo[kLoLevels / 2];   // (A_a, A_{a+1}), a = 0, 2 (levels 0..3)
    bf16x8 hi[kDig / 2];        // (A'_a, A'_{a+1}), a = 0, 2, 4 (levels 4, 5)
    double scale[4];        // 2^{e_M(i) - 16} of the lane's output rows

    template <bool TR, int N>
    __device__ void init(const double* __restrict__ D, int lane)
    {
        // M padded with zeros to 8 x 8
        auto M = [&](int i, int n) { return (i < N && n < N) ? (TR ? D[n * N + i] : D[i * N + n]) : 0.0; };
        const int R = lane % 32, g = R / 8, i = R % 8, h = lane / 32;
        double rowmax[kK];
#pragma unroll
        for (int r = 0; r < kK; ++r) {
            double m = 0.0;
#pragma unroll
            for (int n = 0; n < kK; ++n) {
                m = fmax(m, fabs(M(r, n)));
            }
            rowmax[r] = m;
        }
        const int ei = exp_of(pick(rowmax, i));
        double dig[4][kDig];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            digits(ldexp(M(i, 4 * h + t), -ei), dig[t]);
        }
        bf16x4 l4[kLoLevels], h4[kDig];
#pragma unroll
        for (int a = 0; a < kLoLevels; ++a) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                l4[a][t] = bf16_of(float(pick(dig[t], g - a)));
            }
        }
#pragma unroll
        for (int a = 0; a < kDig; ++a) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                h4[a][t] = g < 2 ? bf16_of(float(pick(dig[t], 4 + g - a))) : short(0);
            }
        }
#pragma unroll
        for (int a = 0; a < kLoLevels / 2; ++a) {
            lo[a] = k_pair(l4[2 * a], l4[2 * a + 1]);
        }
#pragma unroll
        for (int a = 0; a < kDig / 2; ++a) {
            hi[a] = k_pair(h4[2 * a], h4[2 * a + 1]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            scale[r] = ldexp(1.0, exp_of(h == 0 ? rowmax[r] : rowmax[4 + r]) - 16);
        }
    }
};

// sum_t C_t 2^{-8t} of the lane's output r (C_t in the level accumulators).
// Levels 5, 4, 3 are combined in FP32 first: in level-3 units their sum is
// below 2^21, so its FP32 rounding is at most 2^-2 of a level-3 unit, 2^-26
// of a level-0 unit, below 2^-45 of a result of up to 2^20 level-0 units;
// levels 2, 1, 0 are added in FP64.
__device__ __forceinline__ double recombine(const fx16& c_lo, const fx16& c_hi, int r)
{
    const float s3 = fmaf(fmaf(c_hi[4 + r], 0x1p-8f, c_hi[r]), 0x1p-8f, c_lo[12 + r]);   // levels 5, 4, 3
    double s = double(s3);
#pragma unroll
    for (int g = 2; g >= 0; --g) {
        s = fma(s, 0x1p-8, double(c_lo[4 * g + r]));    // levels 2 .. 0
    }
    return s;
}

// out(i, x) for the lane's outputs (i = 4 h + r, x = 32 tile + l % 32) from
// the digits B[a] of w (packed for this lane: n = 4 h + 0..3)
__device__ __forceinline__ void oz_contract(const OzMat& M, const bf16x4 (&B)[kDig], double wscale, double (&out)[4])
{
    fx16 c_lo = {};
    fx16 c_hi = {};
#pragma unroll
    for (int a = 0; a < kLoLevels / 2; ++a) {
        c_lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M.lo[a], k_pair(B[2 * a], B[2 * a + 1]), c_lo, 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < kDig / 2; ++a) {
        c_hi = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M.hi[a], k_pair(B[2 * a], B[2 * a + 1]), c_hi, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out[r] = recombine(c_lo, c_hi, r) * (M.scale[r] * wscale);
    }
}

// the lane's B operands from its four values w[0..3] (n = 4 h + 0..3),
// w / 2^{e_w} with |w| / 2^{e_w} < 1
__device__ __forceinline__ void oz_operands(const float (&w)[4], int ew, bf16x4 (&B)[kDig])
{
    float d[4][kDig];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        digits(ldexp(w[t], -ew), d[t]);
    }
#pragma unroll
    for (int a = 0; a < kDig; ++a) {
        B[a] = bf16_quad(d[0][a], d[1][a], d[2][a], d[3][a]);
    }
}

__device__ __forceinline__ void oz_operands(const double (&w)[4], int ew, bf16x4 (&B)[kDig])
{
    float d[4][kDig];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        digits_mixed(ldexp(w[t], -ew), d[t]);
    }
#pragma unroll
    for (int a = 0; a < kDig; ++a) {
        B[a] = bf16_quad(d[0][a], d[1][a], d[2][a], d[3][a]);
    }
}

// G16: G is 16-byte aligned (used at n = 8 only, where the flux pass reads G
// as 16-byte vectors).  Registers are held to three wavefronts per SIMD
// (159 VGPRs at n = 8, no spills), and 24 N^3 bytes of LDS per wavefront
// allow at least 3.25.
template <int N, bool G16>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void bk5_oz_kernel(
    size_t nelmt, const double* __restrict__ D, const float* __restrict__ G, const float* __restrict__ u,
    double* __restrict__ v)
{
    constexpr int N2 = N * N;
    constexpr int N3 = N2 * N;
    constexpr bool kFull = N == kK;   // no padding: every mask below is true
    // qr, qs, qt of the element; then, in their place, rr, rs, rt.  Two
    // more fields share these boxes:
    // - v, in the box of qr / rr.  Each divergence tile reads all of its
    //   lines of the field it contracts (the 8 line points of each of its
    //   cross points) before it writes v at the same points, and the tiles
    //   cover disjoint lines: direction i overwrites rr only where it has
    //   already read it.
    // - u, as floats, in the top half of the box of qt (floats N3 .. 2 N3 - 1
    //   of it).  Only qt's direction writes that box, and it reads all of
    //   its u values (both tiles) before it writes any.
    __shared__ double s_q[3][N3];
    double* const s_v = s_q[0];
    float* const su = reinterpret_cast<float*>(s_q[2]) + N3;

    const int lane = threadIdx.x;
    const int h = lane / 32;
    OzMat Mg, Md;
    Mg.init<false, N>(D, lane);   // gradient: D(i, n)
    Md.init<true, N>(D, lane);    // divergence: D(n, i)

    for (size_t e = blockIdx.x; e < nelmt; e += gridDim.x) {
        const float* ue = u + e * N3;
        const float* ge = G + e * 6 * N3;

        // u staged once in LDS, as floats, and one exponent for the
        // element's u
        constexpr int R = (N3 + kWave - 1) / kWave;
        float ur[R];
#pragma unroll
        for (int m = 0; m < R; ++m) {
            const int x = m * kWave + lane;
            ur[m] = ld_stream(ue + ((kFull || x < N3) ? x : 0));
        }
        float um = 0.0f;
#pragma unroll
        for (int m = 0; m < R; ++m) {
            const int x = m * kWave + lane;
            if (kFull || x < N3) {
                su[x] = ur[m];
                um = fmaxf(um, fabsf(ur[m]));
            }
        }
        const int eu = exp_of(wave_max(um));
        const double uscale = ldexp(1.0, eu);
        wave_sync();

        // gradient: qr, qs, qt, one direction at a time (unrolled, the loads
        // and digits of all six tiles would be live at once).  Both tiles of
        // a direction read their values first, so that the second tile's
        // digits can be formed while the first one's products run (and no
        // read comes after a write of the direction: qt's writes reach the
        // u staging).  Line indices n >= N and cross points x >= N^2 are
        // padding: their values are zero and their outputs are not stored.
#pragma unroll 1
        for (int d = 0; d < 3; ++d) {
            const OzDir<N> dir(d);
            float w[2][4];
            int b[2];
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
                const int x = 32 * tile + lane % 32;
                b[tile] = dir.base(x) + 4 * h * dir.sn;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool ok = kFull || (4 * h + t < N && x < N2);
                    const float val = su[ok ? b[tile] + t * dir.sn : 0];
                    w[tile][t] = ok ? val : 0.0f;
                }
            }
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
                const int x = 32 * tile + lane % 32;
                bf16x4 B[kDig];
                oz_operands(w[tile], eu, B);
                double q[4];
                oz_contract(Mg, B, uscale, q);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (kFull || (4 * h + r < N && x < N2)) {
                        s_q[d][b[tile] + r * dir.sn] = q[r];
                    }
                }
            }
        }
        wave_sync();

        // the three flux fields (rr, rs, rt) = G (qr, qs, qt), pointwise, in
        // place of the gradients, and each field's largest magnitude.  At
        // n = 8 with G16 each lane takes the points (i, j, k0 .. k0 + 3) of
        // two (i, j) rows and reads G(e, i, j, f, k0 .. k0 + 3) as 16-byte
        // vectors (all twelve loads in flight at once); otherwise one point
        // per lane and round, G read per point.
        double fm[3] = {0.0, 0.0, 0.0};
        auto flux = [&](int pt, float g0, float g1, float g2, float g3, float g4, float g5) {
            const double q0 = s_q[0][pt], q1 = s_q[1][pt], q2 = s_q[2][pt];
            const double rr = double(g0) * q0 + double(g1) * q1 + double(g2) * q2;
            const double rs = double(g1) * q0 + double(g3) * q1 + double(g4) * q2;
            const double rt = double(g2) * q0 + double(g4) * q1 + double(g5) * q2;
            s_q[0][pt] = rr;
            s_q[1][pt] = rs;
            s_q[2][pt] = rt;
            fm[0] = fmax(fm[0], fabs(rr));
            fm[1] = fmax(fm[1], fabs(rs));
            fm[2] = fmax(fm[2], fabs(rt));
        };
        if constexpr (kFull && G16) {
            f32x4 g[2][6];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int q = lane + kWave * r;
                const int i = q / 16, jj = (q / 2) % N, k0 = 4 * (q % 2);
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    g[r][f] = *reinterpret_cast<const f32x4*>(ge + ((i * N + jj) * 6 + f) * N + k0);
                }
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int q = lane + kWave * r;
                const int i = q / 16, jj = (q / 2) % N, k0 = 4 * (q % 2);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    flux((i * N + jj) * N + k0 + t, g[r][0][t], g[r][1][t], g[r][2][t], g[r][3][t], g[r][4][t],
                         g[r][5][t]);
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int x = m * kWave + lane;
                const int pt = (kFull || x < N3) ? x : 0;
                const int ii = pt / N2, jj = (pt / N) % N, kk = pt % N;
                const float* gp = ge + (ii * N + jj) * 6 * N + kk;
                float g[6];
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    g[f] = gp[f * N];
                }
                if (kFull || x < N3) {
                    flux(pt, g[0], g[1], g[2], g[3], g[4], g[5]);
                }
            }
        }
        wave_sync();

        // divergence: per direction, the flux field it contracts, read at its
        // own points
        int efs[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            efs[c] = fm[c] > 0.0 ? exp_of(fm[c]) : INT_MIN;
        }
        wave_max_exp3(efs);
#pragma unroll 1
        for (int d = 0; d < 3; ++d) {
            const OzDir<N> dir(d);
            const int ef = d == 0 ? efs[0] : (d == 1 ? efs[1] : efs[2]);
            const double fscale = ldexp(1.0, ef);
            double ft[2][4];
            int b[2];
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
                const int x = 32 * tile + lane % 32;
                b[tile] = dir.base(x) + 4 * h * dir.sn;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool ok = kFull || (4 * h + t < N && x < N2);
                    const double val = s_q[d][ok ? b[tile] + t * dir.sn : 0];
                    ft[tile][t] = ok ? val : 0.0;
                }
            }
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
                const int x = 32 * tile + lane % 32;
                bf16x4 B[kDig];
                oz_operands(ft[tile], ef, B);
                double o[4];
                oz_contract(Md, B, fscale, o);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (kFull || (4 * h + r < N && x < N2)) {
                        const int pt = b[tile] + r * dir.sn;
                        s_v[pt] = d == 0 ? o[r] : s_v[pt] + o[r];
                    }
                }
            }
            wave_sync();
        }

        double* ve = v + e * N3;
#pragma unroll
        for (int m = 0; m < (N3 + kWave - 1) / kWave; ++m) {
            const int x = m * kWave + lane;
            if (kFull || x < N3) {
                st_stream(ve + x, s_v[x]);
            }
        }
        wave_sync();
    }
}

}  // namespace

namespace {

// One wavefront per workgroup, looping over the elements; as many workgroups
// as are resident at once (three per SIMD at every order), so that none
// waits for another to finish its elements.
template <auto Kernel>
int launch_oz(size_t nelmt, const double* D, const float* G, const float* u, double* v, hipStream_t s)
{
    unsigned slots = resident_blocks<Kernel>(kWave);
    if (slots == 0) {
        slots = kOzGrid;
    }
    const unsigned grid = unsigned(nelmt < size_t(slots) ? nelmt : size_t(slots));
    hipLaunchKernelGGL(Kernel, dim3(grid), dim3(kWave), 0, s, nelmt, D, G, u, v);
    return launch_status();
}

}  // namespace

int bk5_apply_oz(int p, size_t nelmt, const double* D, const float* G, const float* u, double* v, hipStream_t s)
{
    switch (p) {
    case 5:
        return launch_oz<bk5_oz_kernel<6, false>>(nelmt, D, G, u, v, s);
    case 6:
        return launch_oz<bk5_oz_kernel<7, false>>(nelmt, D, G, u, v, s);
    case 7:
        return reinterpret_cast<uintptr_t>(G) % 16 == 0 ? launch_oz<bk5_oz_kernel<8, true>>(nelmt, D, G, u, v, s)
                                                        : launch_oz<bk5_oz_kernel<8, false>>(nelmt, D, G, u, v, s);
    default:
        return BKX_ERROR_ORDER;
    }
}

}  // namespace bkx
