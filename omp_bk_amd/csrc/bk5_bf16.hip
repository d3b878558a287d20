// bk5_bf16.hip -- BK5 collocated Laplacian at p = 7 with split-BF16 MFMA
// contractions accumulated in FP32 (a reduced-precision form; no reference
// counterpart: the reference runs BK5 in float or double, openmp/BK5.cpp).
//
// Every value x entering a contraction is split into two BF16 parts,
// x ~ x_hi + x_lo (x_hi = x truncated to 8 mantissa bits, x_lo = x - x_hi
// truncated again: about 16 significant bits), and so is D.  Per element and
// direction the contraction out(x, i) = sum_n D(i, n) w(n, x) over the 64
// cross points x runs as four v_mfma_f32_16x16x16_bf16, one per 16 cross
// points, with
//   A(16 x 16):  row = cross point, column k = (w part, n): w_hi for k < 8, w_lo for k >= 8
//   B(16 x 16):  row k = (w part, n), column = (D part, i): D_hi(i, n), D_lo(i, n)
// so that C(x, (i, s)) = sum_n (w_hi + w_lo)(n, x) D_s(i, n), and the two D
// parts (columns i and i + 8, in lanes l and l + 8) are added with a DPP
// row rotation.  The products of BF16 parts are exact in FP32; the sums are
// FP32.  The gradient uses D, the transposed contractions D^T; the fluxes
// (rr, rs, rt) = G (qr, qs, qt) are FP32 FMAs.  u, G and v are FP32
// (4 + 24 + 4 = 32 bytes per DoF).  The expected error is that of 16-bit
// operands, about 1e-5 relative.
//
// Mapping: one wavefront per element (n = 8: 512 points); u, the three
// gradients (then fluxes) and the output (in u's box) accumulate in LDS, all
// in-wavefront.
// The MFMA operand layout of v_mfma_f32_16x16x16_bf16 (64 lanes, 16 x 16 x 16):
//   lane l holds A[l % 16][4 (l / 16) + 0..3], B[4 (l / 16) + 0..3][l % 16],
//   and C[4 (l / 16) + 0..3][l % 16].
#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

using bf16x4 = short __attribute__((ext_vector_type(4)));
using fx4 = float __attribute__((ext_vector_type(4)));

constexpr int NQ = 8;
constexpr int NQ2 = 64;
constexpr int NQ3 = 512;
constexpr int ROW = NQ + 1;            // padded k rows
constexpr int PLANE = NQ * ROW;
constexpr int BOX = NQ * PLANE;        // one padded element box of floats

__device__ __forceinline__ int PX(int i, int j, int k) { return i * PLANE + j * ROW + k; }

// upper 16 bits of x: its BF16 truncation
__device__ __forceinline__ short bf16_hi_bits(float x) { return short(__float_as_uint(x) >> 16); }

// (hi, lo) BF16 parts of x as bit patterns
__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo)
{
    const float h = __uint_as_float(__float_as_uint(x) & 0xffff0000u);
    hi = short(__float_as_uint(h) >> 16);
    lo = bf16_hi_bits(x - h);
}

// point (i, j, k) of cross point x = (a, b) and line index n for direction d:
// d = 0: line i, cross (j, k); d = 1: line j, cross (i, k); d = 2: line k, cross (i, j)
__device__ __forceinline__ int point(int d, int n, int x)
{
    const int a = x / NQ, b = x % NQ;
    return d == 0 ? PX(n, a, b) : (d == 1 ? PX(a, n, b) : PX(a, b, n));
}

// B operand of lane l for the matrix M (D or D^T): B[k][c] with k = 4 (l / 16)
// + t, n = k % 8, c = l % 16, i = c % 8, part = c / 8: M_part(i, n).
__device__ __forceinline__ bf16x4 b_operand(const double* __restrict__ D, bool transpose, int lane)
{
    bf16x4 r;
    const int c = lane % 16;
    const int i = c % 8;
    const int part = c / 8;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = 4 * (lane / 16) + t;
        const int n = k % 8;
        const float m = float(transpose ? D[n * NQ + i] : D[i * NQ + n]);
        short hi, lo;
        split_bf16(m, hi, lo);
        r[t] = part == 0 ? hi : lo;
    }
    return r;
}

// out(point of (d, i, x)) (=, +=) sum_n M(i, n) w(point of (d, n, x)) for all 64
// cross points x and 8 outputs i, w read from src (a padded box).  With ACC
// the result is added to dst.
template <bool ACC>
__device__ __forceinline__ void contract(const float* src, float* dst, int d, bf16x4 bop, int lane)
{
    const int grp = lane / 16;
    const bool lo_part = grp >= 2;
    const int n0 = 4 * (grp % 2);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int x = 16 * t + lane % 16;
        bf16x4 a;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            short hi, lo;
            split_bf16(src[point(d, n0 + s, x)], hi, lo);
            a[s] = lo_part ? lo : hi;
        }
        fx4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bop, c, 0, 0, 0);
        // lane l holds C[4 grp + r][l % 16]: column (i, part) = (l % 8, (l % 16) / 8);
        // add the D_lo column (lane l + 8) to the D_hi column (lane l)
        const int col = lane % 16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // row_ror:8 within each row of 16 lanes: lane l receives lane (l + 8) mod 16.
            // (The element is copied to a scalar first: a bit cast of the vector
            // element expression c[r] itself selected element 0 for every r.)
            const float cr = c[r];
            const float other = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(cr), 0x128, 0xf, 0xf, false));
            if (col < 8) {
                const int xr = 16 * t + 4 * grp + r;
                const int pt = point(d, col, xr);
                const float val = cr + other;
                dst[pt] = ACC ? dst[pt] + val : val;
            }
        }
    }
}

// G16: G is 16-byte aligned, and the fluxes read it as 16-byte vectors:
// each lane takes the points (i, j, k0 .. k0 + 3) of two (i, j) rows and
// loads G(e, i, j, f, k0 .. k0 + 3) for f = 0..5 (12 loads per lane instead
// of 48 four-byte ones).
template <int WAVES, bool G16>
__global__ __launch_bounds__(kWave * WAVES) void bk5_bf16_kernel(size_t nelmt, const double* __restrict__ D,
                                                                const float* __restrict__ G,
                                                                const float* __restrict__ u, float* __restrict__ v)
{
    // u's box holds v once the gradients are formed (four boxes per
    // wavefront: LDS then allows the four waves per SIMD the registers do)
    __shared__ float s_u[WAVES][BOX];
    __shared__ float s_q[WAVES][3][BOX];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const size_t e = size_t(blockIdx.x) * WAVES + wave;
    const bool active = e < nelmt;
    const size_t ee = active ? e : 0;
    float* su = s_u[wave];
    float* sv = s_u[wave];
    const int j = lane / NQ;
    const int k = lane % NQ;

    const bf16x4 bg = b_operand(D, false, lane);   // gradient: D(i, n)
    const bf16x4 bt = b_operand(D, true, lane);    // transposed: D(n, i)

    // u in the (j, k) plane of each layer i (loaded unconditionally: an idle
    // wavefront computes on element 0 and stores nothing)
    const float* ue = u + ee * NQ3 + lane;
    float uv[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        uv[i] = ld_stream(ue + i * NQ2);
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        su[PX(i, j, k)] = uv[i];
    }
    wave_sync();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        contract<false>(su, s_q[wave][d], d, bg, lane);
    }
    wave_sync();
    auto flux = [&](int pt, float g0, float g1, float g2, float g3, float g4, float g5) {
        const float qr = s_q[wave][0][pt], qs = s_q[wave][1][pt], qt = s_q[wave][2][pt];
        s_q[wave][0][pt] = g0 * qr + g1 * qs + g2 * qt;
        s_q[wave][1][pt] = g1 * qr + g3 * qs + g4 * qt;
        s_q[wave][2][pt] = g2 * qr + g4 * qs + g5 * qt;
    };
    if constexpr (G16) {
        // lane row q = lane + 64 r: (i, j) = (q / 16, (q / 2) % 8), k0 = 4 (q % 2)
        const float* gE = G + ee * 6 * NQ3;
        f32x4 g[2][6];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int q = lane + kWave * r;
            const int i = q / 16, jj = (q / 2) % NQ, k0 = 4 * (q % 2);
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                // ordinary loads: the six factor loads of a row share its
                // lines (streamed, they ran at 123 against 175 GDoF/s)
                g[r][f] = *reinterpret_cast<const f32x4*>(gE + ((i * NQ + jj) * 6 + f) * NQ + k0);
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int q = lane + kWave * r;
            const int i = q / 16, jj = (q / 2) % NQ, k0 = 4 * (q % 2);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                flux(PX(i, jj, k0 + t), g[r][0][t], g[r][1][t], g[r][2][t], g[r][3][t], g[r][4][t], g[r][5][t]);
            }
        }
    } else {
        // lane (j, k), layer i: G(e, i, j, f, k)
        const float* ge = G + ee * 6 * NQ3 + j * 6 * NQ + k;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            float g[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = ge[i * 6 * NQ2 + f * NQ];
            }
            flux(PX(i, j, k), g[0], g[1], g[2], g[3], g[4], g[5]);
        }
    }
    wave_sync();
    contract<false>(s_q[wave][0], sv, 0, bt, lane);
    wave_sync();
    contract<true>(s_q[wave][1], sv, 1, bt, lane);
    wave_sync();
    contract<true>(s_q[wave][2], sv, 2, bt, lane);
    wave_sync();
    if (active) {
        float* ve = v + e * NQ3 + lane;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            st_stream(ve + i * NQ2, sv[PX(i, j, k)]);
        }
    }
}

}  // namespace

int bk5_apply_bf16s(int p, size_t nelmt, const double* D, const float* G, const float* u, float* v, hipStream_t s)
{
    if (p != 7) {
        return BKX_ERROR_ORDER;
    }
    constexpr int waves = 2;
    const unsigned grid = grid_for(nelmt, waves);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    if (reinterpret_cast<uintptr_t>(G) % 16 == 0) {
        hipLaunchKernelGGL((bk5_bf16_kernel<waves, true>), dim3(grid), dim3(kWave * waves), 0, s, nelmt, D, G, u,
                           v);
    } else {
        hipLaunchKernelGGL((bk5_bf16_kernel<waves, false>), dim3(grid), dim3(kWave * waves), 0, s, nelmt, D, G, u,
                           v);
    }
    return launch_status();
}

}  // namespace bkx
