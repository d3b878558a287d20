// bk3_adj.hip -- BK3 with a per-quadrature-point coefficient kappa scaling
// the six geometric factors, and its adjoint, in FP64 on gfx950.
//
//   A(kappa) u = sum_x P_x^T kappa_x G_x P_x u   per element, where P_x u is
//   the reference gradient (qr, qs, qt) at Gauss point x (the operator of
//   openmp/bk3_sumfact.h:17-187 with G_x replaced by kappa_x G_x).
//
// A(kappa) is symmetric, so for L = 1/2 |A(kappa) u - b|^2 and r = A(kappa) u - b
//   dL/du       = A(kappa) r
//   dL/dkappa_x = (P_x r) . (G_x P_x u)
// and for the vector-Jacobian product of v = A(kappa) u with a cotangent w
//   grad_u = A(kappa) w,   grad_kappa_x = (P_x w) . (G_x P_x u).
//
// One fused kernel per element (no quadrature-point data goes to HBM): the
// gradient of u, F = G P u kept in registers, the divergence of kappa F
// (loss mode: r = A u - b and 1/2 |r|^2 per element), the gradient of r (or
// w), dL/dkappa = P r . F, and the divergence of kappa G P r.  G and kappa
// are read twice.  At about 2.3 resident waves per SIMD some 300 elements
// per XCD are in flight and their factors (28.7 KB per element at p = 6 in
// FP64) overflow the 4 MB L2 between the two reads, so the second read is
// largely served from the MALL (rocprofv3: 1.85x the algorithmic read
// bytes).  It costs little time: served from a resident element instead,
// the kernel ran 1% faster, and with both reads from a resident element
// (no G traffic at all) 20% faster -- the kernel is bound by its
// instruction stream and memory latency at three waves per SIMD, not by
// HBM.  Keeping the first read's factors in registers (56 values per lane)
// needs 256 VGPRs and one wave per SIMD, which measured 10-17% slower; what
// is kept is that every other stream -- u, b (w), the outputs and the
// second, last read of G and kappa -- is non-temporal (p = 6, E = 16384:
// 169 -> 156 us in FP64, 128 -> 119 us with FP32 storage, same bits).  The
// phases are those of the transposing BK3 mapping (bk3_phases.hpp).
//
// TS is the storage type of the inputs u, b (or w), kappa and G: double for
// the FP64 entry points, float for the FP32-storage forms, which compute
// either in FP64 (_f32w: each value converted when it is loaded, so that the
// FP64 operator is applied to the FP32 inputs exactly) or in float-float
// (_f32c, FF set: the scheme of the reference's float-float BK1,
// openmp/BK1_ff.cpp:73-241 -- B and D split once into hi + lo floats, every
// contraction, flux and product a Dot2 compensated sum or TwoProd of
// bkx_ff.hpp, the loaded floats exact).  dL/du, dL/dkappa and the loss are
// FP64 in every form (the float-float ones joined with one add).
#include <mutex>

#include "bk3_phases.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

enum AdjMode { kLoss = 0, kVjp = 1, kApply = 2 };

// Elements per workgroup: floor(256 / q^2) divided by adj_div(p).  The
// kernel holds up to 168 VGPRs and three LDS boxes per element, so at most
// orders fewer elements per workgroup let more workgroups share a CU; the
// divisors are the fastest measured on MI355X (the adjoint at 5.6 M DoF, p =
// 6: one element per workgroup, 31.4 GDoF/s against 25.2 with four).
constexpr int adj_div(int p) { return (p == 4 || p == 7) ? 1 : (p == 8 ? 2 : 4); }

// Waves per SIMD the FP64-arithmetic kernels are held to (the occupancy
// their registers give without spilling; unbounded, the unmasked loads of
// p = 6 take 174 VGPRs and two waves).  The float-float kernels are left
// unbounded.
constexpr int adj_waves(int nm, bool ff) { return ff ? 1 : (nm <= 2 ? 5 : (nm <= 4 ? 4 : (nm <= 7 ? 3 : 2))); }

template <int NM, int NQ>
struct AdjShape {
    static constexpr int nq2 = NQ * NQ;
    static constexpr int full = 256 / nq2 > 0 ? 256 / nq2 : 1;
    static constexpr int epb = full / adj_div(NM - 1) > 0 ? full / adj_div(NM - 1) : 1;
    static constexpr int threads = ((epb * nq2 + kWave - 1) / kWave) * kWave;
};

// Per-point arithmetic of the kernel in FP64 (V double; L double, the type
// a loaded value is kept in) or float-float (V ff; L float, exact).
//   flux(a, b, c, g0, g1, g2) = g0 a + g1 b + g2 c
//   scale(k, x)               = k x
//   dot3(a, x, b, y, c, z)    = a x + b y + c z
//   sub(o, b)                 = o - b
__device__ __forceinline__ double flux(double a, double b, double c, double g0, double g1, double g2)
{
    return g0 * a + g1 * b + g2 * c;
}
__device__ __forceinline__ ff flux(ff a, ff b, ff c, float g0, float g1, float g2)
{
    dot2 d;
    d.add(a, g0);
    d.add(b, g1);
    d.add(c, g2);
    return d.value();
}
__device__ __forceinline__ double scale(double k, double x) { return k * x; }
__device__ __forceinline__ ff scale(float k, ff x) { return ff_mul(x, k); }
__device__ __forceinline__ double dot3(double a, double x, double b, double y, double c, double z)
{
    return a * x + b * y + c * z;
}
__device__ __forceinline__ ff dot3(ff a, ff x, ff b, ff y, ff c, ff z)
{
    dot2 d;
    d.add(a, x);
    d.add(b, y);
    d.add(c, z);
    return d.value();
}
__device__ __forceinline__ double sub(double o, double b) { return o - b; }
__device__ __forceinline__ ff sub(ff o, float b)
{
    dot2 d;
    d.add(o);
    d.add(ff{-b, 0.0f});
    return d.value();
}
__device__ __forceinline__ double as_f64(double x) { return x; }
__device__ __forceinline__ double as_f64(ff x) { return ff_join(x); }
__device__ __forceinline__ double as_v(double x) { return x; }
__device__ __forceinline__ ff as_v(float x) { return ff{x, 0.0f}; }

template <bool FF>
struct AdjTypes {
    using V = double;
    using L = double;
    using M = double;
};
template <>
struct AdjTypes<true> {
    using V = ff;
    using L = float;
    using M = ff;
};

// EO: Bm and Dm point at the even-odd forms of B and D (bkx_eo.hpp).
// FF: float-float arithmetic (TS float); B and D are split into an LDS copy.
template <int NM, int NQ, int MODE, typename TS, bool EO, bool FF = false>
__global__ __launch_bounds__((AdjShape<NM, NQ>::threads))
__attribute__((amdgpu_waves_per_eu(adj_waves(NM, FF)))) void bk3_adj_kernel(
    size_t nelmt, const double* __restrict__ Bg, const double* __restrict__ Dg, const TS* __restrict__ G,
    const TS* __restrict__ kappa, const TS* __restrict__ u, const TS* __restrict__ bw,
    double* __restrict__ out_u, double* __restrict__ out_k, double* __restrict__ loss)
{
    static_assert(!FF || (!EO && sizeof(TS) == 4), "float-float: FP32 storage, direct bases");
    using S = AdjShape<NM, NQ>;
    using V = typename AdjTypes<FF>::V;
    using L = typename AdjTypes<FF>::L;
    using MT = typename AdjTypes<FF>::M;
    using E = bk3p::Elem<NM, NQ, V>;
    constexpr int nq2 = S::nq2;
    __shared__ V s_box[S::epb][3 * E::box];

    const MT* Bm;
    const MT* Dm;
    if constexpr (FF) {
        __shared__ ff s_bases[NM * NQ + NQ * NQ];
        bk3p::split_bases<NM, NQ>(Bg, Dg, s_bases, s_bases + NM * NQ);
        __syncthreads();
        Bm = s_bases;
        Dm = s_bases + NM * NQ;
    } else {
        Bm = Bg;
        Dm = Dg;
    }

    const int slot = threadIdx.x / nq2;
    const int l = threadIdx.x - slot * nq2;
    const bool in_elem = slot < S::epb;
    const size_t e = size_t(blockIdx.x) * S::epb + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    E el;
    el.init(s_box[in_elem ? slot : 0], l, in_elem);
    const bool modal = in_elem && l < E::nm2;   // lane (j,k) of the modal plane

    // Loads are unconditional, from clamped addresses inside a real element
    // (element 0 for idle lanes, column 0 for lanes outside the modal plane),
    // and their values are not masked: an idle slot computes on element 0's
    // data in its own LDS boxes and stores nothing, and a lane outside the
    // modal plane holds modal values that no phase reads (r, whose squares
    // are summed, is masked where it is formed).  A conditional load would
    // be waited for where its block ends, and a select per loaded value cost
    // about 3% of the kernel's time in FP64, 6-11% with FP32 storage.
    const int lm = modal ? l : 0;
    auto modal_load = [&](const TS* base, L (&dst)[NM]) {
        TS raw[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            raw[i] = ld_stream(base + ee * E::nm3 + i * E::nm2 + lm);
        }
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            dst[i] = L(raw[i]);
        }
    };

    // gradient of u
    L x[NM];
    modal_load(u, x);
    V qr[NQ];
    el.template grad<0, EO>(Bm, Dm, x, qr);

    // F = G P u (kept), kappa F -> rr (registers), rs -> B, rt -> C
    const TS* ge = G + ee * 6 * E::nq3 + l;      // G(e, f, p, q, r) at f = p = 0
    const TS* ke = kappa + ee * E::nq3 + l;      // kappa(e, p, q, r) at p = 0
    V F0[NQ], F1[NQ], F2[NQ], rr[NQ];
    if (in_elem) {
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            TS graw[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                graw[f] = ge[f * E::nq3 + p * nq2];
            }
            const TS kraw = ke[p * nq2];
            L g[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = L(graw[f]);
            }
            const L kap = L(kraw);
            const int xx = E::X(p, el.lq, el.lq2);
            const V qs = el.B[xx];
            const V qt = el.C[xx];
            F0[p] = flux(qr[p], qs, qt, g[0], g[1], g[2]);
            F1[p] = flux(qr[p], qs, qt, g[1], g[3], g[4]);
            F2[p] = flux(qr[p], qs, qt, g[2], g[4], g[5]);
            rr[p] = scale(kap, F0[p]);
            el.B[xx] = scale(kap, F1[p]);
            el.C[xx] = scale(kap, F2[p]);
        }
    }
    __syncthreads();

    V r[NM];
    if constexpr (MODE == kVjp) {
        L w[NM];
        modal_load(bw, w);
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            r[i] = as_v(w[i]);
        }
    } else {
        V o[NM];
        el.template div<10, EO>(Bm, Dm, rr, o);
        if constexpr (MODE == kApply) {
            if (active && modal) {
#pragma unroll
                for (int i = 0; i < NM; ++i) {
                    st_stream(out_u + e * E::nm3 + i * E::nm2 + l, as_f64(o[i]));
                }
            }
            return;
        } else {
            L bv[NM];
            modal_load(bw, bv);
            V part;
            if constexpr (FF) {
                dot2 acc;
#pragma unroll
                for (int i = 0; i < NM; ++i) {
                    r[i] = (active && modal) ? sub(o[i], bv[i]) : ff{0.0f, 0.0f};
                    acc.add(r[i], r[i]);
                }
                part = acc.value();
            } else {
                part = 0.0;
#pragma unroll
                for (int i = 0; i < NM; ++i) {
                    r[i] = (active && modal) ? sub(o[i], bv[i]) : 0.0;
                    part += r[i] * r[i];
                }
            }
            // 1/2 |r_e|^2: the element's lanes summed in lane order through box C
            // (free until the third phase of the next gradient)
            if (modal) {
                el.C[l] = part;
            }
            __syncthreads();
            if (active && l == 0) {
                if constexpr (FF) {
                    dot2 s;
                    for (int t = 0; t < E::nm2; ++t) {
                        s.add(el.C[t]);
                    }
                    loss[e] = 0.5 * ff_join(s.value());
                } else {
                    double s = 0.0;
                    for (int t = 0; t < E::nm2; ++t) {
                        s += el.C[t];
                    }
                    loss[e] = 0.5 * s;
                }
            }
            __syncthreads();
        }
    }

    // gradient of r, dL/dkappa = P r . F, kappa G P r
    V qrr[NQ];
    el.template grad<20, EO>(Bm, Dm, r, qrr);
    if (in_elem) {
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            TS graw[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                graw[f] = ld_stream(ge + f * E::nq3 + p * nq2);   // the last use of G and kappa
            }
            const TS kraw = ld_stream(ke + p * nq2);
            L g[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = L(graw[f]);
            }
            const L kap = L(kraw);
            const int xx = E::X(p, el.lq, el.lq2);
            const V qs = el.B[xx];
            const V qt = el.C[xx];
            if (active) {
                st_stream(out_k + e * E::nq3 + p * nq2 + l, as_f64(dot3(qrr[p], F0[p], qs, F1[p], qt, F2[p])));
            }
            rr[p] = scale(kap, flux(qrr[p], qs, qt, g[0], g[1], g[2]));
            el.B[xx] = scale(kap, flux(qrr[p], qs, qt, g[1], g[3], g[4]));
            el.C[xx] = scale(kap, flux(qrr[p], qs, qt, g[2], g[4], g[5]));
        }
    }
    __syncthreads();
    V o2[NM];
    el.template div<30, EO>(Bm, Dm, rr, o2);
    if (active && modal) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            st_stream(out_u + e * E::nm3 + i * E::nm2 + l, as_f64(o2[i]));
        }
    }
}

template <int NM, int MODE, typename TS, bool EO = false, bool FF = false>
int launch_adj(size_t nelmt, const double* B, const double* D, const TS* G, const TS* kappa, const TS* u,
               const TS* bw, double* out_u, double* out_k, double* loss, hipStream_t s)
{
    using S = AdjShape<NM, NM + 1>;
    const unsigned grid = grid_for(nelmt, S::epb);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk3_adj_kernel<NM, NM + 1, MODE, TS, EO, FF>), dim3(grid), dim3(S::threads), 0, s, nelmt, B,
                       D, G, kappa, u, bw, out_u, out_k, loss);
    return launch_status();
}

// deterministic sum of n doubles into *out: one workgroup, fixed order
// (thread t adds x[t], x[t + T], x[t + 2T], ... in turn, then a tree over
// the threads).  The loads of kSumUnroll consecutive terms are issued
// together; the additions keep their order.
constexpr int kSumThreads = 1024;
constexpr int kSumUnroll = 16;

__global__ __launch_bounds__(kSumThreads) void sum_kernel(const double* __restrict__ x, size_t n,
                                                          double* __restrict__ out)
{
    __shared__ double part[kSumThreads];
    double s = 0.0;
    for (size_t i0 = threadIdx.x; i0 < n; i0 += size_t(kSumThreads) * kSumUnroll) {
        double t[kSumUnroll];
#pragma unroll
        for (int k = 0; k < kSumUnroll; ++k) {
            const size_t i = i0 + size_t(k) * kSumThreads;
            t[k] = x[i < n ? i : i0];
        }
#pragma unroll
        for (int k = 0; k < kSumUnroll; ++k) {
            if (i0 + size_t(k) * kSumThreads < n) {
                s += t[k];
            }
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumThreads / 2; w > 0; w /= 2) {
        if (int(threadIdx.x) < w) {
            part[threadIdx.x] += part[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = part[0];
    }
}

}  // namespace

template <typename TS, bool FF = false>
int bk3_apply_kappa(int p, size_t nelmt, const double* B, const double* D, const TS* G, const TS* kappa, const TS* u,
                    double* v, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        return launch_adj<decltype(P)::value + 1, kApply, TS, false, FF>(nelmt, B, D, G, kappa, u, nullptr, v,
                                                                         nullptr, nullptr, s);
    });
}

template <typename TS, bool FF = false>
int bk3_adjoint(int p, size_t nelmt, const double* B, const double* D, const TS* G, const TS* kappa, const TS* u,
                const TS* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        return launch_adj<decltype(P)::value + 1, kLoss, TS, false, FF>(nelmt, B, D, G, kappa, u, b, dldu, dldk,
                                                                        loss, s);
    });
}

// The loss mode with the bases in even-odd form (P: B's form, then D's).
template <typename TS>
int bk3_adjoint_eo(int p, size_t nelmt, const double* P, const TS* G, const TS* kappa, const TS* u, const TS* b,
                   double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto O) {
        constexpr int NM = decltype(O)::value + 1;
        return launch_adj<NM, kLoss, TS, true>(nelmt, P, P + EoMat<NM, NM + 1, 1>::size, G, kappa, u, b, dldu, dldk,
                                               loss, s);
    });
}

template <typename TS, bool FF = false>
int bk3_vjp(int p, size_t nelmt, const double* B, const double* D, const TS* G, const TS* kappa, const TS* u,
            const TS* w, double* gu, double* gk, hipStream_t s)
{
    return with_order<1, 8>(p, [&](auto P) {
        return launch_adj<decltype(P)::value + 1, kVjp, TS, false, FF>(nelmt, B, D, G, kappa, u, w, gu, gk, nullptr,
                                                                       s);
    });
}

int bk3_apply_kappa_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                        const double* u, double* v, hipStream_t s)
{
    return bk3_apply_kappa(p, nelmt, B, D, G, kappa, u, v, s);
}

int bk3_adjoint_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                    const double* u, const double* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return bk3_adjoint(p, nelmt, B, D, G, kappa, u, b, dldu, dldk, loss, s);
}

int bk3_vjp_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                const double* u, const double* w, double* gu, double* gk, hipStream_t s)
{
    return bk3_vjp(p, nelmt, B, D, G, kappa, u, w, gu, gk, s);
}

int bk3_apply_kappa_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                         const float* u, double* v, hipStream_t s)
{
    return bk3_apply_kappa(p, nelmt, B, D, G, kappa, u, v, s);
}

int bk3_adjoint_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                     const float* u, const float* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return bk3_adjoint(p, nelmt, B, D, G, kappa, u, b, dldu, dldk, loss, s);
}

int bk3_vjp_f32w(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                 const float* u, const float* w, double* gu, double* gk, hipStream_t s)
{
    return bk3_vjp(p, nelmt, B, D, G, kappa, u, w, gu, gk, s);
}

int bk3_apply_kappa_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                         const float* u, double* v, hipStream_t s)
{
    return bk3_apply_kappa<float, true>(p, nelmt, B, D, G, kappa, u, v, s);
}

int bk3_adjoint_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                     const float* u, const float* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return bk3_adjoint<float, true>(p, nelmt, B, D, G, kappa, u, b, dldu, dldk, loss, s);
}

int bk3_vjp_f32c(int p, size_t nelmt, const double* B, const double* D, const float* G, const float* kappa,
                 const float* u, const float* w, double* gu, double* gk, hipStream_t s)
{
    return bk3_vjp<float, true>(p, nelmt, B, D, G, kappa, u, w, gu, gk, s);
}

int bk3_adjoint_eo_f64(int p, size_t nelmt, const double* P, const double* G, const double* kappa, const double* u,
                       const double* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return bk3_adjoint_eo(p, nelmt, P, G, kappa, u, b, dldu, dldk, loss, s);
}

int bk3_adjoint_eo_f32w(int p, size_t nelmt, const double* P, const float* G, const float* kappa, const float* u,
                        const float* b, double* dldu, double* dldk, double* loss, hipStream_t s)
{
    return bk3_adjoint_eo(p, nelmt, P, G, kappa, u, b, dldu, dldk, loss, s);
}

int sum_f64(const double* x, size_t n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kSumThreads), 0, s, x, n, out);
    return launch_status();
}

// ---------------------------------------------------------------------------
// The loss of the FP64 adjoint with nothing allocated: the per-element losses
// that bk3_adjoint_f64 writes to a caller's array go to `scratch`, a
// library-owned array of kLossScratchMax doubles (8 MiB) on every GPU the
// library runs on, held until the process ends; then sum_kernel adds them
// into the caller's L.  Every entry the sum reads was written by the adjoint
// kernel of the same call, so nothing is cleared.  In host code `scratch`
// names a host-side stand-in, not a device address: kernels are given the
// address hipGetSymbolAddress reports for the device, resolved once per device
// with that device current.  Calls that overlap on one GPU share the array
// and race.
constexpr size_t kLossScratchMax = BKX_ADJOINT_LOSS_MAX;
static __device__ __align__(256) double scratch[kLossScratchMax];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_by_device[kMaxDevices] = {};

// The device address of `scratch` on device dev; nullptr if it cannot be had.
double* loss_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) {
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_by_device[dev]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) {
            return nullptr;
        }
        if (cur != dev && hipSetDevice(dev) != hipSuccess) {
            return nullptr;
        }
        void* ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) {
            (void)hipSetDevice(cur);
        }
        if (err != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return nullptr;
        }
        scratch_by_device[dev] = static_cast<double*>(ptr);
    }
    return scratch_by_device[dev];
}

}  // namespace

int bk3_adjoint_loss_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* kappa,
                         const double* u, const double* b, double* dldu, double* dldk, double* L, hipStream_t s)
{
    if (nelmt > kLossScratchMax) {
        return BKX_ERROR_SIZE;
    }
    int dev = -1;
    if ((s ? hipStreamGetDevice(s, &dev) : hipGetDevice(&dev)) != hipSuccess) {
        (void)hipGetLastError();
        return BKX_ERROR_DEVICE;
    }
    double* part = loss_scratch(dev);
    if (!part) {
        return BKX_ERROR_DEVICE;
    }
    if (nelmt > 0) {
        const int st = bk3_adjoint(p, nelmt, B, D, G, kappa, u, b, dldu, dldk, part, s);
        if (st != BKX_SUCCESS) {
            return st;
        }
    }
    return sum_f64(part, nelmt, L, s);   // L = 0 for no elements
}

}  // namespace bkx
