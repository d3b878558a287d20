// bk5.hip -- BK5 collocated Laplacian with FP64 arithmetic on gfx950, with
// u and G stored in FP64 (the drop-in) or in FP32 (half the bytes of the
// geometric factors and of u; v is always written in FP64).
//
// Operator (openmp/BK5.cpp:18-103): per element, with D(i, n) = dbasis[i*nq+n]
// and the six symmetric factors G(e, i, j, f, k),
//   qr = sum_n D(i,n) u(n,j,k),  qs = sum_n D(j,n) u(i,n,k),  qt = sum_n D(k,n) u(i,j,n)
//   (rr, rs, rt) = G(i,j,k) (qr, qs, qt)
//   v(i,j,k) = sum_n D(n,i) rr(n,j,k) + D(n,j) rs(i,n,k) + D(n,k) rt(i,j,n).
//
// Mapping.  One wavefront per element (nq >= 6) or per group of
// floor(64 / nq^2) elements (nq <= 5).  A lane owns the (j, k) column of its
// element and keeps the whole i-line of u, of rr and of the result in
// registers, so the i-direction contractions (qr and the rr transpose) are
// register-only with D(i, n) uniform across the wavefront.  The element's u
// is staged once in LDS for the j- and k-derivatives; rs and rt of one
// i-layer at a time go through a small double-buffered LDS plane for the
// transposed j- and k-contractions.  All LDS traffic stays inside one
// wavefront, so there is no workgroup barrier; layers are processed in i
// order so that only one layer of G (6 values per lane) is live at a time.
//
// HBM traffic per element: u once (s nq^3 B, s = sizeof storage), v once
// (8 nq^3 B) and G once (6 s nq^3 B), all with unit-stride lanes: u/v planes
// are nq^2 contiguous values, G(i, j, f, :) rows are nq contiguous values.
#include <type_traits>

#include "bkx_device.hpp"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

// out[c] = sum_n M[c * NQ + n] in[n] (row c of M times in), n ascending,
// one row of scalar operands per output.
template <int NQ, typename P>
__device__ __forceinline__ void contract_rows_f64(P M, const double (&in)[NQ], double (&out)[NQ])
{
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            t += M[c * NQ + n] * in[n];
        }
        out[c] = t;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// out[c] = sum_n M[n * NQ + c] in[n] (column c of M times in), n ascending,
// one row of scalar operands per step.
template <int NQ, typename P>
__device__ __forceinline__ void contract_cols_f64(P M, const double (&in)[NQ], double (&out)[NQ])
{
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        out[c] = 0.0;
    }
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            out[c] += M[n * NQ + c] * in[n];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Staged reads of G: unit-stride loads of 8 (MODE 1) or 16 bytes (MODE 2)
// per lane over a contiguous run of G, written to the LDS and read back per
// lane.  MODE 2 loads are non-temporal (G is read once; on MI355X the
// streaming hint lifts BK5 p = 7 from 87 to 101 GDoF/s in FP64) and are
// used only when every element's G and every layer of it are whole 16-byte
// units (otherwise MODE 1).
template <typename TS, int NQ, int MODE>
struct GStage {
    static constexpr int layer = 6 * NQ * NQ;   // values of one layer of G
    static constexpr bool wide = MODE == 2 && (layer * sizeof(TS)) % 16 == 0;
    static constexpr int bytes = wide ? 16 : 8;
    static constexpr int tpe = bytes / int(sizeof(TS));   // values per unit
    using unit = std::conditional_t<wide, std::conditional_t<sizeof(TS) == 4, f32x4, f64x2>,
                                    std::conditional_t<sizeof(TS) == 4, f32x2, double>>;
    static_assert((layer * sizeof(TS)) % 8 == 0, "a layer is a whole number of 8-byte units");
    __device__ static __forceinline__ unit load(const TS* p)
    {
        const unit* q = reinterpret_cast<const unit*>(p);
        if constexpr (wide) {
            return __builtin_nontemporal_load(q);
        } else {
            return *q;
        }
    }
};

template <int NQ>
struct Bk5Shape {
    static constexpr int nq2 = NQ * NQ;
    static constexpr int nq3 = nq2 * NQ;
    static constexpr int epw = kWave / nq2;       // elements per wavefront
    static constexpr int row = NQ | 1;            // odd row stride in 8-byte words (bank spread)
    static constexpr int plane = NQ * row;        // one padded (j, k) plane
    static constexpr int box = NQ * plane;        // one padded element box
    static_assert(epw >= 1, "one element must fit a wavefront");
};

template <int NQ, int WAVES, typename TS, bool NT>
__global__ __launch_bounds__(kWave * WAVES) void bk5_kernel(size_t nelmt, const double* __restrict__ D,
                                                           const TS* __restrict__ G, const TS* __restrict__ u,
                                                           double* __restrict__ v)
{
    using S = Bk5Shape<NQ>;
    constexpr int nq2 = S::nq2;
    constexpr int nq3 = S::nq3;
    constexpr int epw = S::epw;

    __shared__ double s_u[WAVES][epw][S::box];
    __shared__ double s_x[WAVES][2][2][epw][S::plane];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = lane / nq2;
    const bool in_elem = slot < epw;           // idle lanes: nq^2 does not divide 64
    const int sl = in_elem ? slot : epw - 1;   // idle lanes read (never write) a real slot
    const int jk = lane - slot * nq2;
    const int j = in_elem ? jk / NQ : 0;
    const int k = in_elem ? jk - j * NQ : 0;
    const size_t e = (size_t(blockIdx.x) * WAVES + wave) * epw + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    const int lx = j * S::row + k;             // this lane's slot in a padded plane

    // Per-lane rows and columns of D for the j and k contractions.
    double Dj[NQ], Dk[NQ], DTj[NQ], DTk[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
        Dj[n] = D[j * NQ + n];
        Dk[n] = D[k * NQ + n];
        DTj[n] = D[n * NQ + j];
        DTk[n] = D[n * NQ + k];
    }

    // u column along i -> registers and the padded LDS box.
    const TS* ue = u + ee * nq3 + jk;
    double* su = s_u[wave][sl];
    // unconditional loads (ee is a real element), masked after the
    // conversion: converting inside the condition waits for each load
    TS uraw[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        uraw[i] = load<NT>(ue + i * nq2);
    }
    double ur[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        ur[i] = active ? double(uraw[i]) : 0.0;
    }
    if (in_elem) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            su[i * S::plane + lx] = ur[i];
        }
    }
    wave_sync();

    // G(e, i, j, f, k): the lane's six factors of layer i are at
    // ge + i * 6 nq^2 + f * nq.  One layer is prefetched ahead, kept in the
    // storage type until it is used (a conversion at the load would wait for
    // it there).
    const TS* ge = G + ee * 6 * nq3 + j * 6 * NQ + k;
    TS gn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        gn[f] = active ? load<NT>(ge + f * NQ) : TS(0);
    }
    // acc(m) = v(m, j, k): the i-direction transpose sum_i D(i, m) rr(i) is
    // accumulated as each layer's rr appears (i ascending), the j/k
    // transposes of layer m are added at layer m.
    double acc[NQ];
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
        acc[m] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        double g[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            g[f] = double(gn[f]);
        }
        if (i + 1 < NQ) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                gn[f] = active ? load<NT>(ge + (i + 1) * 6 * nq2 + f * NQ) : TS(0);
            }
        }
        double qr = 0.0, qs = 0.0, qt = 0.0;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qr += D[i * NQ + n] * ur[n];
        }
        const double* sp = su + i * S::plane;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qs += Dj[n] * sp[n * S::row + k];
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            qt += Dk[n] * sp[j * S::row + n];
        }
        const double rr = g[0] * qr + g[1] * qs + g[2] * qt;
        const double rs = g[1] * qr + g[3] * qs + g[4] * qt;
        const double rt = g[2] * qr + g[4] * qs + g[5] * qt;

        double* xs = s_x[wave][i & 1][0][sl];
        double* xt = s_x[wave][i & 1][1][sl];
        if (in_elem) {
            xs[lx] = rs;
            xt[lx] = rt;
        }
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            acc[m] += D[i * NQ + m] * rr;
        }
        wave_sync();
        double a = acc[i];
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            a += DTj[n] * xs[n * S::row + k];
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            a += DTk[n] * xt[j * S::row + n];
        }
        acc[i] = a;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            pin(acc[m]);
        }
    }

    if (active) {
        double* ve = v + e * nq3 + jk;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            st_stream(ve + i * nq2, acc[i]);
        }
    }
}

// Transposing mapping.  The same wavefront per element, but every
// contraction runs along the lane's register column with D(., .) uniform
// across the wavefront (scalar operands); the LDS is used only to move the
// element between three layouts, lane (j,k) / (i,k) / (i,j) holding the
// column along i / j / k:
//   u (L_i) -> LDS;  qr in L_i;  u read as L_j -> qs;  as L_k -> qt
//   qs, qt -> LDS -> read in L_i;  fluxes with G (L_i), rr kept
//   rs, rt -> LDS -> read in L_j / L_k -> transposed j / k contractions
//   -> LDS -> read in L_i and added to the transposed i contraction of rr.
// Per element that is 7 nq^3 LDS writes and 8 nq^3 reads, against
// 3 nq^3 writes and 4 nq^4 reads when the j and k contractions read their
// operands from LDS (the slab mapping above).  MINW is the occupancy
// (wavefronts per SIMD) the register allocation is held to.
// GATHER: u is an L-vector read through the element restriction, u(e, x) =
// u[uidx[e nq^3 + x]] (the gather of restrict.hip fused into the loads).
template <int NQ, int WAVES, typename TS, bool NT, int MINW, int GSTAGE, int PD, bool SEQ, bool GATHER>
__global__ __launch_bounds__(kWave * WAVES) __attribute__((amdgpu_waves_per_eu(MINW))) void bk5t_kernel(size_t nelmt, const double* __restrict__ D,
                                                            const TS* __restrict__ G, const TS* __restrict__ u,
                                                            const int32_t* __restrict__ uidx,
                                                            double* __restrict__ v)
{
    using S = Bk5Shape<NQ>;
    constexpr int nq2 = S::nq2;
    constexpr int nq3 = S::nq3;
    constexpr int epw = S::epw;

    __shared__ double s_a[WAVES][epw][S::box];
    __shared__ double s_b[WAVES][epw][S::box];
    // GSTAGE: one layer of G of every element of the wavefront, loaded with
    // unit-stride 8-byte lanes and read back per lane (G(e, i, :, :, :) is
    // 6 nq^2 contiguous values per element).
    using GS = GStage<TS, NQ, GSTAGE>;
    using stage_t = typename GS::unit;
    constexpr int tpe = GS::tpe;                             // values per unit
    constexpr int layer = 6 * nq2;                           // values per layer
    constexpr int units = epw * layer / tpe;                 // units per wavefront layer
    constexpr int upl = (units + kWave - 1) / kWave;         // units per lane
    __shared__ alignas(16) TS s_g[GSTAGE ? WAVES : 1][GSTAGE ? epw * layer : 1];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = lane / nq2;
    const bool in_elem = slot < epw;
    const int sl = in_elem ? slot : epw - 1;
    const int ab = lane - slot * nq2;
    const int a = in_elem ? ab / NQ : 0;
    const int b = in_elem ? ab - a * NQ : 0;
    const size_t e = (size_t(blockIdx.x) * WAVES + wave) * epw + slot;
    const bool active = in_elem && e < nelmt;
    const size_t ee = active ? e : 0;
    double* sa = s_a[wave][sl];
    double* sb = s_b[wave][sl];
    // padded box address of (i, j, k)
    auto X = [](int i, int j, int k) { return i * S::plane + j * S::row + k; };

    // u in L_i: lane (j, k) = (a, b), column i
    const TS* ue = u + ee * nq3 + a * NQ + b;
    // unconditional loads (ee is a real element), masked after the
    // conversion: converting inside the condition waits for each load
    TS xraw[NQ];
    if constexpr (GATHER) {
        // the indices are read once: streamed
        const int32_t* ie = uidx + ee * nq3 + a * NQ + b;
        int32_t g[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            g[c] = __builtin_nontemporal_load(ie + c * nq2);
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            xraw[c] = u[g[c]];
        }
    } else {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            xraw[c] = load<NT>(ue + c * nq2);
        }
    }
    double x[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        x[c] = active ? double(xraw[c]) : 0.0;
    }
    // first layer of G
    const TS* ge = G + ee * 6 * nq3 + a * 6 * NQ + b;
    TS gn[PD][6];            // PD layers of G in flight (per-lane reads)
    stage_t gs[PD][upl];     // PD layers of G in flight (staged reads)
    // the wavefront's elements are consecutive: its G is one contiguous run
    const size_t e0 = (size_t(blockIdx.x) * WAVES + wave) * epw;
    const int nact = int(nelmt - e0 < size_t(epw) ? nelmt - e0 : size_t(epw));
    auto stage_load = [&](int c, stage_t (&dst)[upl]) {
#pragma unroll
        for (int t = 0; t < upl; ++t) {
            const int x = t * kWave + lane;                  // unit within the wavefront's layer
            const int el = x * tpe / layer;                  // element slot of the unit
            const int off = x * tpe - el * layer;            // value offset within that layer
            const bool ok = x < units && el < nact;
            const TS* src = G + (e0 + (ok ? el : 0)) * 6 * nq3 + c * layer + (ok ? off : 0);
            dst[t] = ok ? GS::load(src) : stage_t{};
        }
    };
    if constexpr (GSTAGE) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            if (d < NQ) {
                stage_load(d, gs[d]);
            }
        }
    } else {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            if (d < NQ) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    gn[d][f] = active ? load<NT>(ge + d * 6 * nq2 + f * NQ) : TS(0);
                }
            }
        }
    }
    if (in_elem) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            sa[X(c, a, b)] = x[c];
        }
    }
    wave_sync();
    double qr[NQ];
    if constexpr (SEQ) {
        // one contraction at a time, each result stored before the next
        // column is read: qs in L_j (lane (i, k) = (a, b)) -> box B, then
        // qt in L_k (lane (i, j) = (a, b)) -> box A, then qr in L_i.  In L_j
        // and L_k a lane reads and writes only its own column of the box, so
        // qt may overwrite u in box A once this lane's column has been read
        // (the box's other columns are read by their own lanes), and u's
        // j-columns were all read before (LDS operations of one wavefront
        // complete in order).  At most two columns of eight values are live.
        {
            const auto Ds = launder<2>(D);
            double y[NQ], w[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                y[n] = sa[X(a, n, b)];
            }
            contract_rows_f64<NQ>(Ds, y, w);
            if (in_elem) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    sb[X(a, c, b)] = w[c];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            const auto Ds = launder<5>(D);
            double z[NQ], w[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                z[n] = sa[X(a, b, n)];
            }
            contract_rows_f64<NQ>(Ds, z, w);
            if (in_elem) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    sa[X(a, b, c)] = w[c];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        contract_rows_f64<NQ>(launder<6>(D), x, qr);
    } else {
    // one pass over the rows of D: qr (L_i, u in registers), qs in L_j
    // (lane (i, k) = (a, b)) and qt in L_k (lane (i, j) = (a, b))
    double qs[NQ], qt[NQ];
    {
        const auto Ds = launder<2>(D);
        double y[NQ], z[NQ];
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            y[n] = sa[X(a, n, b)];
            z[n] = sa[X(a, b, n)];
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            double tr = 0.0, ts = 0.0, tt = 0.0;
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                tr += Ds[c * NQ + n] * x[n];
                ts += Ds[c * NQ + n] * y[n];
                tt += Ds[c * NQ + n] * z[n];
            }
            qr[c] = tr;
            qs[c] = ts;
            qt[c] = tt;
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    wave_sync();
    if (in_elem) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            sb[X(a, c, b)] = qs[c];
            sa[X(a, b, c)] = qt[c];
        }
    }
    }
    wave_sync();
    // fluxes in L_i, one layer of G at a time (prefetched one ahead); rs and
    // rt overwrite the qs and qt this lane has just read
    double rr[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        double g[6];
        if constexpr (GSTAGE) {
            TS* sg = s_g[wave];
#pragma unroll
            for (int t = 0; t < upl; ++t) {
                const int x = t * kWave + lane;
                if (x < units) {
                    *reinterpret_cast<stage_t*>(sg + x * tpe) = gs[c % PD][t];
                }
            }
            if (c + PD < NQ) {
                stage_load(c + PD, gs[c % PD]);
            }
            wave_sync();
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = double(sg[sl * layer + (a * 6 + f) * NQ + b]);
            }
            wave_sync();
        } else {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = double(gn[c % PD][f]);
            }
            if (c + PD < NQ) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    gn[c % PD][f] = active ? load<NT>(ge + (c + PD) * 6 * nq2 + f * NQ) : TS(0);
                }
            }
        }
        const double s_ = sb[X(c, a, b)];
        const double t_ = sa[X(c, a, b)];
        const double r_ = qr[c];
        rr[c] = g[0] * r_ + g[1] * s_ + g[2] * t_;
        const double rs = g[1] * r_ + g[3] * s_ + g[4] * t_;
        const double rt = g[2] * r_ + g[4] * s_ + g[5] * t_;
        if (in_elem) {
            sb[X(c, a, b)] = rs;
            sa[X(c, a, b)] = rt;
        }
    }
    wave_sync();
    if constexpr (SEQ) {
        // transposed j contraction in L_j, stored in place; then the k
        // contraction in L_k (each lane reads and writes its own column)
        {
            const auto Dt = launder<3>(D);
            double y[NQ], w[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                y[n] = sb[X(a, n, b)];
            }
            contract_cols_f64<NQ>(Dt, y, w);
            if (in_elem) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    sb[X(a, c, b)] = w[c];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            const auto Dt = launder<7>(D);
            double z[NQ], w[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                z[n] = sa[X(a, b, n)];
            }
            contract_cols_f64<NQ>(Dt, z, w);
            if (in_elem) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    sa[X(a, b, c)] = w[c];
                }
            }
        }
    } else {
    // transposed j contraction in L_j (lane (i, k)), k contraction in L_k (lane (i, j))
    // (row n of D times y[n], accumulated over n ascending: one contiguous row
    // of scalar operands per step)
    double wj[NQ], wk[NQ];
    {
        const auto Dt = launder<3>(D);
        double y[NQ], z[NQ];
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            y[n] = sb[X(a, n, b)];
            z[n] = sa[X(a, b, n)];
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            wj[c] = 0.0;
            wk[c] = 0.0;
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                wj[c] += Dt[n * NQ + c] * y[n];
                wk[c] += Dt[n * NQ + c] * z[n];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    wave_sync();
    if (in_elem) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            sb[X(a, c, b)] = wj[c];
            sa[X(a, b, c)] = wk[c];
        }
    }
    }
    wave_sync();
    if (active) {
        const auto Df = launder<4>(D);
        double t[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            t[c] = 0.0;
        }
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                t[c] += Df[n * NQ + c] * rr[n];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        double* ve = v + e * nq3 + a * NQ + b;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            st_stream(ve + c * nq2, t[c] + sb[X(c, a, b)] + sa[X(c, a, b)]);
        }
    }
}

// Pipelined transposing mapping (nq >= 6: one element per wavefront).  The
// phases of bk5t_kernel, but each wavefront is resident for the whole launch
// and walks the elements e = w, w + W, w + 2W, ... (W = the number of
// wavefronts).  Its loads run ahead of its arithmetic across element
// boundaries:
//   - G is read through a ring of PD layers (staged by unit-stride 8-byte
//     loads): when layer c of element e has been used, the slot is refilled
//     with layer c + PD of e, or with layer c + PD - nq of the next element;
//   - the next element's u column is loaded as soon as the current u has
//     been used (after the derivatives).
// So the next element's inputs are in flight during the fluxes, the
// transposed contractions and the store of the current one, without extra
// registers for G.  The loads are unconditional: past the wavefront's last
// element they re-read the current one (never used), so that the number of
// memory operations in flight is the same on every path and the compiler's
// wait counts stay exact across the loop (a load on one path only makes it
// wait for every load and store in flight at the top of the loop).  Every
// wavefront exits its loop after its last element.
template <int NQ, typename TS, int PD, int MINW, int GSTAGE>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MINW))) void bk5p_kernel(
    size_t nelmt, size_t stride, const double* __restrict__ D, const TS* __restrict__ G, const TS* __restrict__ u,
    double* __restrict__ v)
{
    using S = Bk5Shape<NQ>;
    static_assert(S::epw == 1, "one element per wavefront");
    static_assert(NQ % PD == 0, "the G ring holds a whole number of rings per element");
    constexpr int nq2 = S::nq2;
    constexpr int nq3 = S::nq3;
    __shared__ double sa[S::box];
    __shared__ double sb[S::box];
    using GS = GStage<TS, NQ, GSTAGE>;
    using stage_t = typename GS::unit;
    constexpr int tpe = GS::tpe;                             // values per unit
    constexpr int layer = 6 * nq2;                           // values per layer of G
    constexpr int units = layer / tpe;                       // units per layer
    constexpr int upl = (units + kWave - 1) / kWave;         // units per lane
    __shared__ alignas(16) TS sg[layer];

    const int lane = threadIdx.x;
    const bool in_elem = nq2 >= kWave || lane < nq2;   // nq^2 lanes hold the element
    const int a = in_elem ? lane / NQ : 0;
    const int b = in_elem ? lane - a * NQ : 0;
    auto X = [](int i, int j, int k) { return i * S::plane + j * S::row + k; };

    size_t e = blockIdx.x;
    if (e >= nelmt) {
        return;
    }
    // layer c of element el into dst (unit-stride 8-byte lanes over the layer)
    auto stage_load = [&](size_t el, int c, stage_t (&dst)[upl]) {
        const TS* src = G + el * 6 * nq3 + c * layer;
#pragma unroll
        for (int t = 0; t < upl; ++t) {
            const int x = t * kWave + lane;
            if constexpr (units % kWave == 0) {
                dst[t] = GS::load(src + x * tpe);
            } else {
                dst[t] = GS::load(src + (x < units ? x : 0) * tpe);
            }
        }
    };
    auto u_load = [&](size_t el, TS (&dst)[NQ]) {
        const TS* ue = u + el * nq3 + (in_elem ? lane : 0);
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            dst[c] = ue[c * nq2];
        }
    };

    // prologue: the first element's u and first PD layers of G
    TS xin[NQ];
    u_load(e, xin);
    stage_t gs[PD][upl];
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        stage_load(e, d, gs[d]);
    }

    for (;;) {
        const bool has_next = e + stride < nelmt;      // wavefront-uniform
        const size_t en = has_next ? e + stride : e;   // loads past the end re-read e
        // u in L_i -> registers and box A
        double x[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            x[c] = double(xin[c]);
        }
        if (in_elem) {
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                sa[X(c, a, b)] = x[c];
            }
        }
        wave_sync();
        // one pass over the rows of D: qr (L_i, u in registers), qs in L_j
        // (lane (i, k) = (a, b)) and qt in L_k (lane (i, j) = (a, b))
        double qr[NQ], qs[NQ], qt[NQ];
        {
            const auto Ds = launder<2>(D);
            double y[NQ], z[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                y[n] = sa[X(a, n, b)];
                z[n] = sa[X(a, b, n)];
            }
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                double tr = 0.0, ts = 0.0, tt = 0.0;
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    tr += Ds[c * NQ + n] * x[n];
                    ts += Ds[c * NQ + n] * y[n];
                    tt += Ds[c * NQ + n] * z[n];
                }
                qr[c] = tr;
                qs[c] = ts;
                qt[c] = tt;
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the next element's u, in flight from here on
        u_load(en, xin);
        wave_sync();
        if (in_elem) {
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                sb[X(a, c, b)] = qs[c];
                sa[X(a, b, c)] = qt[c];
            }
        }
        wave_sync();
        // fluxes in L_i, one staged layer of G at a time; each ring slot is
        // refilled as soon as it has been copied to the LDS
        double rr[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
#pragma unroll
            for (int t = 0; t < upl; ++t) {
                const int xu = t * kWave + lane;
                if (units % kWave == 0 || xu < units) {
                    *reinterpret_cast<stage_t*>(sg + xu * tpe) = gs[c % PD][t];
                }
            }
            if (c + PD < NQ) {
                stage_load(e, c + PD, gs[c % PD]);
            } else {
                stage_load(en, c + PD - NQ, gs[c % PD]);
            }
            wave_sync();
            double g[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                g[f] = double(sg[(a * 6 + f) * NQ + b]);
            }
            wave_sync();
            const double s_ = sb[X(c, a, b)];
            const double t_ = sa[X(c, a, b)];
            const double r_ = qr[c];
            rr[c] = g[0] * r_ + g[1] * s_ + g[2] * t_;
            const double rs = g[1] * r_ + g[3] * s_ + g[4] * t_;
            const double rt = g[2] * r_ + g[4] * s_ + g[5] * t_;
            if (in_elem) {
                sb[X(c, a, b)] = rs;
                sa[X(c, a, b)] = rt;
            }
        }
        wave_sync();
        // transposed j contraction in L_j (lane (i, k)), k contraction in
        // L_k (lane (i, j)): row n of D times y[n], n ascending
        double wj[NQ], wk[NQ];
        {
            const auto Dt = launder<3>(D);
            double y[NQ], z[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                y[n] = sb[X(a, n, b)];
                z[n] = sa[X(a, b, n)];
            }
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                wj[c] = 0.0;
                wk[c] = 0.0;
            }
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    wj[c] += Dt[n * NQ + c] * y[n];
                    wk[c] += Dt[n * NQ + c] * z[n];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wave_sync();
        if (in_elem) {
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                sb[X(a, c, b)] = wj[c];
                sa[X(a, b, c)] = wk[c];
            }
        }
        wave_sync();
        // transposed i contraction of rr plus the j and k parts, stored
        {
            double t[NQ];
            contract_cols_f64<NQ>(launder<4>(D), rr, t);
            if (in_elem) {
                double* ve = v + e * nq3 + lane;
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    st_stream(ve + c * nq2, t[c] + sb[X(c, a, b)] + sa[X(c, a, b)]);
                }
            }
        }
        if (!has_next) {
            break;
        }
        e = en;
        // the boxes are rewritten next: this lane's reads above come first
        wave_sync();
    }
}

// Launches the pipelined kernel with at most as many wavefronts as are
// resident at once (each then runs from the start of the launch).
template <int NQ, typename TS, int PD, int MINW, int GSTAGE = 1>
int launch_bk5p_v(size_t nelmt, const double* D, const TS* G, const TS* u, double* v, hipStream_t s)
{
    const unsigned resident = resident_blocks<bk5p_kernel<NQ, TS, PD, MINW, GSTAGE>>(kWave);
    if (resident == 0) {
        return BKX_ERROR_LAUNCH;
    }
    const size_t grid = nelmt < resident ? nelmt : resident;
    hipLaunchKernelGGL((bk5p_kernel<NQ, TS, PD, MINW, GSTAGE>), dim3(unsigned(grid)), dim3(kWave), 0, s, nelmt, grid, D, G, u, v);
    return launch_status();
}

template <int NQ, int WAVES, typename TS, bool NT, int MINW = 1, int GSTAGE = 0, int PD = 1, bool SEQ = false,
          bool GATHER = false>
int launch_bk5t_v(size_t nelmt, const double* D, const TS* G, const TS* u, double* v, hipStream_t s,
                  const int32_t* uidx = nullptr)
{
    constexpr size_t per_block = size_t(WAVES) * Bk5Shape<NQ>::epw;
    const unsigned grid = grid_for(nelmt, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk5t_kernel<NQ, WAVES, TS, NT, MINW, GSTAGE, PD, SEQ, GATHER>), dim3(grid), dim3(kWave * WAVES), 0, s, nelmt, D, G, u, uidx, v);
    return launch_status();
}

template <int NQ, int WAVES, typename TS, bool NT>
int launch_bk5_v(size_t nelmt, const double* D, const TS* G, const TS* u, double* v, hipStream_t s)
{
    constexpr size_t per_block = size_t(WAVES) * Bk5Shape<NQ>::epw;
    const unsigned grid = grid_for(nelmt, per_block);
    if (grid == 0) {
        return BKX_ERROR_SIZE;
    }
    hipLaunchKernelGGL((bk5_kernel<NQ, WAVES, TS, NT>), dim3(grid), dim3(kWave * WAVES), 0, s, nelmt, D, G, u, v);
    return launch_status();
}

// Launch variants (the autotuner's search space): the slab mapping (0-3) or
// the transposing mapping (4-21), wavefronts per workgroup, ordinary or
// non-temporal loads, an occupancy target, G read per lane or staged
// through LDS by unit-stride loads (9-18), and 1 to 8 layers of G in
// flight; one contraction at a time (22-29); the pipelined kernel (30-33);
// and G staged by 16-byte non-temporal loads (34-41).  Per-lane G reads
// use ordinary loads: the non-temporal hint costs them 10% (p = 7) to 40%
// (p = 3), as a G row of one factor is 64 contiguous bytes and the next
// factor's row, read by the next instruction, fills the rest of the same
// 128-byte line.  Unit-stride staged reads touch every line once, and there
// the hint is a gain (variants 34-41).
constexpr int kBk5Variants = 42;

template <int NQ, typename TS>
int launch_bk5(int variant, size_t nelmt, const double* D, const TS* G, const TS* u, double* v, hipStream_t s)
{
    // the staged variants read G in 8-byte units: they need an 8-byte-aligned
    // G; the 16-byte units (34-41) a 16-byte-aligned one
    if (variant >= 9 && reinterpret_cast<uintptr_t>(G) % 8 != 0) {
        variant = 4;
    } else if (variant >= 34 && reinterpret_cast<uintptr_t>(G) % 16 != 0) {
        variant = 17;
    }
    switch (variant) {
    case 0:
        return launch_bk5_v<NQ, 4, TS, false>(nelmt, D, G, u, v, s);
    case 1:
        return launch_bk5_v<NQ, 1, TS, false>(nelmt, D, G, u, v, s);
    case 2:
        return launch_bk5_v<NQ, 2, TS, false>(nelmt, D, G, u, v, s);
    case 3:
        return launch_bk5_v<NQ, 4, TS, true>(nelmt, D, G, u, v, s);
    case 4:
        return launch_bk5t_v<NQ, 4, TS, false>(nelmt, D, G, u, v, s);
    case 5:
        return launch_bk5t_v<NQ, 1, TS, false>(nelmt, D, G, u, v, s);
    case 6:
        return launch_bk5t_v<NQ, 2, TS, false>(nelmt, D, G, u, v, s);
    case 7:
        return launch_bk5t_v<NQ, 1, TS, true>(nelmt, D, G, u, v, s);
    case 8:
        return launch_bk5t_v<NQ, 4, TS, false, 3>(nelmt, D, G, u, v, s);
    case 9:
        return launch_bk5t_v<NQ, 2, TS, false, 1, true>(nelmt, D, G, u, v, s);
    case 10:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true>(nelmt, D, G, u, v, s);
    case 11:
        return launch_bk5t_v<NQ, 4, TS, false, 1, true>(nelmt, D, G, u, v, s);
    case 12:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 2>(nelmt, D, G, u, v, s);
    case 13:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 3>(nelmt, D, G, u, v, s);
    case 14:
        return launch_bk5t_v<NQ, 2, TS, false, 1, true, 2>(nelmt, D, G, u, v, s);
    case 15:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 4>(nelmt, D, G, u, v, s);
    case 16:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 6>(nelmt, D, G, u, v, s);
    case 17:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 8>(nelmt, D, G, u, v, s);
    case 18:
        return launch_bk5t_v<NQ, 2, TS, false, 1, true, 4>(nelmt, D, G, u, v, s);
    case 19:
        return launch_bk5t_v<NQ, 1, TS, false, 1, false, 2>(nelmt, D, G, u, v, s);
    case 20:
        return launch_bk5t_v<NQ, 1, TS, false, 1, false, 3>(nelmt, D, G, u, v, s);
    case 21:
        return launch_bk5t_v<NQ, 2, TS, false, 1, false, 2>(nelmt, D, G, u, v, s);
    // one contraction at a time (fewer registers live): 22-29
    case 22:
        return launch_bk5t_v<NQ, 1, TS, false, 4, true, 8, true>(nelmt, D, G, u, v, s);
    case 23:
        return launch_bk5t_v<NQ, 1, TS, false, 4, true, 4, true>(nelmt, D, G, u, v, s);
    case 24:
        return launch_bk5t_v<NQ, 1, TS, false, 4, false, 2, true>(nelmt, D, G, u, v, s);
    case 25:
        return launch_bk5t_v<NQ, 4, TS, false, 4, false, 1, true>(nelmt, D, G, u, v, s);
    case 26:
        return launch_bk5t_v<NQ, 1, TS, false, 4, true, 2, true>(nelmt, D, G, u, v, s);
    case 27:
        // four waves per SIMD do not fit the FP64 registers at nq = 8: three there
        return launch_bk5t_v<NQ, 2, TS, false, (NQ == 8 && sizeof(TS) == 8) ? 3 : 4, true, 4, true>(nelmt, D, G, u,
                                                                                                     v, s);
    case 28:
        return launch_bk5t_v<NQ, 1, TS, false, 1, true, 8, true>(nelmt, D, G, u, v, s);
    case 29:
        return launch_bk5t_v<NQ, 4, TS, false, 1, false, 1, true>(nelmt, D, G, u, v, s);
    // pipelined over elements (nq >= 6; at lower orders the staged
    // transposing kernel): 30-33
    case 30:
    case 31:
    case 32:
    case 33:
        if constexpr (Bk5Shape<NQ>::epw == 1) {
            constexpr int half = NQ % 2 == 0 ? NQ / 2 : NQ;
            constexpr int two = NQ % 2 == 0 ? 2 : 1;
            switch (variant) {
            case 30:
                return launch_bk5p_v<NQ, TS, NQ, 1>(nelmt, D, G, u, v, s);
            case 31:
                return launch_bk5p_v<NQ, TS, half, 1>(nelmt, D, G, u, v, s);
            case 32:
                return launch_bk5p_v<NQ, TS, two, 1>(nelmt, D, G, u, v, s);
            default:
                return launch_bk5p_v<NQ, TS, NQ, 3>(nelmt, D, G, u, v, s);
            }
        } else {
            return launch_bk5t_v<NQ, 1, TS, false, 1, true, 2>(nelmt, D, G, u, v, s);
        }
    // G staged by 16-byte non-temporal loads (8-byte units where a layer
    // is not a whole number of 16-byte units): 34-41
    case 34:
        return launch_bk5t_v<NQ, 1, TS, false, 1, 2, 8>(nelmt, D, G, u, v, s);
    case 35:
        return launch_bk5t_v<NQ, 1, TS, false, 1, 2, 4>(nelmt, D, G, u, v, s);
    case 36:
        return launch_bk5t_v<NQ, 1, TS, false, 1, 2, 2>(nelmt, D, G, u, v, s);
    case 37:
        return launch_bk5t_v<NQ, 1, TS, true, 1, 2, 8>(nelmt, D, G, u, v, s);
    case 38:
        return launch_bk5t_v<NQ, 2, TS, false, 1, 2, 4>(nelmt, D, G, u, v, s);
    case 39:
    case 40:
    case 41:
        if constexpr (Bk5Shape<NQ>::epw == 1) {
            constexpr int half = NQ % 2 == 0 ? NQ / 2 : NQ;
            constexpr int two = NQ % 2 == 0 ? 2 : 1;
            switch (variant) {
            case 39:
                return launch_bk5p_v<NQ, TS, NQ, 1, 2>(nelmt, D, G, u, v, s);
            case 40:
                return launch_bk5p_v<NQ, TS, half, 1, 2>(nelmt, D, G, u, v, s);
            default:
                return launch_bk5p_v<NQ, TS, two, 1, 2>(nelmt, D, G, u, v, s);
            }
        } else {
            return launch_bk5t_v<NQ, 1, TS, false, 1, 2, 2>(nelmt, D, G, u, v, s);
        }
    default:
        return BKX_ERROR_ARG;
    }
}

// BK5 with the gather fused into the loads of u (GATHER): the launch
// shapes of the variants that are defaults at some order (18, 27, 34-36,
// 38); any other variant runs as 36.  Variants 34-38 need a 16-byte-aligned
// G, else 17's shape.
template <int NQ>
int launch_bk5_gather(int variant, size_t nelmt, const double* D, const double* G, const int32_t* uidx,
                      const double* uL, double* v, hipStream_t s)
{
    const bool g16 = reinterpret_cast<uintptr_t>(G) % 16 == 0;
    if (!g16 && variant >= 34) {
        return launch_bk5t_v<NQ, 1, double, false, 1, true, 8, false, true>(nelmt, D, G, uL, v, s, uidx);
    }
    switch (variant) {
    case 18:
        return launch_bk5t_v<NQ, 2, double, false, 1, true, 4, false, true>(nelmt, D, G, uL, v, s, uidx);
    case 27:
        return launch_bk5t_v<NQ, 2, double, false, NQ == 8 ? 3 : 4, true, 4, true, true>(nelmt, D, G, uL, v, s,
                                                                                         uidx);
    case 34:
        return launch_bk5t_v<NQ, 1, double, false, 1, 2, 8, false, true>(nelmt, D, G, uL, v, s, uidx);
    case 35:
        return launch_bk5t_v<NQ, 1, double, false, 1, 2, 4, false, true>(nelmt, D, G, uL, v, s, uidx);
    case 38:
        return launch_bk5t_v<NQ, 2, double, false, 1, 2, 4, false, true>(nelmt, D, G, uL, v, s, uidx);
    default:
        if (!g16) {
            return launch_bk5t_v<NQ, 1, double, false, 1, true, 8, false, true>(nelmt, D, G, uL, v, s, uidx);
        }
        return launch_bk5t_v<NQ, 1, double, false, 1, 2, 2, false, true>(nelmt, D, G, uL, v, s, uidx);
    }
}

}  // namespace

int bk5_variants() { return kBk5Variants; }

int bk5_apply_gather_f64(int variant, int p, size_t nelmt, const double* D, const double* G, const int32_t* uidx,
                         const double* uL, double* v, hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) {
        return launch_bk5_gather<decltype(P)::value + 1>(variant, nelmt, D, G, uidx, uL, v, s);
    });
}

int bk5_apply_f64(int variant, int p, size_t nelmt, const double* D, const double* G, const double* u, double* v,
                  hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) {
        return launch_bk5<decltype(P)::value + 1>(variant, nelmt, D, G, u, v, s);
    });
}

int bk5_apply_f32w(int variant, int p, size_t nelmt, const double* D, const float* G, const float* u, double* v,
                   hipStream_t s)
{
    return with_order<1, 7>(p, [&](auto P) {
        return launch_bk5<decltype(P)::value + 1>(variant, nelmt, D, G, u, v, s);
    });
}

}  // namespace bkx
