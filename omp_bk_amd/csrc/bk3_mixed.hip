// bk3_mixed.hip -- BK3 over a batch of elements of mixed order (p = 1..8)
// in FP64 on gfx950, as one persistent launch or as one launch per order.
//
// The reference runs one order per binary (include/bk_common.h:143-178); a
// batch of elements of different orders has no counterpart there.  Layout of
// a mixed batch, in the caller's element order e = 0 .. nelmt-1 with order
// p_e (n_e = p_e + 1 modes, q_e = p_e + 2 points per direction):
//   u, v:  element e's n_e^3 values at uoff[e] = sum_{e' < e} n_e'^3
//   G:     element e's 6 q_e^3 factors (factor-major, as BK3) at
//          goff[e] = sum_{e' < e} 6 q_e'^3
//   bases: for p = 1..8 in turn, B_p (n x q) then D_p (q x q)
// The plan (host side) sorts the elements by order and keeps each element's
// offsets at its position in the sorted list.  The single launch cuts each
// order's part of the list into tiles of rounds x epb elements (epb =
// elements per workgroup of that order's kernel body, chosen per order by
// the tuner), ordered by the memory position of their elements.  The single
// launch runs one workgroup per tile, or (persistent) a resident grid whose
// workgroups take tiles from a queue in the plan.  The per-order launches
// run each order's part of the list with a kernel compiled for that order
// alone.
// Both run the transposing BK3 body of bk13t_body.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <new>
#include <numeric>
#include <vector>

#include "bk13t_body.hpp"
#include "bkx.h"
#include "bkx_launch.hpp"

namespace bkx {
namespace {

struct MixedTile {
    int32_t p;       // order
    int32_t div;     // elements-per-workgroup divisor of that order's body (1, 2 or 4)
    int32_t first;   // first index into the order-sorted element list
    int32_t count;   // elements in the tile
};

constexpr int kMixedThreads = 256;
// every order's body: two LDS boxes per element, and three rows of G loaded
// ahead (the fastest single-order BK3 form at most orders); two in the
// persistent loop, where a third takes its registers past four waves per
// SIMD (134 VGPRs)
constexpr int kMixedBoxes = 2;
constexpr int mixed_gpf(bool loop) { return loop ? 2 : 3; }
// u loaded and v stored with the non-temporal hint (each is touched once)
constexpr int kMixedNt = 1;

template <int P>
constexpr int smem_of()
{
    using S = bk13t::Bk13tShape<P + 1, P + 2, 1>;
    return S::epb * kMixedBoxes * S::box;
}

template <int P>
constexpr int max_smem()
{
    if constexpr (P == 1) {
        return smem_of<1>();
    } else {
        return smem_of<P>() > max_smem<P - 1>() ? smem_of<P>() : max_smem<P - 1>();
    }
}

constexpr int kMixedSmem = max_smem<8>();   // doubles

// EO: bases holds the even-odd forms of B_p and D_p (bkx_eo.hpp), at boff[p].
template <int P, int DIV, bool LOOP, bool EO>
__device__ __forceinline__ void run_tile(int tid, const MixedTile& t, const int64_t* __restrict__ uoff,
                                         const int64_t* __restrict__ goff,
                                         const double* __restrict__ bases, const int64_t* __restrict__ boff,
                                         const double* __restrict__ G, const double* __restrict__ u,
                                         double* __restrict__ v, double* smem)
{
    constexpr int NM = P + 1;
    constexpr int NQ = P + 2;
    using S = bk13t::Bk13tShape<NM, NQ, DIV>;
    const double* B = bases + boff[P];
    const double* D = B + (EO ? EoMat<NM, NQ, 1>::size : NM * NQ);
    const int slot = tid / S::nq2;
    const int l = tid - slot * S::nq2;
    if constexpr (!LOOP) {
        // one round: the tile holds at most epb elements
        const bool active = slot < S::epb && slot < t.count;
        const int e = t.first + (active ? slot : 0);
        bk13t::bk13t_body<NM, NQ, true, DIV, false, kMixedBoxes, mixed_gpf(LOOP), BKX_FORM_F64, double, kMixedNt, EO>(
            B, D, G + goff[e], u + uoff[e], v + uoff[e], active, slot, l, smem);
    } else {
        for (int r0 = 0; r0 < t.count; r0 += S::epb) {
            const int idx = r0 + slot;
            const bool active = slot < S::epb && idx < t.count;
            const int e = t.first + (active ? idx : 0);
            bk13t::bk13t_body<NM, NQ, true, DIV, false, kMixedBoxes, mixed_gpf(LOOP), BKX_FORM_F64, double, kMixedNt, EO>(
                B, D, G + goff[e], u + uoff[e], v + uoff[e], active, slot, l, smem, t.first + r0);
            __syncthreads();
        }
    }
}

template <int DIV, bool LOOP, bool EO>
__device__ __forceinline__ void run_tile_p(int tid, const MixedTile& t, const int64_t* uoff, const int64_t* goff,
                                           const double* bases, const int64_t* boff,
                                           const double* G, const double* u, double* v, double* smem)
{
    switch (t.p) {
    case 1:
        run_tile<1, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 2:
        run_tile<2, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 3:
        run_tile<3, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 4:
        run_tile<4, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 5:
        run_tile<5, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 6:
        run_tile<6, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    case 7:
        run_tile<7, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    default:
        run_tile<8, DIV, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        break;
    }
}

// LOOP = false: one workgroup per tile (a single launch over all orders; the
// hardware dispatches the workgroups in tile order as CUs free up).
// LOOP = true: persistent, each workgroup takes the next tile from the queue
// until none is left (two counters in device memory, zero between launches:
// launches with one plan must not overlap).
// In the persistent form the basis rows are read with vector loads (loads
// after the loop's barriers and stores cannot be proven unclobbered, so they
// do not become scalar loads), which costs registers and occupancy.
template <bool LOOP, bool EO>
__global__ __launch_bounds__(kMixedThreads) void bk3_mixed_kernel(
    int ntiles, const MixedTile* __restrict__ tiles, const int64_t* __restrict__ uoff,
    const int64_t* __restrict__ goff, const double* __restrict__ bases,
    const int64_t* __restrict__ boff, const double* __restrict__ G, const double* __restrict__ u,
    double* __restrict__ v, int* __restrict__ queue)
{
    __shared__ double smem[kMixedSmem];
    auto one = [&](int k, int tid) {
        const MixedTile t = tiles[k];
        // a tile of divisor 4 runs on the DIV = 1 body: its slots past the
        // tile's count stay idle, which is all a divisor changes here (the
        // workgroup has 256 threads whatever the divisor)
        if (t.div == 2) {
            run_tile_p<2, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        } else {
            run_tile_p<1, LOOP, EO>(tid, t, uoff, goff, bases, boff, G, u, v, smem);
        }
    };
    if constexpr (LOOP) {
        // tiles taken from a queue: queue[0] counts the tiles taken, queue[1]
        // the workgroups done; the last workgroup done resets both for the
        // next launch on the stream
        __shared__ int s_tile;
        for (;;) {
            if (threadIdx.x == 0) {
                s_tile = atomicAdd(&queue[0], 1);
            }
            __syncthreads();
            const int k = __builtin_amdgcn_readfirstlane(s_tile);
            __syncthreads();
            if (k >= ntiles) {
                break;
            }
            // The lane index, made to depend on the tile: the bodies' lane
            // arithmetic is then computed in the body that runs instead of
            // being hoisted out of the loop for all of them.
            int tid = threadIdx.x;
            asm("; tile %1" : "+v"(tid) : "s"(k));
            one(k, tid);
        }
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(&queue[1], 1) == int(gridDim.x) - 1) {
                atomicExch(&queue[0], 0);
                atomicExch(&queue[1], 0);
            }
        }
    } else {
        if (int(blockIdx.x) < ntiles) {
            one(blockIdx.x, threadIdx.x);
        }
    }
}

// One order's tiles, with that order's body alone: registers and LDS sized
// for P (the all-order kernel above is sized for its largest body, which
// costs the low orders occupancy), u and v streamed with non-temporal
// accesses and three rows of G in flight, as the fastest single-order BK3
// variants (with their register budget: at least four waves per SIMD).  Workgroup k runs tile k of the order's range, looping when the
// tile holds more than one workgroup's elements.  The arithmetic is the
// all-order kernel's: the two give the same bits.
constexpr int kOrderGpf = 3;   // the default rows of G in flight
constexpr int kOrderNt = 1;

// GPF: rows of G in flight (2, 3 or 4), DIV: elements-per-workgroup divisor
// (1, 2 or 4); the launch shape is chosen per order by the tuner
// (bkx_mixed_plan_set_order_shape).
template <int P, int DIV, bool EO, int GPF = kOrderGpf>
__global__ __launch_bounds__((bk13t::Bk13tShape<P + 1, P + 2, DIV>::threads))
__attribute__((amdgpu_waves_per_eu(4))) void bk3_order_kernel(
    int first, int count, const int64_t* __restrict__ uoff, const int64_t* __restrict__ goff,
    const double* __restrict__ B, const double* __restrict__ D, const double* __restrict__ G,
    const double* __restrict__ u, double* __restrict__ v)
{
    constexpr int NM = P + 1;
    constexpr int NQ = P + 2;
    using S = bk13t::Bk13tShape<NM, NQ, DIV>;
    __shared__ double smem[S::epb * kMixedBoxes * S::box];
    const int slot = threadIdx.x / S::nq2;
    const int l = threadIdx.x - slot * S::nq2;
    const int idx = int(blockIdx.x) * S::epb + slot;
    const bool active = slot < S::epb && idx < count;
    const int e = first + (active ? idx : 0);
    constexpr int gpf = GPF < NQ ? GPF : NQ;   // at most a whole element's rows
    bk13t::bk13t_body<NM, NQ, true, DIV, false, kMixedBoxes, gpf, BKX_FORM_F64, double, kOrderNt, EO>(
        B, D, G + goff[e], u + uoff[e], v + uoff[e], active, slot, l, smem);
}

int n_of(int p) { return p + 1; }
int q_of(int p) { return p + 2; }

}  // namespace
}  // namespace bkx

struct bkx_mixed_plan {
    size_t nelmt = 0;
    size_t ndof = 0;
    size_t ngeo = 0;
    int ntiles = 0;
    int grid = 0;
    bool persistent = false;    // loop over tiles (grid resident workgroups) or one workgroup per tile
    int elem_begin[10] = {};            // per order: range of the order-sorted element list
    int div_p[10] = {};                 // per order: elements-per-workgroup divisor
    int order_div[10] = {};             // per order: the per-order launch's divisor (1, 2, 4)
    int order_gpf[10] = {};             // per order: the per-order launch's rows of G in flight (2, 3, 4)
    int64_t boff_h[10] = {};            // per order: offset of B_p in the bases array
    int64_t eoff_h[10] = {};            // per order: offset of B_p's even-odd form in the prepared array
    std::vector<int64_t> uoff_h, goff_h;
    bkx::MixedTile* tiles = nullptr;
    int* queue = nullptr;       // the persistent launch's tile queue (two counters, zero between launches)
    int64_t* uoff = nullptr;    // by position in the order-sorted element list
    int64_t* goff = nullptr;
    int64_t* boff = nullptr;
    int64_t* eoff = nullptr;
};

struct bkx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <typename T>
int upload(T** dst, const T* src, size_t n)
{
    if (hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
        *dst = nullptr;
        return BKX_ERROR_DEVICE;
    }
    if (n > 0 && hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        return BKX_ERROR_DEVICE;
    }
    return BKX_SUCCESS;
}

template <bool EO>
int launch_tiles(const bkx::MixedTile* tiles, int ntiles, bool persistent, const bkx_mixed_plan* pl,
                 const double* bases, const double* G, const double* in, double* out, hipStream_t s)
{
    if (ntiles == 0) {
        return BKX_SUCCESS;
    }
    const int64_t* off = EO ? pl->eoff : pl->boff;
    if (persistent) {
        hipLaunchKernelGGL((bkx::bk3_mixed_kernel<true, EO>), dim3(std::min(pl->grid, ntiles)), dim3(bkx::kMixedThreads),
                           0, s, ntiles, tiles, pl->uoff, pl->goff, bases, off, G, in, out, pl->queue);
    } else {
        hipLaunchKernelGGL((bkx::bk3_mixed_kernel<false, EO>), dim3(ntiles), dim3(bkx::kMixedThreads), 0, s, ntiles,
                           tiles, pl->uoff, pl->goff, bases, off, G, in, out, pl->queue);
    }
    return bkx::launch_status();
}

template <int P, int DIV, bool EO, int GPF>
void launch_order_div(int first, int count, const bkx_mixed_plan* pl, const double* bases, const double* G,
                      const double* in, double* out, hipStream_t s)
{
    using S = bkx::bk13t::Bk13tShape<P + 1, P + 2, DIV>;
    const double* B = bases + (EO ? pl->eoff_h[P] : pl->boff_h[P]);
    const double* D = B + (EO ? bkx::EoMat<P + 1, P + 2, 1>::size : (P + 1) * (P + 2));
    const unsigned grid = unsigned((count + S::epb - 1) / S::epb);
    hipLaunchKernelGGL((bkx::bk3_order_kernel<P, DIV, EO, GPF>), dim3(grid), dim3(S::threads), 0, s, first, count,
                       pl->uoff, pl->goff, B, D, G, in, out);
}

template <int P, int DIV, bool EO>
void launch_order_gpf(int gpf, int first, int count, const bkx_mixed_plan* pl, const double* bases,
                      const double* G, const double* in, double* out, hipStream_t s)
{
    if (gpf == 2) {
        launch_order_div<P, DIV, EO, 2>(first, count, pl, bases, G, in, out, s);
    } else if (gpf == 4) {
        launch_order_div<P, DIV, EO, 4>(first, count, pl, bases, G, in, out, s);
    } else {
        launch_order_div<P, DIV, EO, 3>(first, count, pl, bases, G, in, out, s);
    }
}

// One order's elements, by that order's own kernel: workgroup k takes
// elements k epb .. k epb + epb - 1 of the order's part of the sorted list.
template <bool EO>
int launch_order(int p, const bkx_mixed_plan* pl, const double* bases, const double* G, const double* in,
                 double* out, hipStream_t s)
{
    const int first = pl->elem_begin[p];
    const int count = pl->elem_begin[p + 1] - first;
    if (count == 0) {
        return BKX_SUCCESS;
    }
    const int div = pl->order_div[p];
    const int gpf = pl->order_gpf[p];
    return bkx::with_order<1, 8>(p, [&](auto P) {
        constexpr int PP = decltype(P)::value;
        if (div == 4) {
            launch_order_gpf<PP, 4, EO>(gpf, first, count, pl, bases, G, in, out, s);
        } else if (div == 2) {
            launch_order_gpf<PP, 2, EO>(gpf, first, count, pl, bases, G, in, out, s);
        } else {
            launch_order_gpf<PP, 1, EO>(gpf, first, count, pl, bases, G, in, out, s);
        }
        return bkx::launch_status();
    });
}

// The entry points' bodies, with the bases (EO: the prepared even-odd
// forms, bkx_bk3_mixed_prepare) as given.
template <bool EO>
int apply_single(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in, double* out,
                 hipStream_t s)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (plan->nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!bases || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return launch_tiles<EO>(plan->tiles, plan->ntiles, plan->persistent, plan, bases, G, in, out, s);
}

template <bool EO>
int apply_by_order(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in, double* out,
                   hipStream_t s)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (plan->nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!bases || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    for (int p = 1; p <= 8; ++p) {
        const int st = launch_order<EO>(p, plan, bases, G, in, out, s);
        if (st != BKX_SUCCESS) {
            return st;
        }
    }
    return BKX_SUCCESS;
}

template <bool EO>
int graph_create(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in, double* out,
                 int by_order, bkx_graph** graph)
{
    if (!plan || !graph) {
        return BKX_ERROR_NULL;
    }
    *graph = nullptr;
    auto* g = new (std::nothrow) bkx_graph;
    if (!g) {
        return BKX_ERROR_DEVICE;
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return BKX_ERROR_DEVICE;
    }
    int st = BKX_SUCCESS;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        st = BKX_ERROR_DEVICE;
    } else {
        st = by_order ? apply_by_order<EO>(plan, bases, G, in, out, s) : apply_single<EO>(plan, bases, G, in, out, s);
        hipGraph_t captured = nullptr;
        const hipError_t end = hipStreamEndCapture(s, &captured);
        if (st == BKX_SUCCESS && end != hipSuccess) {
            st = BKX_ERROR_DEVICE;
        }
        g->graph = captured;
        if (st == BKX_SUCCESS && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
            st = BKX_ERROR_DEVICE;
        }
    }
    (void)hipStreamDestroy(s);
    if (st != BKX_SUCCESS) {
        bkx_graph_destroy(g);
        return st;
    }
    *graph = g;
    return BKX_SUCCESS;
}

}  // namespace

namespace {

// Offsets of the grouped layout: each order's elements contiguous, in
// ascending element index, orders ascending, each order's range of u/v and
// of G starting on a 256-byte boundary.
constexpr int64_t kGroupAlign = 32;   // doubles

int64_t align_up(int64_t x) { return (x + kGroupAlign - 1) / kGroupAlign * kGroupAlign; }

int plan_create(const int32_t* order, size_t nelmt, const int32_t* div_by_order, const int32_t* rounds_by_order,
                int grid, bool group_layout, bkx_mixed_plan** plan);

}  // namespace

extern "C" {

int bkx_mixed_plan_create(const int32_t* order, size_t nelmt, const int32_t* div_by_order,
                          const int32_t* rounds_by_order, int grid, bkx_mixed_plan** plan)
{
    return plan_create(order, nelmt, div_by_order, rounds_by_order, grid, false, plan);
}

int bkx_mixed_plan_create_grouped(const int32_t* order, size_t nelmt, const int32_t* div_by_order,
                                  const int32_t* rounds_by_order, int grid, bkx_mixed_plan** plan)
{
    return plan_create(order, nelmt, div_by_order, rounds_by_order, grid, true, plan);
}

}  // extern "C"

namespace {

int plan_create(const int32_t* order, size_t nelmt, const int32_t* div_by_order, const int32_t* rounds_by_order,
                int grid, bool group_layout, bkx_mixed_plan** plan)
{
    if (!plan || (nelmt > 0 && !order)) {
        return BKX_ERROR_NULL;
    }
    *plan = nullptr;
    if (nelmt > size_t(INT32_MAX)) {
        return BKX_ERROR_SIZE;
    }
    for (int p = 1; p <= 8 && div_by_order; ++p) {
        if (div_by_order[p] != 1 && div_by_order[p] != 2 && div_by_order[p] != 4) {
            return BKX_ERROR_ARG;
        }
    }
    auto* pl = new (std::nothrow) bkx_mixed_plan;
    if (!pl) {
        return BKX_ERROR_DEVICE;
    }
    pl->nelmt = nelmt;
    pl->uoff_h.resize(nelmt);
    pl->goff_h.resize(nelmt);
    std::vector<std::vector<int32_t>> by_p(9);
    int64_t uo = 0, go = 0;
    for (size_t e = 0; e < nelmt; ++e) {
        const int p = order[e];
        if (p < 1 || p > 8) {
            delete pl;
            return BKX_ERROR_ORDER;
        }
        pl->uoff_h[e] = uo;
        pl->goff_h[e] = go;
        uo += int64_t(bkx::n_of(p)) * bkx::n_of(p) * bkx::n_of(p);
        go += int64_t(6) * bkx::q_of(p) * bkx::q_of(p) * bkx::q_of(p);
        by_p[p].push_back(int32_t(e));
    }
    if (group_layout) {
        uo = 0;
        go = 0;
        for (int p = 1; p <= 8; ++p) {
            uo = align_up(uo);
            go = align_up(go);
            for (const int32_t e : by_p[p]) {
                pl->uoff_h[size_t(e)] = uo;
                pl->goff_h[size_t(e)] = go;
                uo += int64_t(bkx::n_of(p)) * bkx::n_of(p) * bkx::n_of(p);
                go += int64_t(6) * bkx::q_of(p) * bkx::q_of(p) * bkx::q_of(p);
            }
        }
    }
    pl->ndof = size_t(uo);
    pl->ngeo = size_t(go);

    std::vector<int32_t> elems;
    std::vector<bkx::MixedTile> grouped;
    for (int p = 1; p <= 8; ++p) {
        const int q = bkx::q_of(p);
        const int div = div_by_order ? div_by_order[p] : 1;
        const int full = std::max(1, 256 / (q * q));
        const int epb = std::max(1, full / div);
        const int rounds = (rounds_by_order && rounds_by_order[p] > 0) ? rounds_by_order[p] : 1;
        if (rounds > 1) {
            pl->persistent = true;
        }
        const int per_tile = int(std::min<int64_t>(int64_t(epb) * rounds, INT32_MAX));
        pl->div_p[p] = div;
        pl->order_div[p] = div;
        pl->order_gpf[p] = bkx::kOrderGpf;
        const int first = int(elems.size());
        pl->elem_begin[p] = first;
        elems.insert(elems.end(), by_p[p].begin(), by_p[p].end());
        const int cnt = int(by_p[p].size());
        for (int t = 0; t < cnt; t += std::min(per_tile, cnt - t)) {
            const int c = std::min(per_tile, cnt - t);
            grouped.push_back(bkx::MixedTile{p, div, first + t, c});
        }
    }
    pl->elem_begin[9] = int(elems.size());
    // each element's offsets at its position in the sorted list: a workgroup
    // finds its elements' data with one load, not through the element index
    std::vector<int64_t> uoff_s(elems.size()), goff_s(elems.size());
    for (size_t i = 0; i < elems.size(); ++i) {
        uoff_s[i] = pl->uoff_h[size_t(elems[i])];
        goff_s[i] = pl->goff_h[size_t(elems[i])];
    }
    pl->ntiles = int(grouped.size());
    int64_t boff[10] = {};
    int64_t bo = 0;
    for (int p = 1; p <= 8; ++p) {
        boff[p] = bo;
        bo += int64_t(bkx::n_of(p)) * bkx::q_of(p) + int64_t(bkx::q_of(p)) * bkx::q_of(p);
    }
    boff[9] = bo;
    std::copy(boff, boff + 10, pl->boff_h);
    int64_t eoff[10] = {};
    int64_t eo = 0;
    for (int p = 1; p <= 8; ++p) {
        eoff[p] = eo;
        eo += int64_t(bkx_bk3_basis_size(p));
    }
    eoff[9] = eo;
    std::copy(eoff, eoff + 10, pl->eoff_h);

    const bool grid_given = grid > 0;
    if (!grid_given) {
        int dev = 0, cus = 256, per_cu = 1;
        if (hipGetDevice(&dev) == hipSuccess) {
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        }
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bkx::bk3_mixed_kernel<true, false>, bkx::kMixedThreads,
                                                         0) !=
                hipSuccess ||
            per_cu < 1) {
            per_cu = 1;
        }
        grid = cus * per_cu;
    }
    // persistent when asked for: rounds of more than one workgroup's
    // elements per tile, or a grid given without rounds.  Otherwise one
    // workgroup per tile, the hardware dispatching them in list order as CUs
    // free up (on MI355X 59.2 GDoF/s for the 200,000-element Zipf batch in
    // the caller's layout, against 55.3 for the persistent loop).
    if (grid_given && rounds_by_order == nullptr) {
        pl->persistent = true;
    }
    pl->grid = std::max(1, std::min(grid, std::max(1, pl->ntiles)));

    // Tile order: by the memory position of the tile's first element, last
    // first, for both launch forms (the persistent one takes tiles from a
    // queue, so any order balances).  The workgroups in flight then work on
    // neighbouring elements: in the caller's layout, which interleaves the
    // orders, the single launch ran at 59.2 against 56.7 GDoF/s longest
    // first, and the persistent one at 55.3 against 50.7 with a static
    // longest-first split.  In the grouped layout the orders lie in
    // ascending order in memory, so this is longest first as well.
    std::vector<int> idx(grouped.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return uoff_s[grouped[a].first] > uoff_s[grouped[b].first]; });
    std::vector<bkx::MixedTile> ordered(grouped.size());
    for (size_t k = 0; k < idx.size(); ++k) {
        ordered[k] = grouped[idx[k]];
    }

    int st = BKX_SUCCESS;
    st = st ? st : upload(&pl->tiles, ordered.data(), ordered.size());
    const int zero[2] = {0, 0};
    st = st ? st : upload(&pl->queue, zero, 2);
    st = st ? st : upload(&pl->uoff, uoff_s.data(), nelmt);
    st = st ? st : upload(&pl->goff, goff_s.data(), nelmt);
    st = st ? st : upload(&pl->boff, boff, 10);
    st = st ? st : upload(&pl->eoff, eoff, 10);
    if (st != BKX_SUCCESS) {
        bkx_mixed_plan_destroy(pl);
        return st;
    }
    *plan = pl;
    return BKX_SUCCESS;
}

}  // namespace

extern "C" {

int bkx_mixed_plan_destroy(bkx_mixed_plan* plan)
{
    if (!plan) {
        return BKX_SUCCESS;
    }
    for (void* ptr : {static_cast<void*>(plan->tiles), static_cast<void*>(plan->queue),
                      static_cast<void*>(plan->uoff),
                      static_cast<void*>(plan->goff), static_cast<void*>(plan->boff),
                      static_cast<void*>(plan->eoff)}) {
        if (ptr) {
            (void)hipFree(ptr);
        }
    }
    delete plan;
    return BKX_SUCCESS;
}

int bkx_mixed_plan_info(const bkx_mixed_plan* plan, size_t* ndof, size_t* ngeo, int* ntiles, int* grid)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (ndof) {
        *ndof = plan->ndof;
    }
    if (ngeo) {
        *ngeo = plan->ngeo;
    }
    if (ntiles) {
        *ntiles = plan->ntiles;
    }
    if (grid) {
        *grid = plan->grid;
    }
    return BKX_SUCCESS;
}

int bkx_mixed_plan_set_order_shape(bkx_mixed_plan* plan, int p, int div, int gpf)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (p < 1 || p > 8) {
        return BKX_ERROR_ORDER;
    }
    if ((div != 1 && div != 2 && div != 4) || gpf < 2 || gpf > 4) {
        return BKX_ERROR_ARG;
    }
    plan->order_div[p] = div;
    plan->order_gpf[p] = gpf;
    return BKX_SUCCESS;
}

int bkx_mixed_plan_get_order_shape(const bkx_mixed_plan* plan, int p, int* div, int* gpf)
{
    if (!plan || !div || !gpf) {
        return BKX_ERROR_NULL;
    }
    if (p < 1 || p > 8) {
        return BKX_ERROR_ORDER;
    }
    *div = plan->order_div[p];
    *gpf = plan->order_gpf[p];
    return BKX_SUCCESS;
}

int bkx_mixed_plan_offsets(const bkx_mixed_plan* plan, int64_t* uoff, int64_t* goff)
{
    if (!plan || !uoff || !goff) {
        return BKX_ERROR_NULL;
    }
    std::copy(plan->uoff_h.begin(), plan->uoff_h.end(), uoff);
    std::copy(plan->goff_h.begin(), plan->goff_h.end(), goff);
    return BKX_SUCCESS;
}

int bkx_bk3_mixed_apply_f64(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in,
                            double* out, bkx_stream_t stream)
{
    return apply_single<false>(plan, bases, G, in, out, reinterpret_cast<hipStream_t>(stream));
}

int bkx_bk3_mixed_apply_by_order_f64(const bkx_mixed_plan* plan, const double* bases, const double* G,
                                     const double* in, double* out, bkx_stream_t stream)
{
    return apply_by_order<false>(plan, bases, G, in, out, reinterpret_cast<hipStream_t>(stream));
}

int bkx_bk3_mixed_apply_order_f64(const bkx_mixed_plan* plan, int p, const double* bases, const double* G,
                                  const double* in, double* out, bkx_stream_t stream)
{
    if (!plan) {
        return BKX_ERROR_NULL;
    }
    if (p < 1 || p > 8) {
        return BKX_ERROR_ORDER;
    }
    if (!bases || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return launch_order<false>(p, plan, bases, G, in, out, reinterpret_cast<hipStream_t>(stream));
}

int bkx_bk3_mixed_graph_create(const bkx_mixed_plan* plan, const double* bases, const double* G, const double* in,
                               double* out, int by_order, bkx_graph** graph)
{
    return graph_create<false>(plan, bases, G, in, out, by_order, graph);
}

size_t bkx_bk3_mixed_prepared_size(void)
{
    size_t n = 0;
    for (int p = 1; p <= 8; ++p) {
        n += bkx_bk3_basis_size(p);
    }
    return n;
}

int bkx_bk3_mixed_prepare(const double* bases, double* prepared)
{
    if (!bases || !prepared) {
        return BKX_ERROR_NULL;
    }
    const double* b = bases;
    double* f = prepared;
    for (int p = 1; p <= 8; ++p) {
        const int nm = bkx::n_of(p), nq = bkx::q_of(p);
        const int st = bkx_bk3_basis_prepare(p, b, b + nm * nq, f);
        if (st != BKX_SUCCESS) {
            return st;
        }
        b += nm * nq + nq * nq;
        f += bkx_bk3_basis_size(p);
    }
    return BKX_SUCCESS;
}

int bkx_bk3_mixed_apply_prepared_f64(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                     const double* in, double* out, bkx_stream_t stream)
{
    return apply_single<true>(plan, prepared, G, in, out, reinterpret_cast<hipStream_t>(stream));
}

int bkx_bk3_mixed_apply_by_order_prepared_f64(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                              const double* in, double* out, bkx_stream_t stream)
{
    return apply_by_order<true>(plan, prepared, G, in, out, reinterpret_cast<hipStream_t>(stream));
}

int bkx_bk3_mixed_graph_create_prepared(const bkx_mixed_plan* plan, const double* prepared, const double* G,
                                        const double* in, double* out, int by_order, bkx_graph** graph)
{
    return graph_create<true>(plan, prepared, G, in, out, by_order, graph);
}

int bkx_graph_launch(bkx_graph* graph, bkx_stream_t stream)
{
    if (!graph || !graph->exec) {
        return BKX_ERROR_NULL;
    }
    return hipGraphLaunch(graph->exec, reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? BKX_SUCCESS
                                                                                           : BKX_ERROR_LAUNCH;
}

int bkx_graph_destroy(bkx_graph* graph)
{
    if (!graph) {
        return BKX_SUCCESS;
    }
    if (graph->exec) {
        (void)hipGraphExecDestroy(graph->exec);
    }
    if (graph->graph) {
        (void)hipGraphDestroy(graph->graph);
    }
    delete graph;
    return BKX_SUCCESS;
}

}  // extern "C"
