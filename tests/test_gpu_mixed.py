"""Mixed-order BK3 batches (p = 1..8 in one E-vector): the single launch,
the persistent launch, the per-order launches and their hipGraphs, against
the FP64 oracle per order (<= 1e-13) and against each other."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import geometry as geo
from omp_bk_amd import mixed

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _batch(nelmt, seed):
    order = mixed.zipf_orders(nelmt, seed=seed)
    plan = mixed.MixedPlan(order)
    bases = mixed.packed_bases()
    G = mixed.affine_factors(order, plan.goff, plan.ngeo, seed=seed)
    u = np.random.default_rng(seed).uniform(-1, 1, plan.ndof)
    return order, plan, bases, G, u


def _oracle_check(order, plan, bases, G, u, v, elems=None):
    """Each order's elements (or the given subset) against bko_bk3_sumfact."""
    boff = 0
    worst = 0.0
    for p in range(1, 9):
        n, q = p + 1, p + 2
        B = bases[boff:boff + n * q].reshape(n, q)
        D = bases[boff + n * q:boff + n * q + q * q].reshape(q, q)
        boff += n * q + q * q
        idx = np.nonzero(order == p)[0]
        if elems is not None:
            idx = np.intersect1d(idx, elems)
        if idx.size == 0:
            continue
        ui = np.concatenate([u[plan.uoff[e]:plan.uoff[e] + n ** 3] for e in idx])
        Gi = np.concatenate([G[plan.goff[e]:plan.goff[e] + 6 * q ** 3] for e in idx])
        vi = np.concatenate([v[plan.uoff[e]:plan.uoff[e] + n ** 3] for e in idx])
        ref = O.bk3(q, B, D, Gi, ui)
        assert np.all(np.isfinite(vi))
        err = O.rel_maxdiff(vi, ref)
        worst = max(worst, err)
        assert err <= 1e-13, (p, err)
    return worst


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


@pytest.mark.parametrize("nelmt", [1, 9, 517])
def test_small_batches(nelmt):
    order, plan, bases, G, u = _batch(nelmt, seed=nelmt)
    b, g, uu = _t(bases), _t(G), _t(u)
    v1 = torch.full_like(uu, float("nan"))
    plan.apply(b, g, uu, v1)
    v2 = torch.full_like(uu, float("nan"))
    plan.apply_by_order(b, g, uu, v2)
    torch.cuda.synchronize()
    _oracle_check(order, plan, bases, G, u, v1.cpu().numpy())
    assert torch.equal(v1, v2)
    persist = mixed.MixedPlan(order, rounds_by_order=[2] * 9, grid=64)
    v3 = torch.full_like(uu, float("nan"))
    persist.apply(b, g, uu, v3)
    torch.cuda.synchronize()
    assert torch.equal(v1, v3)


def test_tuned_order_shapes():
    """The per-order tuner sets a launch shape (divisor, rows of G in
    flight) per order on the plan; the tuned launches match the oracle."""
    order, plan, bases, G, u = _batch(3000, seed=5)
    b, g, uu = _t(bases), _t(G), _t(u)
    v = torch.empty_like(uu)
    shapes, times = mixed.tune(order, b, g, uu, v, steps=3, plan=plan)
    assert set(shapes) == set(int(p) for p in np.unique(order))
    for p, shape in shapes.items():
        assert shape in mixed.ORDER_SHAPES and plan.order_shape(p) == shape
        assert len(times[p]) == len(mixed.ORDER_SHAPES)
    vt = torch.full_like(uu, float("nan"))
    plan.apply_by_order(b, g, uu, vt)
    torch.cuda.synchronize()
    _oracle_check(order, plan, bases, G, u, vt.cpu().numpy())


def test_every_order_shape_same_bits():
    """Every per-order launch shape gives the bits of the default shape,
    also with the bases in even-odd form."""
    order, plan, bases, G, u = _batch(2000, seed=15)
    b, g, uu = _t(bases), _t(G), _t(u)
    P = mixed.prepared_bases(b)
    ref = torch.full_like(uu, float("nan"))
    plan.apply_by_order(b, g, uu, ref)
    ref_eo = torch.full_like(uu, float("nan"))
    plan.apply_by_order(P, g, uu, ref_eo, prepared=True)
    for div, gpf in mixed.ORDER_SHAPES:
        for p in range(1, 9):
            plan.set_order_shape(p, div, gpf)
        v = torch.full_like(uu, float("nan"))
        plan.apply_by_order(b, g, uu, v)
        v_eo = torch.full_like(uu, float("nan"))
        plan.apply_by_order(P, g, uu, v_eo, prepared=True)
        torch.cuda.synchronize()
        assert torch.equal(v, ref), (div, gpf)
        assert torch.equal(v_eo, ref_eo), (div, gpf)
    with pytest.raises(RuntimeError):
        plan.set_order_shape(3, 3, 3)


def test_tuned_single_launch_divisors():
    order, plan, bases, G, u = _batch(3000, seed=6)
    b, g, uu = _t(bases), _t(G), _t(u)
    v = torch.empty_like(uu)
    div, ms = mixed.tune_single(order, b, g, uu, v, steps=3)
    assert len(div) == 9 and set(div) <= {1, 2, 4} and ms > 0
    tuned = mixed.MixedPlan(order, div_by_order=div)
    vt = torch.full_like(uu, float("nan"))
    tuned.apply(b, g, uu, vt)
    torch.cuda.synchronize()
    _oracle_check(order, tuned, bases, G, u, vt.cpu().numpy())
    # every divisor, one workgroup per tile and persistent (three rounds a
    # tile, so that workgroups loop): the same bits as divisor 1
    ref = torch.full_like(uu, float("nan"))
    plan.apply(b, g, uu, ref)
    for d in (2, 4):
        for rounds in (None, [3] * 9):
            pl = mixed.MixedPlan(order, div_by_order=[d] * 9, rounds_by_order=rounds)
            vd = torch.full_like(uu, float("nan"))
            pl.apply(b, g, uu, vd)
            torch.cuda.synchronize()
            assert torch.equal(vd, ref), (d, rounds)
            pl.close()


def test_graph_replays():
    """Each replay of the captured launch, with eager work between replays,
    equals the eager result."""
    order, plan, bases, G, u = _batch(4000, seed=3)
    b, g, uu = _t(bases), _t(G), _t(u)
    ref = torch.empty_like(uu)
    plan.apply(b, g, uu, ref)
    for by_order in (False, True):
        v = torch.full_like(uu, float("nan"))
        graph = plan.graph(b, g, uu, v, by_order=by_order)
        for rep in range(4):
            v.fill_(float("nan"))
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(v, ref), (by_order, rep)
            _ = (uu * 2.0).sum().item()      # eager work between replays
        graph.close()


def test_persistent_queue_repeats():
    """The persistent launch takes its tiles from a queue that each launch
    leaves at zero: five eager applies in a row and four graph replays (with
    eager work between them) equal the one-workgroup-per-tile launch, at a
    small grid, at the resident grid, and with a grid clamped to the tile
    count."""
    order, plan, bases, G, u = _batch(4000, seed=5)
    b, g, uu = _t(bases), _t(G), _t(u)
    ref = torch.empty_like(uu)
    plan.apply(b, g, uu, ref)
    for kw in (dict(rounds_by_order=[2] * 9, grid=64), dict(rounds_by_order=[3] * 9), dict(grid=1 << 20)):
        persist = mixed.MixedPlan(order, **kw)
        v = torch.full_like(uu, float("nan"))
        for rep in range(5):
            v.fill_(float("nan"))
            persist.apply(b, g, uu, v)
            torch.cuda.synchronize()
            assert torch.equal(v, ref), (kw, "eager", rep)
        graph = persist.graph(b, g, uu, v)
        for rep in range(4):
            v.fill_(float("nan"))
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(v, ref), (kw, "graph", rep)
            _ = (uu * 2.0).sum().item()      # eager work between replays
        graph.close()


def test_full_batch_sampled():
    """The 200,000-element Zipf batch: sampled elements of every order
    against the oracle; single launch == per-order launches."""
    order, plan, bases, G, u = _batch(200000, seed=2024)
    b, g, uu = _t(bases), _t(G), _t(u)
    v1 = torch.full_like(uu, float("nan"))
    v2 = torch.full_like(uu, float("nan"))
    plan.apply(b, g, uu, v1)
    plan.apply_by_order(b, g, uu, v2)
    torch.cuda.synchronize()
    assert torch.equal(v1, v2)
    assert torch.isfinite(v1).all()
    rng = np.random.default_rng(0)
    sample = np.concatenate([rng.choice(np.nonzero(order == p)[0], 40, replace=False) for p in range(1, 9)])
    _oracle_check(order, plan, bases, G, u, v1.cpu().numpy(), elems=sample)


def test_prepared_bases_every_launch_form():
    """The bases of every order in even-odd form: the single launch, the
    persistent launch, the per-order launches and their graphs, against the
    oracle per order and the direct launches (1e-14: the even-odd sums round
    differently)."""
    order, plan, bases, G, u = _batch(3000, seed=13)
    b, g, uu = _t(bases), _t(G), _t(u)
    P = mixed.prepared_bases(b)
    direct = torch.full_like(uu, float("nan"))
    plan.apply(b, g, uu, direct)
    torch.cuda.synchronize()
    big = direct.abs().max().item()
    persist = mixed.MixedPlan(order, rounds_by_order=[2] * 9, grid=64)
    for name, fn in (("single", lambda v: plan.apply(P, g, uu, v, prepared=True)),
                     ("by_order", lambda v: plan.apply_by_order(P, g, uu, v, prepared=True)),
                     ("persistent", lambda v: persist.apply(P, g, uu, v, prepared=True))):
        print("mixed prepared", name, flush=True)
        v = torch.full_like(uu, float("nan"))
        fn(v)
        torch.cuda.synchronize()
        assert (v - direct).abs().max().item() <= 1e-14 * big, name
        if name == "single":
            _oracle_check(order, plan, bases, G, u, v.cpu().numpy())
    for by_order in (False, True):
        v = torch.full_like(uu, float("nan"))
        graph = plan.graph(P, g, uu, v, by_order=by_order, prepared=True)
        graph.launch()
        torch.cuda.synchronize()
        assert (v - direct).abs().max().item() <= 1e-14 * big, by_order
        graph.close()
    persist.close()


def test_grouped_layout_same_bits():
    """A plan for the grouped layout (each order's elements contiguous, each
    order's range on a 256-byte boundary): the same elements' data placed
    by its offsets give the bits of the caller-order plan, element by
    element, in every launch form."""
    order, plan, bases, G, u = _batch(3000, seed=21)
    gp = mixed.MixedPlan(order, grouped=True)
    assert gp.dofs == plan.dofs == plan.ndof and gp.ndof >= plan.ndof and gp.ngeo >= plan.ngeo
    for p in range(1, 9):
        idx = np.nonzero(order == p)[0]
        n3, g3 = (p + 1) ** 3, 6 * (p + 2) ** 3
        assert gp.uoff[idx[0]] % 32 == 0 and gp.goff[idx[0]] % 32 == 0
        assert np.array_equal(np.diff(gp.uoff[idx]), np.full(idx.size - 1, n3))
        assert np.array_equal(np.diff(gp.goff[idx]), np.full(idx.size - 1, g3))
    ug = np.zeros(gp.ndof)
    Gg = np.zeros(gp.ngeo)
    for e in range(order.size):
        n3, g3 = (order[e] + 1) ** 3, 6 * (order[e] + 2) ** 3
        ug[gp.uoff[e]:gp.uoff[e] + n3] = u[plan.uoff[e]:plan.uoff[e] + n3]
        Gg[gp.goff[e]:gp.goff[e] + g3] = G[plan.goff[e]:plan.goff[e] + g3]
    b, g, uu = _t(bases), _t(G), _t(u)
    bg, ugt = _t(Gg), _t(ug)
    ref = torch.empty_like(uu)
    plan.apply(b, g, uu, ref)
    ref = ref.cpu().numpy()
    P = mixed.prepared_bases(b)
    refp = torch.empty_like(uu)
    plan.apply(P, g, uu, refp, prepared=True)
    refp = refp.cpu().numpy()
    persist = mixed.MixedPlan(order, rounds_by_order=[2] * 9, grid=64, grouped=True)
    runs = [(gp.apply, False), (gp.apply_by_order, False), (persist.apply, False), (gp.apply, True),
            (gp.apply_by_order, True)]
    for fn, prep in runs:
        v = torch.full_like(ugt, float("nan"))
        fn(P if prep else b, bg, ugt, v, prepared=prep)
        v = v.cpu().numpy()
        want = refp if prep else ref
        for e in range(order.size):
            n3 = (order[e] + 1) ** 3
            assert np.array_equal(v[gp.uoff[e]:gp.uoff[e] + n3], want[plan.uoff[e]:plan.uoff[e] + n3]), (fn, prep, e)
    v = torch.full_like(ugt, float("nan"))
    graph = gp.graph(b, bg, ugt, v, by_order=True)
    graph.launch()
    torch.cuda.synchronize()
    got = v.cpu().numpy()
    assert all(np.array_equal(got[gp.uoff[e]:gp.uoff[e] + (order[e] + 1) ** 3],
                              ref[plan.uoff[e]:plan.uoff[e] + (order[e] + 1) ** 3]) for e in range(order.size))
    graph.close()
    persist.close()
    gp.close()


def test_bases_size_checked():
    """The packed bases (708 doubles) and their even-odd forms (756) are
    checked against the form asked for: a short or mismatched array raises
    instead of being read out of bounds on the device."""
    order = np.array([1, 3, 5, 8], dtype=np.int32)
    plan = mixed.MixedPlan(order)
    B = torch.as_tensor(mixed.packed_bases(), device=DEV)
    P = mixed.prepared_bases(B)
    G = torch.zeros(plan.ngeo, dtype=torch.float64, device=DEV)
    u = torch.zeros(plan.ndof, dtype=torch.float64, device=DEV)
    v = torch.empty_like(u)
    with pytest.raises(ValueError):
        plan.apply(B, G, u, v, prepared=True)
    with pytest.raises(ValueError):
        plan.apply(P, G, u, v)
    with pytest.raises(ValueError):
        plan.apply_by_order(B[:-1], G, u, v)
    plan.apply(B, G, u, v)
    plan.apply(P, G, u, v, prepared=True)
    torch.cuda.synchronize()
