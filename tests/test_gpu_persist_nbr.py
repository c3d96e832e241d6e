"""Persistent warpings with neighbour flags (launch rule persistent = 2).

One launch runs a whole warping without grid barriers: a tile starts a
chunk of K iterations once its 8 neighbours have finished the previous
chunk, and one extra workgroup sums each chunk's error partials (the same
fixed order as the barrier kernel) and publishes the stopping decision.
Tiles iterate the next chunk speculatively and check the decision before
their first store; a stop inside a chunk re-runs that many iterations from
the chunk's input.  Every case must give the oracle's flow bits and
iteration counts: all launch shapes, warpings that stop inside a chunk and
on its last iteration, the iteration cap inside and at a chunk's end, a cap
of 0, grids of several tiles per workgroup (1920x1080), mixed launch modes
per scale, and the residency report with its host-paced repeat."""
import numpy as np
import pytest

import oracle as O
from amdflow import PERSIST_HOST_PACED, PERSIST_PLAIN, TV_L1, persist_launch_mode, \
    set_persist_launch_mode, synth
from amdflow import debug as D

pytestmark = pytest.mark.gpu

NBR = 2
SHAPES = [(1, 2), (2, 2), (1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8), (8, 8),
          (1, 0), (2, 0), (4, 0)]


def _pair(w, h, cap, seed):
    I0, I1, _, _ = synth.make_pair(w, h, cap, seed, bumps=True)
    return synth.quantize(I0), synth.quantize(I1)


def _solve(q0, q1, rules, **prm):
    h, w = q0.shape
    with TV_L1(w, h, **prm) as tv:
        for min_px, K, R, P in rules:
            tv.set_launch_rule(min_px, K, R, P)
        tv.runDualTVL1Multiscale(q0, q1)
        return tv.getU().copy(), tv.iterations()


def _oracle(q0, q1, **prm):
    return O.tvl1(q0, q1, O.default_params(**prm))


@pytest.mark.parametrize("K,R", SHAPES)
def test_every_shape_matches_oracle(K, R):
    print("nbr shape", K, R, flush=True)
    q0, q1 = _pair(240, 180, 5.0, 13)
    u, it = _solve(q0, q1, [(0, K, R, NBR)])
    uo, ito = _oracle(q0, q1)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


@pytest.mark.parametrize("K,R,eps,cap", [(4, 0, 0.05, 300), (4, 4, 0.002, 300), (8, 8, 0.01, 13),
                                         (2, 8, 0.01, 7), (4, 0, 0.01, 8), (4, 4, 0.0, 10),
                                         (8, 8, 0.03, 300), (1, 0, 0.01, 1)])
def test_stops_inside_and_at_chunk_ends(K, R, eps, cap):
    """Stopping thresholds and caps chosen so that warpings end inside a
    chunk (a re-run), on a chunk's last iteration, and at the cap inside a
    chunk and at its end (epsilon 0: the cap alone)."""
    print("nbr stops", K, R, eps, cap, flush=True)
    q0, q1 = _pair(320, 240, 6.0, 17)
    prm = dict(nscales=4, epsilon=eps, max_iterations=cap)
    u, it = _solve(q0, q1, [(0, K, R, NBR)], **prm)
    uo, ito = _oracle(q0, q1, **prm)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)
    if cap % K:
        assert (it == cap).any() or (it % K != 0).any()


def test_zero_iteration_cap():
    q0, q1 = _pair(160, 120, 3.0, 19)
    u, it = _solve(q0, q1, [(0, 4, 4, NBR)], nscales=3, max_iterations=0)
    uo, ito = _oracle(q0, q1, nscales=3, max_iterations=0)
    assert (it == 0).all() and np.array_equal(it, ito)
    assert np.array_equal(u, uo)


@pytest.mark.parametrize("K,R", [(2, 8), (4, 0), (4, 4)])
def test_1080p_every_scale(K, R):
    """1920x1080 at every scale: at the finest scales each workgroup takes
    several tiles per chunk."""
    print("nbr 1080p", K, R, flush=True)
    cfg = synth.CONFIGS["tvl1_1920x1080"]
    q0, q1 = _pair(cfg["width"], cfg["height"], cfg["cap"], 2024)
    u, it = _solve(q0, q1, [(0, K, R, NBR)], nscales=5)
    uo, ito = _oracle(q0, q1, nscales=5)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_mixed_modes_per_scale():
    """Host-paced, barrier and neighbour-flag warpings in one solve."""
    q0, q1 = _pair(640, 480, 8.0, 23)
    rules = [(200000, 2, 8, 0), (60000, 4, 4, 1), (15000, 4, 0, NBR), (0, 2, 0, NBR)]
    u, it = _solve(q0, q1, rules, nscales=5)
    uo, ito = _oracle(q0, q1, nscales=5)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_repeated_solves_on_one_handle():
    """Flags and decisions are tagged per warping: later solves on the same
    handle (and so on the same flag words) stay exact."""
    q0, q1 = _pair(320, 240, 5.0, 29)
    q2, q3 = _pair(320, 240, 5.0, 31)
    with TV_L1(320, 240, nscales=4) as tv:
        tv.set_launch_rule(0, 4, 0, NBR)
        for a, b in ((q0, q1), (q2, q3), (q0, q1)):
            tv.runDualTVL1Multiscale(a, b)
            uo, ito = _oracle(a, b, nscales=4)
            assert np.array_equal(tv.iterations(), ito)
            assert np.array_equal(tv.getU(), uo)


def test_not_resident_takes_the_host_paced_repeat():
    """The first wait of every tile workgroup reports the grid as not
    resident (test hook; every workgroup still ends normally): the solve is
    repeated host-paced with the oracle's bits."""
    try:
        set_persist_launch_mode(PERSIST_PLAIN)
        D.persist_force_not_resident(True)
        q0, q1 = _pair(320, 240, 6.0, 42)
        u, it = _solve(q0, q1, [(0, 4, 0, NBR)], nscales=5)
        assert persist_launch_mode() == PERSIST_HOST_PACED
        uo, ito = _oracle(q0, q1, nscales=5)
        assert np.array_equal(it, ito)
        assert np.array_equal(u, uo)
    finally:
        D.persist_force_not_resident(False)
        set_persist_launch_mode(PERSIST_PLAIN)


def test_handles_created_after_destroyed_ones():
    """A new handle may receive device memory that an earlier handle's
    control block and flags used: the warpings are numbered per process, so
    no word an earlier warping left there passes for the current one's."""
    q0, q1 = _pair(200, 150, 4.0, 37)
    uo, ito = _oracle(q0, q1, nscales=4)
    for k in range(4):
        u, it = _solve(q0, q1, [(0, (1, 2, 4, 1)[k], (4, 0, 8, 0)[k], NBR)], nscales=4)
        assert np.array_equal(it, ito), k
        assert np.array_equal(u, uo), k
