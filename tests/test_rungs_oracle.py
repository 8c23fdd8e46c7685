"""CPU checks of the oracle's active range and kick ladder (the restatement of
particles_sort_by_rung, src/particles.cpp:1029-1068, and
kick_step_hierarchical, src/driver.cpp:382-593).

Pins:
  * the rung partition keeps exactly the particles the reference's in-place
    two-pointer loop keeps (restated here literally, on small inputs), and
    the same range start;
  * the block time-step scheme: over a whole cycle of the integer time every
    particle is drifted for exactly the cycle's duration, whatever rungs it
    visited (a dropped, doubled or mis-sized drift breaks it), with and
    without the clipping of the top rung;
  * the order of the ladder on a known case: ascending kicks precede the top
    kick, descending kicks push the range and follow it.
"""
import numpy as np
import pytest

from conftest import clustered_box


def two_pointer_partition(rung, begin, min_rung):
    """The loop of src/particles.cpp:1046-1063 as written there."""
    r = list(rung)
    idx = list(range(len(r)))
    lo, hi = begin, len(r)
    while lo < hi:
        if r[lo] >= min_rung:
            while lo != hi:
                hi -= 1
                if r[hi] < min_rung:
                    r[hi], r[lo] = r[lo], r[hi]
                    idx[hi], idx[lo] = idx[lo], idx[hi]
                    break
        lo += 1
    return hi, idx


@pytest.mark.parametrize("seed", range(6))
def test_rung_partition_matches_two_pointer(oc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 400))
    rung = rng.integers(0, 5, n).astype(np.int8)
    box = oc.OracleBox(np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros((n, 3)), rung)
    begin = int(rng.integers(0, n + 1))
    box.rung_begin = begin
    for m in (int(rng.integers(1, 5)), int(rng.integers(1, 5))):
        prev = box.rung_begin
        ref_begin, ref_idx = two_pointer_partition(box.rung, prev, m)
        ids_before = box.id.copy()
        got = box.sort_by_rung(m)
        assert got == ref_begin
        # same members above and below the new start, untouched below the old start
        assert np.array_equal(np.sort(box.id[got:]), np.sort(ids_before[np.array(ref_idx[got:], dtype=np.int64)]))
        assert np.array_equal(box.id[:prev], ids_before[:prev])
        assert np.all(box.rung[got:] >= m) and np.all(box.rung[prev:got] < m)
    assert box.sort_by_rung(0) == 0


def _setup(n, seed):
    import cosmictiger_amd as ct
    from cosmictiger_amd import workloads
    x, y, z = clustered_box(n, seed, nhalo=10)
    vel = workloads.gaussian_velocities(n, seed, sigma=1e-3)
    h = ct.softening(n)
    eta = 0.25 / np.sqrt(693.0 / 256.0)
    return x, y, z, vel, h, eta


def _t0_for_rungs(oc, x, y, z, h, n, factor):
    """A top step factor times the smallest time step of the first kick, so
    that the rungs spread over 0..ceil(log2 factor)."""
    T = oc.OracleTree(x, y, z, 32, 0.55, h)
    p = oc.default_params(0.55, h, 1.0 / n)
    f, _, _ = oc.kick(T, p, np.zeros((n, 3), np.float32), np.zeros(n, np.int8))
    g = np.sqrt(f["gx"] ** 2 + f["gy"] ** 2 + f["gz"] ** 2)
    return factor * float(np.min(0.25 / np.sqrt(693.0 / 256.0) * np.sqrt(h / g)))


@pytest.mark.parametrize("nocrop,cycles", [(True, 1), (False, 3)])
def test_ladder_drifts_every_particle_for_the_whole_cycle(oc, nocrop, cycles):
    n = 3000
    x, y, z, vel, h, eta = _setup(n, 7)
    t0 = _t0_for_rungs(oc, x, y, z, h, n, 6.0)
    box = oc.OracleBox(x, y, z, vel, np.zeros(n, np.int8))
    trace = []
    calls, tau = oc.run_steps(box, 200, t0, 0.55, 32, h, 1.0 / n, eta, nocrop=nocrop, trace=trace, cycles=cycles)
    assert abs(tau - cycles * t0) <= 1e-12 * t0
    bi = box.by_id()
    # the drifts of each particle tile the cycles: sum = cycles * t0 exactly
    # (dt = 2^-rung t0 in float, sums of such are exact)
    assert np.all(bi["tdrift"] == cycles * float(np.float32(t0)))
    # several rungs were visited
    assert max(m for _, m in calls) >= 2
    assert len(trace) > len(calls)
    if not nocrop:
        # the top rung was clipped at a later cycle: rung 0 emptied
        assert np.all(bi["rung"] >= 1)


def test_ladder_order_on_a_known_case(oc):
    """One cycle of a box whose first kick puts particles on rungs 0..3: the
    first call kicks the top (0) and then descends 1, 2, 3, pushing the range
    before each descending sort; the call at minrung 2 kicks 3 ascending
    (popping), then the top at 2, then descends to 3."""
    n = 3000
    x, y, z, vel, h, eta = _setup(n, 7)
    t0 = _t0_for_rungs(oc, x, y, z, h, n, 6.0)
    box = oc.OracleBox(x, y, z, vel, np.zeros(n, np.int8))
    trace = []
    calls, _ = oc.run_steps(box, 3, t0, 0.55, 32, h, 1.0 / n, eta, trace=trace)
    assert [c[0] for c in calls] == [0, 3, 2]
    levels = [(t[0], t[1], t[2]) for t in trace]
    # call 1: top at 0, descending 1..3; call 2: top at 3; call 3: ascending 3, top 2, descending 3
    assert levels == [(0, False, True), (1, False, False), (2, False, False), (3, False, False), (3, False, True),
                      (3, True, False), (2, False, True), (3, False, False)]
    # ascending kicks neither drift nor descend; every other kick drifts
    for lv, asc, top, desc, _, drifted in trace:
        assert drifted == (not asc)
        assert desc == (not asc)
    # the active range only shrinks while descending
    rb = [t[4] for t in trace[:4]]
    assert rb == sorted(rb) and rb[0] == 0
