"""Batched kicks: every box of a batch gets exactly what it gets when kicked
alone (each box's arithmetic is the same sequence of operations whichever
other boxes share the launches), eagerly and from a replayed HIP graph; the
batch drift equals the single-box drift bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vel", "rung", "gx", "gy", "gz", "phi")


def _boxes(ct):
    from cosmictiger_amd import workloads
    specs = [(9000, 0.0, 3, 0.7), (30000, 0.6, 4, 0.55), (1, 0.0, 5, 0.7), (15000, 0.8, 6, 0.4), (4096, 0.3, 7, 0.6)]
    out = []
    for n, frac, seed, theta in specs:
        if frac > 0:
            x, y, z = workloads.nfw_box(n, seed, clustered_fraction=frac, nhalos=12)
        else:
            x, y, z = workloads.uniform_box(n, seed)
        vel = workloads.gaussian_velocities(n, seed, sigma=0.02)
        out.append((x, y, z, vel, theta))
    return out


def _make(ct, spec, bucket):
    x, y, z, vel, theta = spec
    n = len(x)
    b = ct.Box(x, y, z, vel, np.zeros(n, np.int8))
    h = ct.softening(n)
    b.tree_build(theta, bucket, h)
    return b, ct.default_kick_params(theta, h, 1.0 / n)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_alone(gpu, graph):
    specs = _boxes(gpu)
    alone = []
    for sp in specs:
        b, p = _make(gpu, sp, 64)
        r = b.kick(p)
        alone.append((b.read(), r))
        b.close()
    boxes, params = zip(*[_make(gpu, sp, 64) for sp in specs])
    bt = gpu.Batch(boxes)
    rets = bt.kick(params, graph=graph)
    if graph:
        # second kick of the same boxes from the captured graph, against a second eager kick alone
        rets = bt.kick(params, graph=True)
        alone2 = []
        for sp in specs:
            b, p = _make(gpu, sp, 64)
            b.kick(p)
            r = b.kick(p)
            alone2.append((b.read(), r))
            b.close()
        alone = alone2
    for i, b in enumerate(boxes):
        got = b.read()
        ref, rref = alone[i]
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (i, f)
        for f in ("max_rung", "pot", "kin", "dkin", "xmom", "ymom", "zmom", "nmom"):
            assert getattr(rets[i], f) == getattr(rref, f), (i, f)
    bt.close()
    for b in boxes:
        b.close()


def test_batch_drift_equals_alone(gpu):
    specs = _boxes(gpu)[:3]
    alone = []
    for sp in specs:
        b, p = _make(gpu, sp, 64)
        b.drift(0, 0.8, 0.01, 0.1)
        alone.append(b.read())
        b.close()
    boxes, _ = zip(*[_make(gpu, sp, 64) for sp in specs])
    bt = gpu.Batch(boxes)
    bt.drift(0, [0.8] * 3, [0.01] * 3, [0.1] * 3)
    for i, b in enumerate(boxes):
        got = b.read()
        for f in ("x", "y", "z", "vel"):
            assert np.array_equal(got[f], alone[i][f]), (i, f)
    bt.close()
    for b in boxes:
        b.close()


def test_batch_rebuild_and_errors(gpu):
    """A tree rebuilt between batch kicks is picked up; bad arguments fail."""
    specs = _boxes(gpu)[:2]
    boxes, params = zip(*[_make(gpu, sp, 64) for sp in specs])
    bt = gpu.Batch(boxes)
    bt.kick(params, graph=True)
    boxes[0].tree_build(0.5, 32, gpu.softening(boxes[0].n))
    p0 = gpu.default_kick_params(0.5, gpu.softening(boxes[0].n), 1.0 / boxes[0].n)
    rets = bt.kick([p0, params[1]], graph=True)
    ref, rp = _make(gpu, specs[0], 32)
    ref.tree_build(0.5, 32, gpu.softening(ref.n))
    r_alone = ref.kick(p0)
    # same tree, particles in a different order inside the leaves: FP32 sums
    # in another order (potential bar 1e-5 GM N per particle, GM N = 1)
    assert abs(rets[0].pot - r_alone.pot) <= 0.5 * ref.n * 1e-5
    with pytest.raises(gpu.CTError):
        gpu.Batch([boxes[0], boxes[0]])
    bt.close()
    ref.close()
    for b in boxes:
        b.close()


def test_batch_reset_equals_fresh_batch(gpu):
    """A batch pointed at another set of boxes (ct_batch_reset, arenas kept)
    kicks them exactly as a new batch of those boxes does."""
    specs = _boxes(gpu)
    first = [_make(gpu, sp, 64) for sp in specs[:2]]
    second_a = [_make(gpu, sp, 64) for sp in specs[2:]]
    second_b = [_make(gpu, sp, 64) for sp in specs[2:]]
    bt = gpu.Batch([b for b, _ in first])
    bt.kick([p for _, p in first])
    bt.reset([b for b, _ in second_a])
    ra = bt.kick([p for _, p in second_a])
    bt.close()
    fresh = gpu.Batch([b for b, _ in second_b])
    rb = fresh.kick([p for _, p in second_b])
    fresh.close()
    for (ba, _), (bb, _), xa, xb in zip(second_a, second_b, ra, rb):
        assert xa.as_dict() == xb.as_dict()
        da, db = ba.read(), bb.read()
        for f in FIELDS:
            assert np.array_equal(da[f], db[f]), f
    for b, _ in first + second_a + second_b:
        b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_step_equals_kick_then_drift(gpu, graph):
    """ct_batch_step (kick and drift, one captured graph with graph=True) over
    three steps with a different dt each time equals a batch kick followed by
    a batch drift, bit for bit, and replays see the new drift fields."""
    specs = _boxes(gpu)[:3]
    A = [_make(gpu, sp, 64) for sp in specs]
    B = [_make(gpu, sp, 64) for sp in specs]
    ba = gpu.Batch([b for b, _ in A])
    bb = gpu.Batch([b for b, _ in B])
    for k, dt in enumerate((1e-3, 2e-3, 5e-4)):
        pa = [p for _, p in A]
        for p in pa:
            p.first_call = k == 0
        ra = ba.step(pa, 0, 1.0, dt, graph=graph)
        rb = bb.kick(pa)
        bb.drift(0, 1.0, dt)
        for xa, xb in zip(ra, rb):
            assert xa.as_dict() == xb.as_dict()
        for (x, _), (y, _) in zip(A, B):
            da, db = x.read(), y.read()
            for f in FIELDS:
                assert np.array_equal(da[f], db[f]), (k, f)
    ba.close()
    bb.close()
    for b, _ in A + B:
        b.close()
