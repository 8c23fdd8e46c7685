"""GPU checks of the active range, the rung bookkeeping and the hierarchical
step against the oracle (particles_sort_by_rung / current_range /
push_rungs / pop_rungs / set_minrung / rung_counts, src/particles.cpp;
kick_step_hierarchical, src/driver.cpp:382-593).

Bars:
  * rung partition: the same particles in the same order as the oracle's
    stable partition, the same range start (index work: exact);
  * tree over the active range: host and device builds give the same nodes
    (integer and fixed32 fields exact) and the oracle's tree over the active
    particles alone (offsets of the range start aside);
  * kicks of a sub-range at min_rung > 0, not top, descending: forces,
    velocities and rungs as the default kick's parity test (forces rms
    relative 2e-5, rungs exact where the oracle's log2(t0/dt) is 1e-4 away
    from an integer);
  * the ladder over a whole cycle: the same (minrung, max_rung) per call as
    the oracle's restatement, the same rung per particle, velocities and
    displacements within 1e-4 relative rms (FP32 forces through ~20 kicks).
"""
import numpy as np
import pytest

from conftest import clustered_box

pytestmark = pytest.mark.gpu


def _state(n, seed):
    import cosmictiger_amd as ct
    from cosmictiger_amd import workloads
    x, y, z = clustered_box(n, seed, nhalo=10)
    vel = workloads.gaussian_velocities(n, seed, sigma=1e-3)
    return x, y, z, vel, ct.softening(n)


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


def test_sort_by_rung_and_range_stack(gpu, oc):
    n = 5000
    rng = np.random.default_rng(3)
    x, y, z, vel, h = _state(n, 3)
    rung = rng.integers(0, 6, n).astype(np.int8)
    box = gpu.Box(x, y, z, vel, rung)
    ob = oc.OracleBox(x, y, z, vel, rung)
    assert box.current_range() == (0, n)
    for m in (2, 4, 5):
        box.push_rungs()
        ob.push_rungs()
        assert box.sort_by_rung(m) == ob.sort_by_rung(m)
        d = box.read()
        assert np.array_equal(d["id"], ob.id)
        assert np.array_equal(d["rung"], ob.rung)
        assert np.array_equal(d["x"], ob.x) and np.array_equal(d["vel"], ob.vel)
        assert box.current_range() == (ob.rung_begin, n)
    for _ in range(3):
        box.pop_rungs()
        ob.pop_rungs()
        assert box.current_range()[0] == ob.rung_begin
    with pytest.raises(gpu.CTError):
        box.pop_rungs()
    assert box.sort_by_rung(0) == 0
    counts = box.rung_counts()
    assert np.array_equal(counts, np.bincount(rung))
    box.set_minrung(3)
    ob.set_minrung(3)
    assert np.array_equal(_by_id(box.read())["rung"], ob.by_id()["rung"])
    box.close()


def test_drift_covers_the_active_range_only(gpu):
    n = 4000
    x, y, z, vel, h = _state(n, 4)
    rung = (np.arange(n) % 3).astype(np.int8)
    box = gpu.Box(x, y, z, vel, rung)
    box.sort_by_rung(2)
    before = _by_id(box.read())
    box.drift(1, 1.0, 0.01)  # rung-1 particles lie below the range: nothing moves
    mid = _by_id(box.read())
    assert np.array_equal(before["x"], mid["x"])
    box.drift(2, 1.0, 0.01)  # every particle of the range moves
    after = _by_id(box.read())
    moved = after["x"] != mid["x"]
    assert np.array_equal(moved, (before["rung"] == 2) & (before["vel"][:, 0] != 0))
    box.close()


@pytest.mark.parametrize("where", ["device", "host"])
def test_tree_over_active_range(gpu, oc, where):
    n = 20000
    rng = np.random.default_rng(5)
    x, y, z, vel, h = _state(n, 5)
    rung = rng.integers(0, 3, n).astype(np.int8)
    box = gpu.Box(x, y, z, vel, rung)
    rb = box.sort_by_rung(1)
    assert 0 < rb < n
    box.tree_build_ex(0.55, 64, h, gpu.TREE_DEVICE if where == "device" else gpu.TREE_HOST)
    nd = box.nodes()
    assert nd["pbegin"][0] == rb and nd["pend"][0] == n
    d = box.read()
    T = oc.OracleTree(d["x"][rb:], d["y"][rb:], d["z"][rb:], 64, 0.55, h)
    on = T.node_arrays()
    assert len(on) == len(nd)
    # the oracle's nodes are in its depth-first order; match by (depth, box)
    def key(a, off):
        return [(int(a["depth"][i]), tuple(a["box_begin"][i]), tuple(a["box_end"][i]), int(a["pbegin"][i]) - off,
                 int(a["pend"][i]) - off, tuple(a["pos"][i]), float(a["radius"][i])) for i in range(len(a))]
    assert sorted(key(nd, rb)) == sorted(key(on, 0))
    box.close()


def test_kick_of_a_sub_range_matches_oracle(gpu, oc):
    """A descending, non-top kick at min_rung 1 over the particles of rung >= 1
    (the second kick of the ladder's first call)."""
    n = 20000
    rng = np.random.default_rng(6)
    x, y, z, vel, h = _state(n, 6)
    rung = rng.integers(0, 3, n).astype(np.int8)
    box = gpu.Box(x, y, z, vel, rung)
    ob = oc.OracleBox(x, y, z, vel, rung)
    box.sort_by_rung(1)
    ob.sort_by_rung(1)
    box.tree_build(0.55, 64, h)
    p = gpu.default_kick_params(0.55, h, 1.0 / n)
    p.min_rung = 1
    p.top = False
    p.ascending = False
    p.descending = True
    p.first_call = False
    p.t0 = 2e-3
    q = oc.default_params(0.55, h, 1.0 / n)
    for f in ("min_rung", "top", "ascending", "descending", "first_call", "t0"):
        setattr(q, f, getattr(p, f))
    kr = box.kick(p)
    okr = ob.tree_kick(q, 0.55, 64, h)
    g = _by_id(box.read())
    o = ob.by_id()
    act = np.zeros(n, bool)
    act[ob.forces["id"]] = True
    fo = {k: np.zeros(n) for k in ("gx", "gy", "gz")}
    for k in fo:
        fo[k][ob.forces["id"]] = ob.forces[k]
    go = np.stack([fo["gx"], fo["gy"], fo["gz"]], 1)[act]
    gg = np.stack([g["gx"], g["gy"], g["gz"]], 1)[act].astype(np.float64)
    err = np.linalg.norm(gg - go, axis=1) / np.linalg.norm(go, axis=1)
    print("sub-range kick: %d active, force rel err rms %.3e" % (act.sum(), np.sqrt(np.mean(err ** 2))))
    assert np.sqrt(np.mean(err ** 2)) <= 2e-5
    # particles outside the range are untouched
    assert np.array_equal(g["vel"][~act], np.asarray(vel, np.float32).reshape(n, 3)[~act])
    dv = np.linalg.norm(g["vel"][act].astype(np.float64) - o["vel"][act], axis=1)
    dv0 = np.linalg.norm(o["vel"][act].astype(np.float64) - np.asarray(vel, np.float64).reshape(n, 3)[act], axis=1)
    assert np.sqrt(np.mean(dv ** 2)) <= 1e-4 * np.sqrt(np.mean(dv0 ** 2)) + 1e-12
    assert np.array_equal(g["rung"], o["rung"])
    assert kr.max_rung == okr.max_rung
    assert abs(kr.kin - okr.kin) <= 1e-4 * abs(okr.kin)
    box.close()


def test_ladder_over_a_cycle_matches_oracle(gpu, oc):
    n = 12000
    x, y, z, vel, h = _state(n, 7)
    eta = 0.25 / np.sqrt(693.0 / 256.0)
    # top step: 6 times the smallest time step of the first kick (rungs 0..3)
    T = oc.OracleTree(x, y, z, 64, 0.55, h)
    f, _, _ = oc.kick(T, oc.default_params(0.55, h, 1.0 / n), np.zeros((n, 3), np.float32), np.zeros(n, np.int8))
    t0 = 6.0 * float(np.min(eta * np.sqrt(h / np.sqrt(f["gx"] ** 2 + f["gy"] ** 2 + f["gz"] ** 2))))
    box = gpu.Box(x, y, z, vel, np.zeros(n, np.int8))

    def gpu_step(minrung, max_rung, minrung0, tau):
        sp = gpu.StepParams(minrung=minrung, max_rung=max_rung, minrung0=minrung0, tau=tau, t0=t0, theta=0.55, bucket=64,
                            hsoft=h, GM=1.0 / n, eta=eta, nocrop=1)
        r = box.step_hierarchical(sp)
        return r.kr, r.minrung

    ob = oc.OracleBox(x, y, z, vel, np.zeros(n, np.int8))
    calls_o, tau_o = oc.run_steps(ob, 64, t0, 0.55, 64, h, 1.0 / n, eta)
    calls_g, tau_g = oc.run_steps(None, 64, t0, 0.55, 64, h, 1.0 / n, eta, step=gpu_step)
    print("calls", calls_g)
    assert calls_g == calls_o
    assert max(m for _, m in calls_g) >= 2
    g = _by_id(box.read())
    o = ob.by_id()
    assert np.count_nonzero(g["rung"] != o["rung"]) == 0
    v0 = np.asarray(vel, np.float64).reshape(n, 3)
    dv = np.linalg.norm(g["vel"].astype(np.float64) - o["vel"], axis=1)
    dv0 = np.linalg.norm(o["vel"] - v0, axis=1)
    print("velocity change rel err rms %.3e" % (np.sqrt(np.mean(dv ** 2)) / np.sqrt(np.mean(dv0 ** 2))))
    assert np.sqrt(np.mean(dv ** 2)) <= 1e-4 * np.sqrt(np.mean(dv0 ** 2))

    def disp(a):
        return np.stack([((a[k].astype(np.int64) - x0.astype(np.int64) + 2 ** 31) % 2 ** 32) - 2 ** 31
                         for k, x0 in (("x", x), ("y", y), ("z", z))], 1).astype(np.float64)
    dg, do = disp(g), disp(o)
    e = np.linalg.norm(dg - do, axis=1)
    print("displacement rel err rms %.3e" % (np.sqrt(np.mean(e ** 2)) / np.sqrt(np.mean(np.sum(do ** 2, 1)))))
    assert np.sqrt(np.mean(e ** 2)) <= 1e-4 * np.sqrt(np.mean(np.sum(do ** 2, 1)))
    box.close()


def _read_reference_layout(path, offset=0):
    """The particles_save record (src/particles.cpp:1206-1227), parsed by hand."""
    with open(path, "rb") as f:
        f.seek(offset)
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        x = np.frombuffer(f.read(4 * n), np.uint32)
        y = np.frombuffer(f.read(4 * n), np.uint32)
        z = np.frombuffer(f.read(4 * n), np.uint32)
        vel = np.frombuffer(f.read(12 * n), np.float32).reshape(n, 3)
        rung = np.frombuffer(f.read(n), np.int8)
        ns, rb = np.frombuffer(f.read(8), np.int32)
        stack = np.frombuffer(f.read(4 * int(ns)), np.int32)
        return {"x": x, "y": y, "z": z, "vel": vel, "rung": rung, "rung_begin": int(rb), "stack": stack.tolist(), "end": f.tell()}


def test_box_save_load_round_trip(gpu, tmp_path):
    n = 6000
    rng = np.random.default_rng(8)
    x, y, z, vel, h = _state(n, 8)
    rung = rng.integers(0, 4, n).astype(np.int8)
    box = gpu.Box(x, y, z, vel, rung)
    box.push_rungs()
    box.sort_by_rung(1)
    box.push_rungs()
    box.sort_by_rung(3)
    box.tree_build(0.55, 64, h)
    path = str(tmp_path / "box.dat")
    header = np.arange(7, dtype=np.float64)  # stands for a checkpoint header in front of the record
    header.tofile(path)
    box.save(path, append=True)
    rec = _read_reference_layout(path, header.nbytes)
    d = box.read()
    for k in ("x", "y", "z", "vel", "rung"):
        assert np.array_equal(rec[k], d[k]), k
    first, _ = box.current_range()
    assert rec["rung_begin"] == first and len(rec["stack"]) == 2
    b2, end = gpu.Box.load(path, header.nbytes)
    assert end == rec["end"]
    d2 = b2.read()
    for k in ("x", "y", "z", "vel", "rung"):
        assert np.array_equal(d2[k], d[k]), k
    assert np.array_equal(d2["id"], np.arange(n))
    assert b2.current_range() == box.current_range()
    # the restored box steps exactly as the original: same tree, same kick
    b2.tree_build(0.55, 64, h)
    p = gpu.default_kick_params(0.55, h, 1.0 / n)
    p.min_rung = 3
    r1, r2 = box.kick(p), b2.kick(p)
    assert r1.as_dict() == r2.as_dict()
    e1, e2 = box.read(), b2.read()
    o1, o2 = np.argsort(e1["x"].astype(np.uint64) * 2 ** 32 + e1["y"]), np.argsort(e2["x"].astype(np.uint64) * 2 ** 32 + e2["y"])
    assert np.array_equal(e1["vel"][o1], e2["vel"][o2])
    b2.pop_rungs()
    b2.pop_rungs()
    assert b2.current_range()[0] == 0
    box.close()
    b2.close()


def test_box_load_rejects_corrupt_files(gpu, tmp_path):
    path = str(tmp_path / "bad.dat")
    np.array([5], np.int64).tofile(path)  # count without data
    with pytest.raises(gpu.CTError):
        gpu.Box.load(path)
    with pytest.raises(gpu.CTError):
        gpu.Box.load(str(tmp_path / "missing.dat"))
