"""The lightcone drift on the device (ct_drift_lightcone) against the oracle's
restatement of src/drift.cu:75-131: the same crossings, bit for bit in
position and velocity; the same HEALPix pixel except where a direction lies
on a pixel boundary to within the atan2 rounding; and the same drift as
ct_drift."""
import numpy as np
import pytest

from test_lightcone_oracle import _shell_box

pytestmark = pytest.mark.gpu


def _sorted(e):
    return e[np.lexsort((e["z"], e["y"], e["x"]))]


def _boundary_case(oc, nside, e, pix):
    """pix is among the pixels of directions within ~1e-7 of the entry's."""
    X = np.array([e["x"], e["y"], e["z"]], dtype=np.float64) / 2.0 ** 31
    d = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64) * 1e-7
    return pix in set(oc.vec2pix_nest(nside, X[None, :] + d).tolist())


@pytest.mark.parametrize("n,dt,nside", [(30000, 0.02, 16), (200000, 0.05, 256)])
def test_crossings_match_oracle(gpu, oc, n, dt, nside):
    ct = gpu
    x, y, z, vel, rung = _shell_box(n, 4)
    tau_max = 1.5
    tau1 = tau_max - 0.8
    tau0 = tau1 - dt
    print("lightcone n=%d dt=%g nside=%d" % (n, dt, nside), flush=True)
    want = oc.lightcone(x, y, z, vel, rung, 0, 1.0, dt, tau0, tau1, tau_max, nside)
    box = ct.Box(x, y, z, vel, rung)
    got = box.drift_lightcone(0, 1.0, dt, tau0, tau1, tau_max, nside)
    assert len(got) == len(want) > 4096
    g, w = _sorted(got), _sorted(want)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(g[f], w[f]), f
    bad = np.nonzero(g["pix"] != w["pix"])[0]
    assert len(bad) <= max(2, len(g) // 1000)
    for i in bad:
        assert _boundary_case(oc, nside, w[i], int(g["pix"][i])), (i, g[i], w[i])
    # the drift itself is ct_drift's, and the oracle's
    d = box.read()
    ox, oy, oz = oc.drift(x, y, z, vel, rung, 0, 1.0, dt)
    idx = d["id"]
    assert np.array_equal(d["x"], ox[idx]) and np.array_equal(d["y"], oy[idx]) and np.array_equal(d["z"], oz[idx])
    box.close()


def test_off_outside_the_last_unit_and_argument_checks(gpu, oc):
    ct = gpu
    x, y, z, vel, rung = _shell_box(5000, 5)
    box = ct.Box(x, y, z, vel, rung)
    # tau_max - tau1 >= 1: no lightcone, but the drift happens
    got = box.drift_lightcone(0, 1.0, 0.02, 0.1, 0.12, 1.2, 16)
    assert len(got) == 0
    d = box.read()
    ox, _, _ = oc.drift(x, y, z, vel, rung, 0, 1.0, 0.02)
    assert np.array_equal(d["x"], ox[d["id"]])
    # nested pixels need a power-of-two nside
    with pytest.raises(ct.CTError):
        box.drift_lightcone(0, 1.0, 0.02, 0.68, 0.7, 1.5, 12)
    box.close()


def test_only_the_active_range_crosses(gpu, oc):
    """After particles_sort_by_rung the lightcone, like the drift, covers only
    the active range [rung_begin, N) (src/drift.cu:170-177)."""
    ct = gpu
    n = 30000
    x, y, z, vel, _ = _shell_box(n, 6)
    rung = (np.arange(n) % 3).astype(np.int8)  # rungs 0, 1, 2
    box = ct.Box(x, y, z, vel, rung)
    rb = box.sort_by_rung(1)
    assert rb == n - np.count_nonzero(rung >= 1)
    dt, tau_max = 0.02, 1.5
    tau1 = tau_max - 0.8
    # rung 0 lies outside the active range: no crossings (and no drift)
    assert oc.lightcone(x, y, z, vel, rung, 0, 1.0, dt, tau1 - dt, tau1, tau_max, 16).size > 0
    assert len(box.drift_lightcone(0, 1.0, dt, tau1 - dt, tau1, tau_max, 16)) == 0
    got = box.drift_lightcone(1, 1.0, dt, tau1 - dt, tau1, tau_max, 16)
    # the oracle over the rung-1 particles only: rung-0 particles are outside
    # the range and rung-2 ones are not this rung
    want = oc.lightcone(x, y, z, vel, rung, 1, 1.0, dt, tau1 - dt, tau1, tau_max, 16)
    assert len(got) == len(want) > 0
    g, w = _sorted(got), _sorted(want)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(g[f], w[f]), f
    box.close()
