"""CPU checks of the oracle's HEALPix nested pixels (src/healpix.cu:36-172)
and of its lightcone crossings of the drift (src/drift.cu:75-131).

Pins:
  * pixels: the 12 base-pixel centres of HEALPix; pix2vec then vec2pix gives
    back every pixel at several resolutions; the nested hierarchy (the pixel at
    nside is the pixel at 2 nside divided by 4); every direction lies within a
    pixel radius of its pixel's centre; uniform directions fill the pixels
    equally (equal areas);
  * crossings: every entry lies on the light shell |X(t)| = t_i - tau0 - t; a
    particle moving radially crosses where the closed form puts it; a particle
    far from the shell gives none; velocities are stored unscaled.
"""
import numpy as np
import pytest


def test_base_pixel_centres(oc):
    c = oc.pix2vec_nest(1, np.arange(12))
    z = np.repeat([2.0 / 3.0, 0.0, -2.0 / 3.0], 4)
    phi = np.concatenate([np.pi / 4 + np.arange(4) * np.pi / 2, np.arange(4) * np.pi / 2, np.pi / 4 + np.arange(4) * np.pi / 2])
    st = np.sqrt(1 - z * z)
    want = np.stack([st * np.cos(phi), st * np.sin(phi), z], axis=1)
    assert np.allclose(c, want, atol=1e-15)
    assert np.array_equal(oc.vec2pix_nest(1, want), np.arange(12))


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32, 128])
def test_every_pixel_round_trips(oc, nside):
    pix = np.arange(12 * nside * nside)
    assert np.array_equal(oc.vec2pix_nest(nside, oc.pix2vec_nest(nside, pix)), pix)


def _directions(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_nested_hierarchy(oc):
    v = _directions(20000, 1)
    nside = 1
    coarse = oc.vec2pix_nest(nside, v)
    while nside < 8192:
        fine = oc.vec2pix_nest(2 * nside, v)
        assert np.array_equal(fine // 4, coarse), nside
        coarse = fine
        nside *= 2


@pytest.mark.parametrize("nside", [1, 4, 64])
def test_direction_near_its_pixel_centre(oc, nside):
    v = _directions(20000, 2)
    c = oc.pix2vec_nest(nside, oc.vec2pix_nest(nside, v))
    ang = np.arccos(np.clip(np.sum(v * c, axis=1), -1, 1))
    # a HEALPix pixel spans at most ~1.36 / nside radians corner to centre
    assert ang.max() < 1.4 / nside


def test_equal_areas(oc):
    nside = 4
    npix = 12 * nside * nside
    n = 400000
    counts = np.bincount(oc.vec2pix_nest(nside, _directions(n, 3)), minlength=npix)
    assert counts.sum() == n and len(counts) == npix
    expect = n / npix
    chi2 = np.sum((counts - expect) ** 2 / expect)
    # chi-square with npix - 1 = 191 degrees of freedom: mean 191, sd 19.5
    assert chi2 < 191 + 6 * 19.5


def _shell_box(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    z = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    vel = rng.normal(scale=0.05, size=(n, 3)).astype(np.float32)
    rung = np.zeros(n, np.int8)
    return x, y, z, vel, rung


def test_entries_lie_on_the_light_shell(oc):
    x, y, z, vel, rung = _shell_box(30000, 4)
    a, dt, tau_max = 1.0, 0.02, 1.5
    tau1 = tau_max - 0.8
    tau0 = tau1 - dt
    e = oc.lightcone(x, y, z, vel, rung, 0, a, dt, tau0, tau1, tau_max, 16)
    assert len(e) > 50
    X = np.stack([e["x"], e["y"], e["z"]], axis=1).astype(np.float64) / 2.0 ** 31
    ti = np.float64(np.float32(tau_max))  # i0 = 0: the shell of tau_max
    resid = np.linalg.norm(X, axis=1) - (ti - np.float64(np.float32(tau0)) - e["t"])
    assert np.abs(resid).max() < 1e-8
    assert (e["t"] >= -1e-12).all() and (e["t"] <= dt * 1.01).all()
    assert (np.linalg.norm(X, axis=1) <= 1.0 + 1e-9).all()
    # pixels are those of the entries' directions
    assert np.array_equal(e["pix"], oc.vec2pix_nest(16, X))


def test_radial_mover_closed_form(oc):
    # one particle at p0, moving radially away from the origin image at speed u
    p0 = np.array([0.375, 0.25, 0.125])
    r0 = np.linalg.norm(p0)
    nhat = p0 / r0
    u = 0.25
    vel = (u * nhat).astype(np.float32)[None, :]
    fx = lambda c: np.array([np.uint32(c * 2 ** 32)])
    tau_max = 1.0
    t_cross = 0.01
    # the shell reaches the particle t_cross after tau0: tau0 + t + |p0 + u t| = tau_max
    tau0 = np.float32(tau_max - r0 - (1 + u) * t_cross)
    dt = np.float32(0.03)
    tau1 = np.float32(tau0 + dt)
    e = oc.lightcone(fx(p0[0]), fx(p0[1]), fx(p0[2]), vel, np.zeros(1, np.int8), 0, 1.0, dt, tau0, tau1, tau_max, 8)
    assert len(e) == 1
    v = vel[0].astype(np.float64)
    x0 = np.array([np.float64(np.uint32(c * 2 ** 32)) / 2 ** 32 for c in p0])
    # closed form for motion along x0 (v is parallel to x0 up to float rounding)
    t_exact = (tau_max - np.float64(tau0) - np.linalg.norm(x0)) / (1 + np.linalg.norm(v))
    assert abs(e["t"][0] - t_exact) < 1e-6
    X = np.array([e["x"][0], e["y"][0], e["z"][0]]) / 2.0 ** 31
    assert np.allclose(X, x0 + v * e["t"][0], atol=2e-9)
    assert np.array_equal([e["vx"][0], e["vy"][0], e["vz"][0]], vel[0])
    # a particle whose every image is far from the shell gives nothing
    far = oc.lightcone(fx(0.5), fx(0.5), fx(0.5), np.zeros((1, 3), np.float32), np.zeros(1, np.int8), 0, 1.0, dt, tau0, tau1,
                       tau_max, 8)
    assert len(far) == 0
    # outside the last unit of conformal time there is no lightcone
    none = oc.lightcone(fx(p0[0]), fx(p0[1]), fx(p0[2]), vel, np.zeros(1, np.int8), 0, 1.0, dt, tau0, tau1, tau1 + 1.5, 8)
    assert len(none) == 0
