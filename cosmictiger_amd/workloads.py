"""Synthetic particle boxes of the kick benchmarks (seeded, deterministic).

The reference initialises particles with GSL taus streams per thread
(src/particles.cpp:881-912); those streams are not reproducible here, so the
boxes come from numpy's PCG64 with an explicit seed.  Positions are fixed32
(2^32 u truncated, cosmictiger/fixed.hpp:98-101).

* ``uniform_box``: uniform random positions in the periodic unit box.
* ``nfw_box``: a fraction of the particles in NFW halos (masses log-uniform
  over 3 decades, particle counts proportional to mass, r200 from an
  overdensity of 200 times the mean, concentration uniform in [5, 15], radii by
  inverse CDF of mu(x) = ln(1+x) - x/(1+x) truncated at r200, isotropic
  directions, uniform centres, periodic wrap) and the rest uniform.
"""
import numpy as np


def to_fixed32(u):
    u = np.mod(np.asarray(u, dtype=np.float64), 1.0)
    return np.minimum(np.floor(u * 4294967296.0), 4294967295.0).astype(np.uint64).astype(np.uint32)


def uniform_box(n, seed):
    rng = np.random.default_rng(seed)
    return tuple(rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))


def _nfw_mu(x):
    return np.log1p(x) - x / (1.0 + x)


def _nfw_radii(rng, c, count):
    """x in [0, c] with density proportional to the NFW mass profile."""
    u = rng.uniform(0.0, 1.0, count) * _nfw_mu(c)
    lo = np.zeros(count)
    hi = np.full(count, float(c))
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        m = _nfw_mu(mid)
        lo = np.where(m < u, mid, lo)
        hi = np.where(m < u, hi, mid)
    return 0.5 * (lo + hi)


def nfw_box(n, seed, clustered_fraction=0.7, nhalos=2000, mass_decades=3.0, cmin=5.0, cmax=15.0):
    rng = np.random.default_rng(seed)
    nc = int(round(clustered_fraction * n))
    nu = n - nc
    pos = np.empty((n, 3))
    if nc > 0 and nhalos > 0:
        masses = 10.0 ** rng.uniform(0.0, mass_decades, nhalos)
        share = masses / masses.sum() * nc
        counts = np.floor(share).astype(np.int64)
        rem = nc - counts.sum()
        if rem > 0:
            order = np.argsort(-(share - counts))
            counts[order[:rem]] += 1
        centres = rng.uniform(0.0, 1.0, (nhalos, 3))
        conc = rng.uniform(cmin, cmax, nhalos)
        off = 0
        for h in range(nhalos):
            k = int(counts[h])
            if k == 0:
                continue
            r200 = (3.0 * k / (4.0 * np.pi * 200.0 * n)) ** (1.0 / 3.0)
            rs = r200 / conc[h]
            r = _nfw_radii(rng, conc[h], k) * rs
            v = rng.normal(size=(k, 3))
            v /= np.linalg.norm(v, axis=1)[:, None]
            pos[off:off + k] = centres[h] + v * r[:, None]
            off += k
        assert off == nc
    pos[nc:] = rng.uniform(0.0, 1.0, (nu, 3))
    perm = rng.permutation(n)
    pos = pos[perm]
    return to_fixed32(pos[:, 0]), to_fixed32(pos[:, 1]), to_fixed32(pos[:, 2])


def gaussian_velocities(n, seed, sigma=1e-3):
    rng = np.random.default_rng(seed + 7919)
    return (rng.normal(0.0, sigma, (n, 3))).astype(np.float32)
