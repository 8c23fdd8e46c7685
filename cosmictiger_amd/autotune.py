"""Autotuner of the kick's leaf bucket size for a set of resident boxes.

The reference picks the bucket by hand (``--bucket_size``, default 128,
src/options.cpp:107) and sweeps it in its ``--test=buckets`` mode
(src/test.cpp:402-476), timing a kick per size.  Here the sweep is automatic:
every candidate bucket is applied to all the boxes (tree rebuilt), the boxes
are kicked together (one batched kick to warm up, then timed kicks), and the
fastest bucket is kept.  The force error changes little with the bucket (the
opening angle sets it), so the choice is a pure throughput one.
"""
import time

import numpy as np

DEFAULT_BUCKETS = (32, 48, 64, 96, 128, 160)


def tune_bucket(ct, boxes, thetas, hsofts, params, candidates=DEFAULT_BUCKETS, steps=2, graph=True, log=None):
    """Returns (best bucket, {bucket: seconds per batched kick}).  The boxes
    are left with trees built at the best bucket."""
    times = {}
    batch = ct.Batch(boxes)
    try:
        for bucket in candidates:
            for b, th, h in zip(boxes, thetas, hsofts):
                b.tree_build(th, bucket, h)
            batch.kick(params, graph=graph)
            t0 = time.perf_counter()
            for _ in range(steps):
                batch.kick(params, graph=graph)
            times[bucket] = (time.perf_counter() - t0) / steps
            if log:
                log("autotune: bucket %3d  %.2f ms per batched kick" % (bucket, 1e3 * times[bucket]))
    finally:
        batch.close()
    best = min(times, key=times.get)
    for b, th, h in zip(boxes, thetas, hsofts):
        b.tree_build(th, best, h)
    return best, times


def ensemble_specs(nbox=64, nmin=32 ** 3, nmax=128 ** 3, seed=0):
    """The ensemble of independent boxes: N log-uniform in [nmin, nmax], box i
    seeded with 1000 + i, clustered fraction uniform in [0, 0.8]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nbox)).astype(np.int64)
    fracs = rng.uniform(0.0, 0.8, nbox)
    return [(int(n), float(f), 1000 + i) for i, (n, f) in enumerate(zip(ns, fracs))]
