"""Serving on the device: every request's final particles (fixed32 positions,
velocities) are bit-identical whether its box was stepped alone or batched
with whatever else was resident at the time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix(gpu):
    from cosmictiger_amd import serve
    reqs = serve.poisson_requests(7, rate=400.0, seed=11, nmin=1500, nmax=40000, max_steps=3)
    ex = serve.GpuExecutor(gpu, theta=0.6, bucket=64, dt=1e-3)
    srv = serve.Server(ex, hbm_budget=8 * 2 ** 30, host_threads=4, batched=True)
    out = srv.run(reqs)
    srv.close()
    assert out["served"] == len(reqs)
    assert srv.max_batch > 1
    for r in reqs:
        alone = serve.Request(r.rid, r.n, r.steps, r.seed, r.clustered, 0.0)
        s1 = serve.Server(ex, hbm_budget=8 * 2 ** 30, host_threads=1, batched=False)
        s1.run([alone])
        s1.close()
        for f in ("x", "y", "z", "vel", "rung"):
            assert np.array_equal(alone.result[f], r.result[f]), (r.rid, f)
    ex.close()
