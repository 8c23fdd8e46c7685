"""The serving scheduler on the CPU, with a fake executor: admission by HBM
footprint, rejection of requests that can never fit, FIFO order, continuous
batching of the boxes that are ready, cancellation, and results that do not
depend on batching."""
from cosmictiger_amd import serve


def _reqs(count=12, rate=200.0, seed=3):
    return serve.poisson_requests(count, rate, seed=seed, nmin=1000, nmax=200000, max_steps=5)


def test_all_served_with_their_steps():
    ex = serve.FakeExecutor(ns_per_particle=20.0)
    srv = serve.Server(ex, hbm_budget=10 ** 12, host_threads=4)
    reqs = _reqs()
    out = srv.run(reqs)
    srv.close()
    assert out["served"] == len(reqs) and out["rejected"] == 0
    for r in reqs:
        assert len(r.step_latencies) == r.steps
        assert r.result["steps"] == r.steps
        assert all(x > 0 for x in r.step_latencies)
    assert out["steps"] == sum(r.steps for r in reqs)
    assert srv.max_batch > 1  # boxes that were ready together shared a step


def test_budget_admission_and_rejection():
    reqs = _reqs(10)
    big = max(reqs, key=lambda r: r.n)
    budget = serve.footprint(big.n) + serve.footprint(min(r.n for r in reqs))
    huge = serve.Request(99, 10 ** 9, 1, 1, 0.0, 0.0)
    ex = serve.FakeExecutor(ns_per_particle=50.0)
    srv = serve.Server(ex, hbm_budget=budget, host_threads=4)
    out = srv.run(reqs + [huge])
    srv.close()
    assert huge.rejected and huge.result is None
    assert out["rejected"] == 1 and out["served"] == len(reqs)
    assert srv.max_used <= budget
    # admission is in arrival order
    admitted = sorted((r for r in reqs), key=lambda r: r.admitted_at)
    assert [r.rid for r in admitted] == [r.rid for r in sorted(reqs, key=lambda r: r.arrival)]


def test_results_independent_of_batching():
    a = _reqs(8, seed=5)
    b = _reqs(8, seed=5)
    s1 = serve.Server(serve.FakeExecutor(), 10 ** 12, batched=True)
    s2 = serve.Server(serve.FakeExecutor(), 10 ** 12, batched=False)
    s1.run(a)
    s2.run(b)
    s1.close()
    s2.close()
    assert [r.result for r in a] == [r.result for r in b]


def test_cancel_frees_the_box():
    ex = serve.FakeExecutor(ns_per_particle=200.0)
    srv = serve.Server(ex, 10 ** 12, host_threads=2)
    reqs = _reqs(6, rate=1e6)
    victim = max(reqs, key=lambda r: r.steps)
    victim.steps = 50
    srv.cancel(victim.rid)
    out = srv.run(reqs)
    srv.close()
    assert victim.cancelled and victim.result is None
    assert ("release", victim.rid) in ex.log
    assert out["cancelled"] == 1 and out["served"] == len(reqs) - 1
