"""Batched and served proving: every proof identical to the oracle's and to
the same request proved alone, whatever else runs beside it."""
import time

import numpy as np
import pytest

from helpers import oracle_params, oracle_prove, oracle_sample, oracle_verify, product

pytestmark = pytest.mark.gpu

KW = dict(security_level=60, pow_bits=8, ff=3, rs=2)


def prm(W):
    return W.params(0, security_level=60, pow_bits=8, folding_factor=3, rs_domain_initial_reduction_factor=2)


def make_requests(W, specs):
    reqs = []
    for i, (n, npts) in enumerate(specs):
        poly = oracle_sample(100 + i, 1 << n)
        pts = [oracle_sample(1000 + 10 * i + j, 4 * n).reshape(n, 4) for j in range(npts)]
        reqs.append((W.ProofRequest(poly, pts), poly, pts))
    return reqs


SPECS = [(10, 1), (13, 2), (11, 0), (15, 3), (12, 1), (14, 4), (10, 2), (16, 1), (13, 1), (11, 2), (15, 1), (12, 2)]


def test_prove_batch_matches_oracle():
    W = product()
    reqs = make_requests(W, SPECS)
    assert W.prove_batch(prm(W), [r for r, _, _ in reqs], workers=5) == 0
    for r, poly, pts in reqs:
        assert r.job.status == 0
        want, evals, _ = oracle_prove(oracle_params(r.n, **KW), poly, pts)
        assert r.proof() == want
        assert np.array_equal(r.evals[: 4 * len(pts)].reshape(len(pts), 4) if pts else evals, evals)
        assert oracle_verify(oracle_params(r.n, **KW), r.proof(), pts, evals)[0] == 0


def test_served_results_equal_alone_results():
    W = product()
    reqs = make_requests(W, SPECS)
    alone = []
    for r, poly, pts in reqs:
        cfg = W.WhirConfig(W.params(r.n, security_level=60, pow_bits=8, folding_factor=3,
                                    rs_domain_initial_reduction_factor=2), device=0)
        alone.append(W.prove_once(cfg, poly, pts)[0])
        cfg.close()
    srv = W.Server(prm(W), workers=4)
    rng = np.random.default_rng(5)
    for r, _, _ in reqs:
        srv.submit(r)
        time.sleep(float(rng.exponential(0.002)))
    for (r, _, _), a in zip(reqs, alone):
        assert srv.wait(r) == 0
        assert r.proof() == a
    srv.close()


def test_admission_with_a_tiny_memory_budget_runs_one_at_a_time():
    W = product()
    reqs = make_requests(W, SPECS[:6])
    srv = W.Server(prm(W), workers=4, hbm_budget=1)
    for r, _, _ in reqs:
        srv.submit(r)
    seen_running = 0
    while any(r.job.status == W.PENDING for r, _, _ in reqs):
        q, running, _ = srv.stats()
        seen_running = max(seen_running, running)
        time.sleep(0.0005)
    for r, poly, pts in reqs:
        assert srv.wait(r) == 0
        want, _, _ = oracle_prove(oracle_params(r.n, **KW), poly, pts)
        assert r.proof() == want
    assert seen_running <= 1
    srv.close()


def test_bad_job_reports_an_error_and_others_finish():
    W = product()
    reqs = make_requests(W, [(10, 1), (12, 1)])
    bad = W.ProofRequest(oracle_sample(7, 1 << 3), [])   # n = 3 < folding factor 3 + schedule -> rejected
    bad.job.num_variables = 2
    srv = W.Server(prm(W), workers=2)
    srv.submit(reqs[0][0])
    srv.submit(bad)
    srv.submit(reqs[1][0])
    assert srv.wait(bad) != 0
    assert srv.wait(reqs[0][0]) == 0 and srv.wait(reqs[1][0]) == 0
    srv.close()
