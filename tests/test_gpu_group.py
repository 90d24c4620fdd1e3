"""Groups: several proofs of one configuration committed and proved in
lockstep, one launch per device step for all of them (whirgpu_prove_group).
Every proof of a group must equal the oracle's proof of the same inputs (and
the same request proved alone), whatever else is in the group."""
import numpy as np
import pytest

from helpers import oracle_params, oracle_prove, oracle_sample, oracle_verify, product

pytestmark = pytest.mark.gpu

BENCH = dict(security_level=100, pow_bits=16, ff=4, rs=3)


def cfg_for(W, n, security_level=100, pow_bits=16, ff=4, rs=3, soundness=2):
    return W.WhirConfig(W.params(n, security_level=security_level, pow_bits=pow_bits, folding_factor=ff,
                                 rs_domain_initial_reduction_factor=rs, soundness=soundness), device=0)


def inputs(n, count, npts, seed):
    polys = [oracle_sample(seed + 100 * i, 1 << n) for i in range(count)]
    pts = [[oracle_sample(seed + 100 * i + 1 + j, 4 * n).reshape(n, 4) for j in range(npts)] for i in range(count)]
    return polys, pts


def check_against_oracle(n, kw, polys, pts, got):
    for (proof, evals), poly, p in zip(got, polys, pts):
        want, want_evals, _ = oracle_prove(oracle_params(n, **kw), poly, p)
        assert all(np.array_equal(a, b) for a, b in zip(evals, want_evals))
        assert proof == want
        rc, err = oracle_verify(oracle_params(n, **kw), proof, p, want_evals)
        assert rc == 0, err


@pytest.mark.parametrize("n,count,npts", [(16, 5, 1), (14, 3, 2), (18, 4, 1), (13, 2, 5)])
def test_group_matches_oracle_at_benchmark_parameters(n, count, npts):
    W = product()
    cfg = cfg_for(W, n)
    polys, pts = inputs(n, count, npts, seed=n * 7 + count)
    got = W.prove_group(cfg, polys, pts)
    check_against_oracle(n, BENCH, polys, pts, got)
    cfg.close()


def test_full_group_equals_single_proofs():
    # sixteen proofs (the group limit) against the same requests proved one at a time
    W = product()
    n = 12
    kw = dict(security_level=60, pow_bits=8, ff=3, rs=2)
    cfg = cfg_for(W, n, **kw)
    polys, pts = inputs(n, 16, 1, seed=500)
    got = W.prove_group(cfg, polys, pts)
    for (proof, _), poly, p in zip(got, polys, pts):
        alone, _ = W.prove_once(cfg, poly, p)
        assert proof == alone
    # and a second group on the same context reuses its slots
    again = W.prove_group(cfg, polys[:7], pts[:7])
    assert [a for a, _ in again] == [a for a, _ in got[:7]]
    cfg.close()


def test_group_with_host_driven_rounds():
    # 270 queries in round 0: the round is driven by the host for each proof of the group
    W = product()
    n = 12
    kw = dict(security_level=128, pow_bits=16, ff=4, rs=3, soundness=0)
    cfg = cfg_for(W, n, **kw)
    polys, pts = inputs(n, 3, 1, seed=900)
    got = W.prove_group(cfg, polys, pts)
    check_against_oracle(n, kw, polys, pts, got)
    cfg.close()


def test_group_of_small_instances():
    # n <= folding factor: classic first rounds, proof by proof inside the group
    W = product()
    n = 4
    kw = dict(security_level=32, pow_bits=5, ff=4, rs=1)
    cfg = cfg_for(W, n, **kw)
    polys, pts = inputs(n, 3, 2, seed=40)
    got = W.prove_group(cfg, polys, pts)
    check_against_oracle(n, kw, polys, pts, got)
    cfg.close()


def test_group_at_n24_device_inputs():
    W = product()
    n = 24
    cfg = cfg_for(W, n)
    polys = [W.sample_field(10 + i, 1 << n) for i in range(2)]
    pts = [[W.sample_field(20 + i, 4 * n).reshape(n, 4)] for i in range(2)]
    bufs = [W.DeviceBuffer.from_array(p) for p in polys]
    got = W.prove_group(cfg, None, pts, device_ptrs=[b.ptr for b in bufs])
    for b in bufs:
        b.free()
    check_against_oracle(n, BENCH, polys, pts, got)
    cfg.close()


def test_group_rejects_mixed_constraint_counts():
    W = product()
    n = 10
    cfg = cfg_for(W, n, security_level=32, pow_bits=5, ff=2, rs=1)
    polys, pts = inputs(n, 2, 1, seed=3)
    pts[1] = pts[1] + pts[1]
    sts = [W.InitialStatement(cfg, p) for p in polys]
    for st, ps in zip(sts, pts):
        for z in ps:
            st.evaluate(z)
    import ctypes
    handles = (ctypes.c_void_p * 2)(*[s.handle.value for s in sts])
    proofs = [W.WhirProof(cfg) for _ in range(2)]
    ph = (ctypes.c_void_p * 2)(*[p.handle.value for p in proofs])
    chs = (W.ChallengerState * 2)(cfg.challenger(), cfg.challenger())
    assert W.lib().whirgpu_prove_group(cfg.handle, 2, handles, ph, chs) == 1
    assert b"same number of constraints" in W.lib().whirgpu_last_error()
    cfg.close()
