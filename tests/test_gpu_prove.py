"""Whole commit + prove on the device against the oracle: byte-identical
proofs, and the oracle's verifier accepts the device's proofs."""
import numpy as np
import pytest

from helpers import oracle_params, oracle_prove, oracle_sample, oracle_verify, product

pytestmark = pytest.mark.gpu

CASES = [
    (10, dict(security_level=32, pow_bits=5, ff=2, rs=1)),
    (10, dict(security_level=32, pow_bits=0, ff=3, rs=2, soundness=0)),
    (9, dict(security_level=32, pow_bits=5, ff=2, rs=1, soundness=1)),
    (11, dict(security_level=32, pow_bits=10, ff=4, ff0=3, rs=3)),
    (8, dict(security_level=32, pow_bits=5, ff=1, rs=1)),
    (6, dict(security_level=32, pow_bits=5, ff=2, rs=2)),
    (4, dict(security_level=32, pow_bits=5, ff=4, rs=1)),
    (14, dict(security_level=100, pow_bits=16, ff=4, rs=3)),
    (16, dict(security_level=100, pow_bits=16, ff=4, rs=3)),
    (13, dict(security_level=32, pow_bits=5, ff=2, rs=1, leaf_width=24)),
    (12, dict(security_level=60, pow_bits=8, ff=3, rs=1, digest_elems=32)),
]


def product_cfg(n, security_level=100, pow_bits=16, ff=4, ff0=None, rate=1, rs=3, soundness=2, svo=1,
                leaf_width=16, digest_elems=8):
    W = product()
    prm = W.params(n, security_level=security_level, pow_bits=pow_bits, folding_factor=ff, folding_factor_first=ff0,
                   starting_log_inv_rate=rate, rs_domain_initial_reduction_factor=rs, soundness=soundness,
                   svo=bool(svo), leaf_hash_width=leaf_width, digest_elems=digest_elems)
    return W.WhirConfig(prm, device=0)


def inputs(n, npts, seed=0):
    poly = oracle_sample(seed, 1 << n)
    pts = [oracle_sample(seed + 1 + j, 4 * n).reshape(n, 4) for j in range(npts)]
    return poly, pts


@pytest.mark.parametrize("n,kw", CASES)
# five points: one full four-point weight launch, then one accumulating launch
@pytest.mark.parametrize("npts", [0, 1, 2, 5])
def test_proof_bytes_match_oracle(n, kw, npts):
    W = product()
    poly, pts = inputs(n, npts, seed=n)
    cfg = product_cfg(n, **kw)
    got, evals = W.prove_once(cfg, poly, pts)
    want, want_evals, _ = oracle_prove(oracle_params(n, **kw), poly, pts)
    for a, b in zip(evals, want_evals):
        assert np.array_equal(a, b)
    assert got == want
    rc, err = oracle_verify(oracle_params(n, **kw), got, pts, want_evals)
    assert rc == 0, err


# Schedules whose query counts exceed the device transcript's 256 (RT_MAX_QUERIES):
# UniqueDecoding at lambda 128 gives 270 queries in round 0 (host-driven WHIR
# round) or 270 final queries (host-driven final round).
FALLBACK_CASES = [
    ("round", 12, dict(security_level=128, pow_bits=16, ff=4, rs=3, soundness=0)),
    ("final", 10, dict(security_level=128, pow_bits=16, ff=4, rs=1, soundness=0)),
    ("final", 6, dict(security_level=128, pow_bits=16, ff=2, rs=1, soundness=0)),
]


@pytest.mark.parametrize("path,n,kw", FALLBACK_CASES)
@pytest.mark.parametrize("npts", [1, 3])
def test_host_driven_rounds_match_oracle(path, n, kw, npts):
    W = product()
    cfg = product_cfg(n, **kw)
    sch = cfg.schedule()
    n_rounds, final_queries = sch[2], sch[3]
    if path == "round":
        assert n_rounds >= 1 and sch[7 + 2] > 256, sch   # round 0's queries
    else:
        assert final_queries > 256, sch
    poly, pts = inputs(n, npts, seed=40 + n)
    got, evals = W.prove_once(cfg, poly, pts)
    want, want_evals, _ = oracle_prove(oracle_params(n, **kw), poly, pts)
    assert all(np.array_equal(a, b) for a, b in zip(evals, want_evals))
    assert got == want
    rc, err = oracle_verify(oracle_params(n, **kw), got, pts, want_evals)
    assert rc == 0, err
    # a second proof on the same context reuses its buffers and gives the same bytes
    again, _ = W.prove_once(cfg, poly, pts)
    assert again == want


# 2^25 coefficients: the command-line default of the reference (src/bin/main.rs:
# -d 25 -k 5 -l 90, rate 1, rs-reduction 3) and the benchmark's folding 4; both
# need pass-1 lines of 2^11 in the four-step encoding
@pytest.mark.parametrize("kw", [dict(security_level=90, pow_bits=16, ff=5, rs=3),
                                dict(security_level=100, pow_bits=16, ff=4, rs=3)])
def test_n25_matches_oracle(kw):
    W = product()
    n = 25
    poly = W.sample_field(0, 1 << n)
    pts = [W.sample_field(1, 4 * n).reshape(n, 4)]
    cfg = product_cfg(n, **kw)
    d = W.DeviceBuffer.from_array(poly)
    got, evals = W.prove_once(cfg, None, pts, device_ptr=d.ptr)
    d.free()
    want, want_evals, _ = oracle_prove(oracle_params(n, **kw), poly, pts)
    assert np.array_equal(evals[0], want_evals[0])
    assert got == want
    assert oracle_verify(oracle_params(n, **kw), got, pts, want_evals)[0] == 0


def test_config1_n18_matches_oracle():
    # n = 18, folding 4, rate 1/2, rs-reduction 3, lambda 100, one seeded constraint
    W = product()
    poly, pts = inputs(18, 1)
    cfg = product_cfg(18)
    got, _ = W.prove_once(cfg, poly, pts)
    want, evals, _ = oracle_prove(oracle_params(18), poly, pts)
    assert got == want
    assert oracle_verify(oracle_params(18), got, pts, evals)[0] == 0


def test_device_resident_input_and_repeat():
    W = product()
    n = 14
    poly, pts = inputs(n, 2, seed=3)
    cfg = product_cfg(n)
    d = W.DeviceBuffer.from_array(poly)
    a, _ = W.prove_once(cfg, poly, pts)
    b, _ = W.prove_once(cfg, None, pts, device_ptr=d.ptr)
    c, _ = W.prove_once(cfg, None, pts, device_ptr=d.ptr)
    assert a == b == c


def test_statement_clone_proves_like_the_original():
    # InitialStatement::clone: constraints evaluated once, then commit + prove on
    # clones (as the reference's benchmark does); every clone gives the oracle's
    # proof, and committing on a clone leaves the original unchanged
    W = product()
    n = 14
    poly, pts = inputs(n, 2, seed=5)
    cfg = product_cfg(n)
    want, _, _ = oracle_prove(oracle_params(n), poly, pts)
    d = W.DeviceBuffer.from_array(poly)
    for st0 in (W.InitialStatement(cfg, poly), W.InitialStatement(cfg, None, device_ptr=d.ptr)):
        for p in pts:
            st0.evaluate(p)
        for _ in range(2):
            st = st0.clone()
            ch = cfg.challenger()
            proof = W.WhirProof(cfg)
            pd = W.CommitmentWriter(cfg).commit(proof, ch, st)
            W.Prover(cfg).prove(proof, ch, st, pd)
            assert proof.to_bytes() == want
            proof.close()
            st.close()
        st0.close()
    d.free()


def test_config2_n24_matches_oracle():
    # the 2^24-variable benchmark proof: two seeded constraints, Poseidon2-16 Merkle
    W = product()
    n = 24
    poly = W.sample_field(0, 1 << n)
    pts = [W.sample_field(1 + j, 4 * n).reshape(n, 4) for j in range(2)]
    cfg = product_cfg(n)
    d = W.DeviceBuffer.from_array(poly)
    got, evals = W.prove_once(cfg, None, pts, device_ptr=d.ptr)
    want, want_evals, _ = oracle_prove(oracle_params(n), poly, pts)
    assert all(np.array_equal(a, b) for a, b in zip(evals, want_evals))
    assert got == want
    assert oracle_verify(oracle_params(n), got, pts, want_evals)[0] == 0
