"""Protocol-level checks of the oracle: schedule, transcript, prove/verify."""
import numpy as np
import pytest

from helpers import from_monty, oracle, oracle_params, oracle_prove, oracle_sample, oracle_verify, to_monty, u32p

# Appendix-A schedule of the survey, derived from src/whir/parameters.rs:81-234
# by hand: (n, rs-red) -> (commit OOD, start folding pow, final (queries, pow,
# final sumcheck rounds), rounds (vars, log domain, queries, ood, pow, fold pow)).
SCHEDULES = {
    (24, 3): (2, 7, (8, 12, 4), [(20, 25, 91, 2, 15, 6), (16, 22, 44, 2, 15, 11), (12, 21, 18, 2, 13, 16),
                                 (8, 20, 11, 2, 15, 21)]),
    (24, 1): (2, 7, (7, 9, 4), [(20, 25, 91, 2, 15, 12), (16, 24, 22, 2, 15, 17), (12, 23, 13, 2, 17, 22),
                                (8, 22, 9, 2, 18, 27)]),
    (18, 3): (2, 1, (18, 11, 6), [(14, 19, 91, 2, 15, 0), (10, 16, 44, 2, 15, 5)]),
    (18, 1): (2, 1, (13, 9, 6), [(14, 19, 91, 2, 15, 6), (10, 18, 22, 2, 13, 11)]),
    (16, 3): (2, 0, (18, 11, 4), [(12, 17, 91, 2, 15, 0), (8, 14, 44, 2, 15, 3)]),
    (20, 3): (2, 3, (11, 12, 4), [(16, 21, 91, 2, 15, 2), (12, 18, 44, 2, 15, 7), (8, 17, 18, 2, 11, 12)]),
    (22, 3): (2, 5, (11, 12, 6), [(18, 23, 91, 2, 15, 4), (14, 20, 44, 2, 15, 9), (10, 19, 18, 2, 11, 14)]),
}


def oracle_schedule(n, rs, **kw):
    import ctypes
    out = (ctypes.c_int * 256)()
    k = oracle().orc_config(oracle_params(n, rs=rs, **kw), out, 256)
    assert k > 0
    return list(out[:k])


@pytest.mark.parametrize("key", sorted(SCHEDULES))
def test_schedule_matches_appendix(key):
    n, rs = key
    ood, spow, (fq, fpow, fsr), rounds = SCHEDULES[key]
    s = oracle_schedule(n, rs)
    assert s[0] == ood and s[1] == spow
    assert (s[3], s[4], s[5]) == (fq, fpow, fsr)
    assert s[2] == len(rounds)
    for i, (nv, ld, q, o, pw, fp) in enumerate(rounds):
        r = s[7 + 8 * i: 15 + 8 * i]
        assert (r[4], r[6], r[2], r[3], r[0], r[1]) == (nv, ld, q, o, pw, fp)


def test_round_generators_have_the_folded_domain_order():
    s = oracle_schedule(24, 3)
    P = 0x7F000001
    for i in range(s[2]):
        r = s[7 + 8 * i: 15 + 8 * i]
        order_log = r[6] - r[5]
        g = int(r[7])
        assert pow(g, 1 << order_log, P) == 1 and pow(g, 1 << (order_log - 1), P) != 1


def test_domain_separator_shape():
    import ctypes
    out = np.zeros(4096, np.uint32)
    k = oracle().orc_domain_separator(oracle_params(18, rs=3), u32p(out), 4096)
    pat = [int(v) for v in from_monty(out[:k])]
    # observe(8, MerkleDigest) = 0 + 8 + 1; OOD: sample(2, OodQuery=7) = 9, observe(2, OodAnswers=1) = 4;
    # sample(1, InitialCombinationRandomness) = 1; then the first sumcheck round observe(3, SumcheckPoly=2) = 6
    assert pat[:5] == [9, 9, 4, 1, 6]
    assert pat[-1] == 5   # hint(DeferredWeightEvaluations = 3) + 2


CASES = [
    # (n, params) covering folding schedules, soundness types, PoW and rs-reduction
    (10, dict(security_level=32, pow_bits=5, ff=2, rs=1)),
    (10, dict(security_level=32, pow_bits=0, ff=3, rs=2, soundness=0)),
    (9, dict(security_level=32, pow_bits=5, ff=2, rs=1, soundness=1)),
    (11, dict(security_level=32, pow_bits=10, ff=4, ff0=3, rs=3)),
    (8, dict(security_level=32, pow_bits=5, ff=1, rs=1)),
    (6, dict(security_level=32, pow_bits=5, ff=2, rs=2)),
    (14, dict(security_level=100, pow_bits=16, ff=4, rs=3)),
]


def _inputs(n, npts, seed=0):
    poly = oracle_sample(seed, 1 << n)
    pts = [oracle_sample(seed + 1 + j, 4 * n).reshape(n, 4) for j in range(npts)]
    return poly, pts


@pytest.mark.parametrize("n,kw", CASES)
@pytest.mark.parametrize("npts", [0, 1, 2])
def test_prove_verify(n, kw, npts):
    prm = oracle_params(n, **kw)
    poly, pts = _inputs(n, npts)
    proof, evals, _ = oracle_prove(prm, poly, pts)
    rc, err = oracle_verify(prm, proof, pts, evals)
    assert rc == 0, err


@pytest.mark.parametrize("n,kw", CASES[:4] + CASES[6:])
def test_classic_and_svo_transcripts_agree(n, kw):
    # src/sumcheck/tests.rs:473-487 asserts equal verifier randomness for both strategies
    poly, pts = _inputs(n, 2)
    a = oracle_prove(oracle_params(n, svo=1, **kw), poly, pts)[0]
    b = oracle_prove(oracle_params(n, svo=0, **kw), poly, pts)[0]
    assert a == b


def test_tampered_proofs_are_rejected():
    n, kw = 12, dict(security_level=32, pow_bits=5, ff=2, rs=1)
    prm = oracle_params(n, **kw)
    poly, pts = _inputs(n, 1)
    proof, evals, _ = oracle_prove(prm, poly, pts)
    assert oracle_verify(prm, proof, pts, evals)[0] == 0
    rng = np.random.default_rng(7)
    rejected = 0
    for pos in rng.choice(len(proof), 40, replace=False):
        b = bytearray(proof)
        b[pos] ^= 1 << int(rng.integers(0, 8))
        rejected += oracle_verify(prm, bytes(b), pts, evals)[0] != 0
    assert rejected == 40
    bad_eval = evals.copy()
    bad_eval[0, 0] = (int(bad_eval[0, 0]) + 1) % 0x7F000001
    assert oracle_verify(prm, proof, pts, bad_eval)[0] != 0


def test_wrong_witness_rejected_by_challenger():
    import ctypes
    L = oracle()
    # grind 8 bits, then check that the witness passes and witness+1 (almost surely) does not
    ops = np.array([0] * 5 + [3], np.uint32)
    args = np.array(list(to_monty([1, 2, 3, 4, 5])) + [8], np.uint32)
    out = np.zeros(4, np.uint32)
    assert L.orc_challenger_script(1, u32p(ops), u32p(args), 6, u32p(out)) == 1
    w = int(out[0])
    ops2 = np.array([0] * 6 + [2], np.uint32)
    for cand, ok in ((w, True), (w + 1, False)):
        args2 = np.array(list(to_monty([1, 2, 3, 4, 5, cand])) + [8], np.uint32)
        o = np.zeros(2, np.uint32)
        L.orc_challenger_script(1, u32p(ops2), u32p(args2), 7, u32p(o))
        assert (int(o[0]) == 0) == ok or not ok
    # smallest: every candidate below w fails
    for cand in range(w):
        args2 = np.array(list(to_monty([1, 2, 3, 4, 5, cand])) + [8], np.uint32)
        o = np.zeros(2, np.uint32)
        L.orc_challenger_script(1, u32p(ops2), u32p(args2), 7, u32p(o))
        assert int(o[0]) != 0
