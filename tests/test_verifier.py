"""The product's proof decoder and verifier (host code of libwhirgpu.so) on
proofs the oracle writes, without a device.

- decoding then encoding gives back the same bytes;
- every honest proof is accepted, over schedules with 0-5 rounds, both leaf
  widths, all three soundness assumptions, ConstantFromSecondRound, 0-3
  constraints and proof of work from 0 to 16 bits;
- a wrong claimed evaluation, a wrong point or a different statement count is
  rejected;
- on single-bit changes of a proof the product's verifier (decoder
  included) and the oracle's reach the same decision, and every change is
  rejected except where tiny parameters let a changed PoW witness pass;
- truncated, extended and non-canonical inputs are refused by the decoder.
"""
import ctypes

import numpy as np
import pytest

from helpers import oracle_params, oracle_prove, oracle_sample, oracle_verify, product, u32p

CASES = [
    (10, dict(security_level=32, pow_bits=5, ff=2, rs=1), 1),
    (11, dict(security_level=32, pow_bits=10, ff=4, ff0=3, rs=3), 2),
    (9, dict(security_level=32, pow_bits=5, ff=2, rs=1, soundness=1), 1),
    (10, dict(security_level=32, pow_bits=0, ff=3, rs=2, soundness=0), 3),
    (12, dict(security_level=60, pow_bits=8, ff=3, rs=2, leaf_width=24), 1),
    (14, dict(security_level=80, pow_bits=12, ff=4, rs=3), 2),
    (16, dict(rs=3), 1),
    (4, dict(security_level=32, pow_bits=5, ff=4, rs=1), 2),
    (13, dict(security_level=40, pow_bits=6, ff=2, rs=1), 0),
]


def product_params(n, security_level=100, pow_bits=16, ff=4, ff0=None, rate=1, rs=3, soundness=2, svo=1,
                   leaf_width=16, digest_elems=8):
    W = product()
    return W.params(n, security_level=security_level, pow_bits=pow_bits, folding_factor=ff,
                    folding_factor_first=ff0, starting_log_inv_rate=rate, rs_domain_initial_reduction_factor=rs,
                    soundness=soundness, svo=bool(svo), leaf_hash_width=leaf_width, digest_elems=digest_elems)


def honest_proof(n, kw, npts, seed):
    poly = oracle_sample(seed, 1 << n)
    pts = [oracle_sample(seed + 1 + j, 4 * n).reshape(n, 4) for j in range(npts)]
    proof, evals, _ = oracle_prove(oracle_params(n, **kw), poly, pts)
    return proof, pts, evals


def host_config(n, kw):
    W = product()
    return W.WhirConfig(product_params(n, **kw), device=-1)


def verifies(cfg, proof, pts, evals):
    return product().verify_bytes(cfg, proof, pts, evals)


@pytest.mark.parametrize("n,kw,npts", CASES)
def test_decode_encode_round_trip_and_accept(n, kw, npts):
    W = product()
    proof, pts, evals = honest_proof(n, kw, npts, seed=7 * n + npts)
    cfg = host_config(n, kw)
    assert W.WhirProof.from_bytes(cfg, proof).to_bytes() == proof
    ch = cfg.challenger()
    W.Verifier(cfg).verify(proof, ch, pts, evals)
    cfg.close()


@pytest.mark.parametrize("n,kw,npts", CASES[:4])
def test_wrong_statement_is_rejected(n, kw, npts):
    W = product()
    proof, pts, evals = honest_proof(n, kw, max(npts, 1), seed=100 + n)
    cfg = host_config(n, kw)
    bad = evals.copy()
    bad[0, 0] = (int(bad[0, 0]) + 1) % 0x7F000001
    assert not verifies(cfg, proof, pts, bad)
    other = [p.copy() for p in pts]
    other[0][0, 1] = (int(other[0][0, 1]) + 12345) % 0x7F000001
    assert not verifies(cfg, proof, other, evals)
    assert not verifies(cfg, proof, pts[1:], evals[1:])   # a constraint left out
    with pytest.raises(W.WhirError, match="error 7"):
        W.Verifier(cfg).verify(proof, cfg.challenger(), pts, bad)
    cfg.close()


@pytest.mark.parametrize("n,kw,npts,all_rejected", [CASES[0] + (True,), CASES[4] + (True,), CASES[7] + (False,)])
def test_bit_flips_are_judged_like_the_oracle(n, kw, npts, all_rejected):
    # the product's verifier and the oracle's reach the same decision on every
    # changed proof.  With 4 PoW bits and a final domain of two positions
    # (n=4, the last case) a flipped final witness passes its check with
    # probability 1/16 and the queries land on the same rows, so a few flips
    # there are valid proofs, for both verifiers; elsewhere every flip fails.
    proof, pts, evals = honest_proof(n, kw, npts, seed=300 + n)
    cfg = host_config(n, kw)
    rng = np.random.default_rng(n)
    positions = sorted(set(int(x) for x in rng.integers(0, 8 * len(proof), 150)))
    accepted = []
    for bit in positions:
        t = bytearray(proof)
        t[bit // 8] ^= 1 << (bit % 8)
        t = bytes(t)
        mine = verifies(cfg, t, pts, evals)
        rc, _ = oracle_verify(oracle_params(n, **kw), t, pts, evals)
        assert mine == (rc == 0), f"bit {bit}: product {'accepts' if mine else 'rejects'}, oracle rc {rc}"
        if mine:
            accepted.append(bit)
    if all_rejected:
        assert not accepted
    else:
        assert len(accepted) <= len(positions) // 10
    cfg.close()


def test_decoder_refuses_malformed_bytes():
    W = product()
    n, kw, npts = CASES[0]
    proof, pts, evals = honest_proof(n, kw, npts, seed=5)
    cfg = host_config(n, kw)
    for bad in (proof[:-1], proof[: len(proof) // 2], proof + b"\0", b"", b"\x02" + proof[1:]):
        with pytest.raises(W.WhirError, match="error 8"):
            W.WhirProof.from_bytes(cfg, bad)
    # a field word not below p: the first OOD answer after the initial digest
    off = 1 + 32 + 8
    t = bytearray(proof)
    t[off:off + 4] = (0x7F000001).to_bytes(4, "little")
    with pytest.raises(W.WhirError, match="not below p"):
        W.WhirProof.from_bytes(cfg, bytes(t))
    # a length far beyond the input is refused before anything is allocated
    t = bytearray(proof)
    t[1 + 32:1 + 32 + 8] = (1 << 60).to_bytes(8, "little")
    with pytest.raises(W.WhirError, match="length beyond the input"):
        W.WhirProof.from_bytes(cfg, bytes(t))
    cfg.close()


def test_verifier_challenger_state():
    # two verifications of one proof leave the same challenger (the GPU tests
    # compare it with the prover's), and a rejected proof leaves the caller's
    # challenger as it was
    W = product()
    n, kw, npts = CASES[1]
    proof, pts, evals = honest_proof(n, kw, npts, seed=9)
    cfg = host_config(n, kw)
    a = cfg.challenger()
    W.Verifier(cfg).verify(proof, a, pts, evals)
    b = cfg.challenger()
    W.Verifier(cfg).verify(proof, b, pts, evals)
    va, vb = np.zeros(8, np.uint32), np.zeros(8, np.uint32)
    W.check(W.lib().whirgpu_challenger_sample(cfg.handle, ctypes.byref(a), u32p(va), 8))
    W.check(W.lib().whirgpu_challenger_sample(cfg.handle, ctypes.byref(b), u32p(vb), 8))
    assert np.array_equal(va, vb)
    # a rejected proof leaves the caller's challenger unchanged
    c = cfg.challenger()
    before = bytes(c)
    bad = evals.copy()
    bad[0, 0] ^= 1
    assert not W.verify_bytes(cfg, proof, pts, bad)
    with pytest.raises(W.WhirError):
        W.Verifier(cfg).verify(proof, c, pts, bad)
    assert bytes(c) == before
    cfg.close()
