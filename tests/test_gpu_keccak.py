"""The Keccak MMCS variant (src/whir/mod.rs:236-383) on the device: Keccak
leaves and tree kernels, the host-driven transcript and host grinding give
proofs byte-identical to the oracle's, which both verifiers accept; the
verifier's challenger ends where the prover's ended."""
import ctypes

import numpy as np
import pytest

from helpers import oracle, oracle_params, oracle_prove, oracle_sample, oracle_verify, product, u32p

pytestmark = pytest.mark.gpu

CASES = [
    # the reference's end-to-end Keccak case (src/whir/mod.rs:373-383)
    (10, dict(security_level=32, pow_bits=0, ff=4, rs=1, soundness=2), 2),
    (11, dict(security_level=32, pow_bits=6, ff=3, rs=2, soundness=1), 1),
    (12, dict(security_level=40, pow_bits=8, ff=4, ff0=3, rs=3), 3),
    (9, dict(security_level=32, pow_bits=4, ff=2, rs=1, leaf_width=24), 0),
    (8, dict(security_level=32, pow_bits=5, ff=1, rs=1, soundness=0), 5),
    (4, dict(security_level=32, pow_bits=5, ff=4, rs=1), 1),
    (16, dict(security_level=100, pow_bits=16, ff=4, rs=3), 2),
    (18, dict(security_level=60, pow_bits=10, ff=5, rs=3), 1),
]


def keccak_cfg(n, kw, device=0):
    W = product()
    prm = W.params(n, security_level=kw.get("security_level", 100), pow_bits=kw.get("pow_bits", 16),
                   folding_factor=kw.get("ff", 4), folding_factor_first=kw.get("ff0"),
                   rs_domain_initial_reduction_factor=kw.get("rs", 3), soundness=kw.get("soundness", 2),
                   leaf_hash_width=kw.get("leaf_width", 16), mmcs="keccak")
    return W.WhirConfig(prm, device=device)


def prove(W, cfg, poly, pts):
    st = W.InitialStatement(cfg, poly)
    evals = np.array([st.evaluate(p) for p in pts], dtype=np.uint32).reshape(len(pts), 4)
    ch = cfg.challenger()
    proof = W.WhirProof(cfg)
    pd = W.CommitmentWriter(cfg).commit(proof, ch, st)
    W.Prover(cfg).prove(proof, ch, st, pd)
    data = proof.to_bytes()
    proof.close()
    st.close()
    return data, evals, ch


def samples(W, ch, k=8):
    out = np.zeros(k, np.uint32)
    W.check(W.lib().whirgpu_keccak_challenger_sample(None, ctypes.byref(ch), u32p(out), k))
    return out


@pytest.mark.parametrize("n,kw,npts", CASES)
def test_keccak_proof_bytes_match_oracle(n, kw, npts):
    W = product()
    print(f"keccak n={n} {kw} npts={npts}", flush=True)
    poly = oracle_sample(n + 500, 1 << n)
    pts = [oracle_sample(n + 501 + j, 4 * n).reshape(n, 4) for j in range(npts)]
    cfg = keccak_cfg(n, kw)
    got, evals, pch = prove(W, cfg, poly, pts)
    prm = oracle_params(n, mmcs=1, **kw)
    want, want_evals, _ = oracle_prove(prm, poly, pts)
    assert np.array_equal(evals, want_evals.reshape(npts, 4))
    assert got == want
    rc, msg = oracle_verify(prm, got, pts, evals)
    assert rc == 0, msg
    vch = cfg.challenger()
    W.Verifier(cfg).verify(got, vch, pts, evals)
    assert np.array_equal(samples(W, vch), samples(W, pch))
    cfg.close()


def test_keccak_n22_proof_verifies():
    # beyond the oracle's quick sizes: the product's own verifier, and a tamper
    W = product()
    n = 22
    cfg = keccak_cfg(n, dict(security_level=100, pow_bits=16, ff=4, rs=3))
    poly = W.sample_field(3, 1 << n)
    pts = [W.sample_field(4 + j, 4 * n).reshape(n, 4) for j in range(2)]
    data, evals, pch = prove(W, cfg, poly, pts)
    vch = cfg.challenger()
    W.Verifier(cfg).verify(data, vch, pts, evals)
    assert np.array_equal(samples(W, vch), samples(W, pch))
    bad = evals.copy()
    bad[0, 0] ^= 1
    assert not W.verify_bytes(cfg, data, pts, bad)
    cfg.close()


def test_keccak_context_refuses_the_duplex_paths():
    W = product()
    cfg = keccak_cfg(10, CASES[0][1])
    ch = W.ChallengerState()
    W.check(W.lib().whirgpu_challenger_new(ctypes.byref(ch)))
    with pytest.raises(W.WhirError, match="Keccak"):
        W.check(W.lib().whirgpu_challenger_observe_domain_separator(cfg.handle, ctypes.byref(ch)))
    with pytest.raises(W.WhirError, match="Keccak"):
        W.prove_group(cfg, [oracle_sample(1, 1 << 10)], [[]])
    # and at the C ABI, with a duplex challenger for the one proof
    st = W.InitialStatement(cfg, oracle_sample(1, 1 << 10))
    proof = W.WhirProof(cfg)
    sts = (ctypes.c_void_p * 1)(st.handle.value)
    prs = (ctypes.c_void_p * 1)(proof.handle.value)
    with pytest.raises(W.WhirError, match="Keccak"):
        W.check(W.lib().whirgpu_prove_group(cfg.handle, 1, sts, prs, ctypes.byref(ch)))
    proof.close()
    st.close()
    cfg.close()


# input lengths around the rate: the witness's bytes in the last lane of a
# block with the padding in the next block (132 mod 136), byte offsets that run
# over a lane (5-7 mod 8) and over a block (133-135 mod 136), several prefix
# blocks
GRIND_LENGTHS = [0, 1, 4, 31, 32, 36, 100, 128, 131, 132, 133, 134, 135, 136, 137, 268, 272, 405, 600]


@pytest.mark.parametrize("bits", [1, 7, 12, 16])
def test_keccak_grind_smallest_witness_matches_oracle(bits):
    W = product()
    L = oracle()
    cfg = keccak_cfg(10, CASES[0][1])
    rng = np.random.default_rng(bits)
    for ln in GRIND_LENGTHS:
        data = bytes(rng.integers(0, 256, ln, dtype=np.uint8))
        nxt = np.zeros(4, np.uint32)
        buf = (ctypes.c_uint8 * max(ln, 1)).from_buffer_copy(data if ln else b"\0")
        want = L.orc_keccak_grind_bytes(buf, ln, bits, u32p(nxt))
        ch = W.KeccakChallengerState()
        W.check(W.lib().whirgpu_keccak_challenger_new(ctypes.byref(ch)))
        ctypes.memmove(ch.input, data, ln)
        ch.input_len = ln
        got = np.zeros(1, np.uint32)
        W.check(W.lib().whirgpu_keccak_grind(cfg.handle, ctypes.byref(ch), bits, u32p(got)))
        assert int(got[0]) == want, (bits, ln)
        assert np.array_equal(samples(W, ch, 4), nxt), (bits, ln)
    cfg.close()
