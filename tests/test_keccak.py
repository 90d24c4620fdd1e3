"""The Keccak MMCS variant (src/whir/mod.rs:236-383) without a device.

Pinned:
- Keccak-f[1600], oracle and product (two independent writings: the product's
  tabulated round constants and rotation offsets, the oracle's derived from
  FIPS 202's LFSR and (x, y) walk), against the published all-zero-state
  known answers and against each other on random states;
- the oracle's byte sponge against Keccak-256's published digests of "" and
  "abc", and (domain byte 0x06) against hashlib's SHA3-256 on every length
  from 0 to 420 bytes (one to four blocks of the 136-byte rate);
- the product's host Keccak challenger against the oracle's on random
  observe/sample scripts that cross the rate and re-flush the output.

Parity unpinned (no fixture or closed form fixes them; they follow the
Plonky3 sources the reference depends on, p3-challenger / p3-keccak /
p3-symmetric 0.4, as written up in the oracle's header): the byte order of a
field element and of a digest in the byte challenger, HashChallenger's
chaining and pop-from-the-end order, SerializingChallenger32's sampling and
the packing of a row's Montgomery words into u64 lanes.  Both sides follow
the same reading, and every proof below is accepted by the oracle's verifier
and the product's.
"""
import ctypes
import hashlib

import numpy as np
import pytest

from helpers import oracle, oracle_params, oracle_prove, oracle_sample, oracle_verify, product, u32p

ZERO_STATE_ONCE = (0xF1258F7940E1DDE7, 0x84D5CCF933C0478A)
ZERO_STATE_TWICE = (0x2D5C954DF96ECB3C, 0x6A332CD07057B56D)


def oracle_f(state):
    a = np.array(state, dtype=np.uint64)
    oracle().orc_keccak_f1600(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return a


def product_f(state):
    a = np.array(state, dtype=np.uint64)
    W = product()
    W.check(W.lib().whirgpu_keccak_f1600(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    return a


def oracle_sponge(msg, pad):
    out = (ctypes.c_uint8 * 32)()
    buf = (ctypes.c_uint8 * max(len(msg), 1)).from_buffer_copy(msg if msg else b"\0")
    oracle().orc_keccak256(buf, len(msg), pad, out)
    return bytes(out)


@pytest.mark.parametrize("perm", [oracle_f, product_f])
def test_keccak_f1600_zero_state_known_answers(perm):
    once = perm(np.zeros(25, np.uint64))
    assert (int(once[0]), int(once[1])) == ZERO_STATE_ONCE
    twice = perm(once)
    assert (int(twice[0]), int(twice[1])) == ZERO_STATE_TWICE


def test_keccak_f1600_product_equals_oracle():
    rng = np.random.default_rng(5)
    for _ in range(64):
        s = rng.integers(0, 2 ** 64, 25, dtype=np.uint64)
        assert np.array_equal(product_f(s), oracle_f(s))


def test_keccak256_known_answers():
    assert oracle_sponge(b"", 0x01).hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"
    assert oracle_sponge(b"abc", 0x01).hex() == "4e03657aea45a94fc7d47ba826c8d667c0d1e6e33a64a036ec44f58fa12d6c45"


def test_sponge_absorbs_like_sha3_on_every_length():
    rng = np.random.default_rng(11)
    for n in range(0, 421):
        msg = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        assert oracle_sponge(msg, 0x06) == hashlib.sha3_256(msg).digest(), n


def random_script(rng, steps):
    ops, args = [], []
    for _ in range(steps):
        r = rng.random()
        if r < 0.55:
            for _ in range(int(rng.integers(1, 60))):   # up to 240 bytes: crosses the rate
                ops.append(0)
                args.append(int(rng.integers(0, 0x7F000001)))
        else:
            for _ in range(int(rng.integers(1, 20))):   # up to 80 bytes: re-flushes the chained output
                ops.append(1)
                args.append(0)
    return ops, args


def test_keccak_challenger_product_equals_oracle():
    W = product()
    L = oracle()
    rng = np.random.default_rng(3)
    for trial in range(12):
        ops, args = random_script(rng, 30)
        want = np.zeros(len(ops), np.uint32)
        k = L.orc_keccak_challenger_script(u32p(np.array(ops, np.uint32)), u32p(np.array(args, np.uint32)),
                                           len(ops), u32p(want))
        assert k == ops.count(1)
        ch = W.KeccakChallengerState()
        W.check(W.lib().whirgpu_keccak_challenger_new(ctypes.byref(ch)))
        got = []
        i = 0
        while i < len(ops):
            j = i
            while j < len(ops) and ops[j] == ops[i]:
                j += 1
            if ops[i] == 0:
                W.check(W.lib().whirgpu_keccak_challenger_observe(None, ctypes.byref(ch),
                                                                  u32p(np.array(args[i:j], np.uint32)), j - i))
            else:
                out = np.zeros(j - i, np.uint32)
                W.check(W.lib().whirgpu_keccak_challenger_sample(None, ctypes.byref(ch), u32p(out), j - i))
                got.extend(int(x) for x in out)
            i = j
        assert got == [int(x) for x in want[:k]], trial


def test_keccak_challenger_bits_and_bytes_product_equals_oracle():
    # sample_bits (script op 2) and observing a digest's bytes (as four field
    # words' little-endian bytes) through the product's host entry points
    W = product()
    L = oracle()
    rng = np.random.default_rng(8)
    for trial in range(10):
        ops, args = [], []
        ch = W.KeccakChallengerState()
        W.check(W.lib().whirgpu_keccak_challenger_new(ctypes.byref(ch)))
        got = []
        for _ in range(25):
            r = rng.random()
            if r < 0.4:
                v = int(rng.integers(0, 0x7F000001))
                ops.append(0)
                args.append(v)
                # the canonical value's bytes, as observe of a field element writes them
                canon = v * pow(pow(2, 32, 0x7F000001), -1, 0x7F000001) % 0x7F000001
                b = (ctypes.c_uint8 * 4)(*canon.to_bytes(4, "little"))
                W.check(W.lib().whirgpu_keccak_challenger_observe_bytes(None, ctypes.byref(ch), b, 4))
            else:
                bits = int(rng.integers(1, 32))
                ops.append(2)
                args.append(bits)
                out = np.zeros(1, np.uint32)
                W.check(W.lib().whirgpu_keccak_challenger_sample_bits(None, ctypes.byref(ch), bits, u32p(out)))
                got.append(int(out[0]))
        want = np.zeros(len(ops), np.uint32)
        k = L.orc_keccak_challenger_script(u32p(np.array(ops, np.uint32)), u32p(np.array(args, np.uint32)),
                                           len(ops), u32p(want))
        assert got == [int(x) for x in want[:k]], trial


def test_keccak_sample_bits_reads_raw_bytes():
    # sample_bits (ops 2) and a field sample (ops 1) start from the same four
    # bytes; the field sample keeps 31 bits of them (rejecting >= p), the
    # 16-bit sample keeps 16 of them: the two agree in their low 16 bits
    # whenever the field sample was not rejected.
    L = oracle()
    agree = 0
    for seed in range(40):
        obs = oracle_sample(900 + seed, 3)
        res = []
        for op in (1, 2):
            ops = np.array([0, 0, 0, op], np.uint32)
            args = np.array(list(obs) + [16], np.uint32)
            out = np.zeros(1, np.uint32)
            assert L.orc_keccak_challenger_script(u32p(ops), u32p(args), 4, u32p(out)) == 1
            res.append(int(out[0]))
        mont_r = pow(2, 32, 0x7F000001)
        canon = res[0] * pow(mont_r, -1, 0x7F000001) % 0x7F000001
        agree += (canon & 0xFFFF) == res[1]
    assert agree >= 38


KECCAK_CASES = [
    # the reference's Keccak end-to-end case: n=10, Constant(4), 2 points,
    # CapacityBound, no PoW, rs 1, security 32, rate 1 (src/whir/mod.rs:373-383)
    (10, dict(security_level=32, pow_bits=0, ff=4, rs=1, soundness=2), 2),
    (11, dict(security_level=32, pow_bits=6, ff=3, rs=2, soundness=1), 1),
    (12, dict(security_level=40, pow_bits=8, ff=4, ff0=3, rs=3), 3),
    (9, dict(security_level=32, pow_bits=4, ff=2, rs=1, leaf_width=24), 0),
]


def keccak_proof(n, kw, npts, seed):
    poly = oracle_sample(seed, 1 << n)
    pts = [oracle_sample(seed + 1 + j, 4 * n).reshape(n, 4) for j in range(npts)]
    prm = oracle_params(n, mmcs=1, **kw)
    proof, evals, _ = oracle_prove(prm, poly, pts)
    return prm, proof, pts, evals


def host_keccak_config(n, kw):
    W = product()
    prm = W.params(n, security_level=kw.get("security_level", 100), pow_bits=kw.get("pow_bits", 16),
                   folding_factor=kw.get("ff", 4), folding_factor_first=kw.get("ff0"),
                   rs_domain_initial_reduction_factor=kw.get("rs", 3), soundness=kw.get("soundness", 2),
                   leaf_hash_width=kw.get("leaf_width", 16), mmcs="keccak")
    return W.WhirConfig(prm, device=-1)


@pytest.mark.parametrize("n,kw,npts", KECCAK_CASES)
def test_oracle_keccak_proofs_verify_on_both_sides(n, kw, npts):
    W = product()
    prm, proof, pts, evals = keccak_proof(n, kw, npts, seed=60 + n)
    rc, msg = oracle_verify(prm, proof, pts, evals)
    assert rc == 0, msg
    cfg = host_keccak_config(n, kw)
    assert W.WhirProof.from_bytes(cfg, proof).to_bytes() == proof
    W.Verifier(cfg).verify(proof, cfg.challenger(), pts, evals)
    if npts:
        bad = evals.copy()
        bad[0, 2] ^= 1
        assert not W.verify_bytes(cfg, proof, pts, bad)
        assert oracle_verify(prm, proof, pts, bad)[0] != 0
    cfg.close()


def test_keccak_and_poseidon2_proofs_differ_and_do_not_cross_verify():
    W = product()
    n, kw, npts = KECCAK_CASES[0]
    prm, proof, pts, evals = keccak_proof(n, kw, npts, seed=70)
    pproof, pevals, _ = oracle_prove(oracle_params(n, digest_elems=4, **kw), oracle_sample(70, 1 << n), pts)
    assert np.array_equal(pevals, evals)   # the statement does not depend on the MMCS
    assert pproof != proof
    cfg = host_keccak_config(n, kw)
    assert not W.verify_bytes(cfg, pproof, pts, evals)
    cfg.close()


def test_keccak_bit_flips_are_judged_like_the_oracle():
    W = product()
    n, kw, npts = KECCAK_CASES[2]
    prm, proof, pts, evals = keccak_proof(n, kw, npts, seed=80)
    cfg = host_keccak_config(n, kw)
    rng = np.random.default_rng(17)
    for bit in sorted(set(int(x) for x in rng.integers(0, 8 * len(proof), 120))):
        t = bytearray(proof)
        t[bit // 8] ^= 1 << (bit % 8)
        t = bytes(t)
        mine = W.verify_bytes(cfg, t, pts, evals)
        rc, _ = oracle_verify(prm, t, pts, evals)
        assert mine == (rc == 0), bit
        assert not mine, bit
    cfg.close()
