"""Plonky3 conventions (CONVENTIONS.md C1-C18) against fixtures emitted by the
real Plonky3 @ c38eb05 on a machine with a Rust toolchain.

The fixture is tests/golden/plonky3/conventions.json.  This image has no Rust
toolchain, so the file does not exist here and the module is skipped; once it
is committed, every key present is checked against the product's host build
and against the oracle, and a mismatch names the convention to correct.
Values are raw Montgomery words (u32) unless stated.  Inputs of the keys that
need them:
- iota: field elements 0, 1, 2, ... (canonical values, stored as Montgomery words);
- challenger_script: a fresh DuplexChallenger over Poseidon2-16 (constants of
  SmallRng seed 1) running "ops": [[0, v]] observe v, [[1]] sample, [[2, b]]
  sample_bits(b); "outputs" holds one word per sample (raw) or sample_bits
  (plain integer);
- merkle_8x16_iota: the 8 x 16 base matrix of iota (row-major), leaves hashed
  by PaddingFreeSponge<P16, 16, 8, 8>, nodes TruncatedPermutation<P16, 2, 8, 16>:
  {"root", "open3_row", "open3_path"};
- dft_batch_8x2: {"input": 16 words row-major, "output": dft_batch of it};
- proof_n10_hex: bincode of the WhirProof for n = 10, folding 2, rate 1/2,
  rs-reduction 1, lambda 32, 5 PoW bits, CapacityBound, SVO, the polynomial of
  SmallRng seed 0 and one constraint at the point of SmallRng seed 1
  (40 draws), Poseidon2 constants of SmallRng seed 1 (grinding witnesses are
  the smallest ones, CONVENTIONS.md C12);
- keccak_merkle_8x16_iota: the same matrix under the Keccak MMCS, digests as
  lists of 4 u64: {"root", "open3_row", "open3_path"};
- keccak_challenger_script: as challenger_script, for a fresh
  SerializingChallenger32<F, HashChallenger<u8, Keccak256Hash, 32>>;
- proof_n10_keccak_hex: the proof_n10_hex statement and parameters proved with
  the Keccak MMCS and challenger (domain separator with 4 digest elements).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from helpers import oracle, oracle_params, oracle_prove, oracle_sample, product, u32p

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plonky3", "conventions.json")

pytestmark = pytest.mark.skipif(not os.path.exists(FIXTURE),
                                reason="no Plonky3 fixture (needs a Rust machine; see CONVENTIONS.md)")


@pytest.fixture(scope="module")
def fx():
    with open(FIXTURE) as f:
        return json.load(f)


def words(v):
    return np.asarray(v, dtype=np.uint32)


def need(fx, key):
    if key not in fx:
        pytest.skip(f"fixture has no {key}")
    return fx[key]


def iota(n):
    W = product()
    out = np.arange(n, dtype=np.uint64)
    return ((out << np.uint64(32)) % np.uint64(W.P)).astype(np.uint32)   # canonical i as a Montgomery word


def test_c1_one_is_montgomery(fx):
    assert need(fx, "koalabear_one_raw") == 0x01FFFFFE


def test_c3_two_adic_generator(fx):
    assert need(fx, "two_adic_generator_24") == int(oracle().orc_two_adic_generator(24))


def test_c2_ef_mul(fx):
    e = need(fx, "ef_mul")
    out = np.zeros(4, np.uint32)
    oracle().orc_ef_mul(u32p(words(e["a"])), u32p(words(e["b"])), u32p(out))
    assert np.array_equal(out, words(e["ab"]))


def test_c4_smallrng_stream(fx):
    want = need(fx, "smallrng_seed0_u64")
    u64p = ctypes.POINTER(ctypes.c_uint64)
    st = np.zeros(4, np.uint64)
    out = np.zeros(8, np.uint64)
    oracle().orc_rng_state(0, st.ctypes.data_as(u64p))
    oracle().orc_xoshiro_outputs(st.ctypes.data_as(u64p), out.ctypes.data_as(u64p), 8)
    assert [int(x) for x in out] == [int(x) for x in want]


def test_c5_field_sampling(fx):
    want = words(need(fx, "koalabear_random_seed0"))
    assert np.array_equal(product().sample_field(0, len(want)), want)
    assert np.array_equal(oracle_sample(0, len(want)), want)


@pytest.mark.parametrize("width", [16, 24])
def test_c6_poseidon2_constants(fx, width):
    c = need(fx, f"poseidon2_{width}_seed1")
    got = product().poseidon2_from_seed(1)
    assert list(getattr(got, f"w{width}_external_initial")) == [w for r in c["external_initial"] for w in r]
    assert list(getattr(got, f"w{width}_external_terminal")) == [w for r in c["external_terminal"] for w in r]
    assert list(getattr(got, f"w{width}_internal")) == list(c["internal"])


@pytest.mark.parametrize("width", [16, 24])
def test_c7_permutation(fx, width):
    want = words(need(fx, f"perm{width}_iota"))
    W = product()
    s = iota(width)
    W.check(W.lib().whirgpu_poseidon2_permute(ctypes.byref(W.poseidon2_from_seed(1)), width, u32p(s)))
    assert np.array_equal(s, want)


@pytest.mark.parametrize("key,width,n", [("sponge16_iota16", 16, 16), ("sponge16_iota64", 16, 64),
                                         ("sponge24_iota16", 24, 16)])
def test_c8_sponge(fx, key, width, n):
    want = words(need(fx, key))
    out = np.zeros(8, np.uint32)
    oracle().orc_sponge(1, width, u32p(iota(n)), n, u32p(out))
    assert np.array_equal(out, want)


def test_c9_c10_merkle(fx):
    m = need(fx, "merkle_8x16_iota")
    root = np.zeros(8, np.uint32)
    row = np.zeros(16, np.uint32)
    path = np.zeros(24, np.uint32)
    oracle().orc_merkle(1, 16, u32p(iota(128)), 8, 16, u32p(root), 3, u32p(row), u32p(path))
    assert np.array_equal(root, words(m["root"]))
    assert np.array_equal(row, words(m["open3_row"]))
    assert np.array_equal(path, words(np.asarray(m["open3_path"]).reshape(-1)))


def test_c11_challenger(fx):
    sc = need(fx, "challenger_script")
    ops = words([o[0] for o in sc["ops"]])
    args = words([o[1] if len(o) > 1 else 0 for o in sc["ops"]])
    out = np.zeros(len(sc["outputs"]), np.uint32)
    k = oracle().orc_challenger_script(1, u32p(ops), u32p(args), len(ops), u32p(out))
    assert k == len(sc["outputs"])
    assert [int(x) for x in out] == [int(x) for x in sc["outputs"]]


def test_c13_dft_order(fx):
    d = need(fx, "dft_batch_8x2")
    mat = words(d["input"]).copy()
    oracle().orc_dft_batch(u32p(mat), 8, 2)
    assert np.array_equal(mat, words(d["output"]))


def test_c14_proof_bytes(fx):
    want = bytes.fromhex(need(fx, "proof_n10_hex"))
    n = 10
    poly = oracle_sample(0, 1 << n)
    pts = [oracle_sample(1, 4 * n).reshape(n, 4)]
    got, _, _ = oracle_prove(oracle_params(n, security_level=32, pow_bits=5, ff=2, rs=1), poly, pts)
    assert got == want


def lanes_as_words(lanes):
    out = []
    for v in lanes:
        out += [int(v) & 0xFFFFFFFF, int(v) >> 32]
    return words(out)


def test_c16_keccak_merkle(fx):
    m = need(fx, "keccak_merkle_8x16_iota")
    root = np.zeros(8, np.uint32)
    row = np.zeros(16, np.uint32)
    path = np.zeros(24, np.uint32)
    oracle().orc_merkle_keccak(u32p(iota(128)), 8, 16, u32p(root), 3, u32p(row), u32p(path))
    assert np.array_equal(root, lanes_as_words(m["root"]))
    assert np.array_equal(row, words(m["open3_row"]))
    assert np.array_equal(path, np.concatenate([lanes_as_words(d) for d in m["open3_path"]]))


def test_c17_keccak_challenger(fx):
    sc = need(fx, "keccak_challenger_script")
    ops = words([o[0] for o in sc["ops"]])
    args = words([o[1] if len(o) > 1 else 0 for o in sc["ops"]])
    out = np.zeros(len(sc["outputs"]), np.uint32)
    k = oracle().orc_keccak_challenger_script(u32p(ops), u32p(args), len(ops), u32p(out))
    assert k == len(sc["outputs"])
    assert [int(x) for x in out] == [int(x) for x in sc["outputs"]]


def test_c18_keccak_proof_bytes(fx):
    want = bytes.fromhex(need(fx, "proof_n10_keccak_hex"))
    n = 10
    poly = oracle_sample(0, 1 << n)
    pts = [oracle_sample(1, 4 * n).reshape(n, 4)]
    got, _, _ = oracle_prove(oracle_params(n, security_level=32, pow_bits=5, ff=2, rs=1, mmcs=1), poly, pts)
    assert got == want
