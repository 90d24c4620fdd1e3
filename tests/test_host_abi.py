"""The C-ABI library: exported symbols and the host-side protocol logic
(schedule, domain separator, challenger, Poseidon2 on the host, input RNG)
against the oracle.  No device is used: contexts are created host-only."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT, from_monty, oracle, oracle_params, oracle_sample, product, to_monty, u32p

CASES = [
    (24, dict(rs=3)), (24, dict(rs=1)), (18, dict(rs=3)), (16, dict(rs=3)), (20, dict(rs=2)),
    (10, dict(security_level=32, pow_bits=5, ff=2, rs=1)),
    (11, dict(security_level=32, pow_bits=10, ff=4, ff0=3, rs=3)),
    (9, dict(security_level=32, pow_bits=5, ff=2, rs=1, soundness=1)),
    (10, dict(security_level=32, pow_bits=0, ff=3, rs=2, soundness=0)),
]


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            syms |= set(re.findall(r"\b(whirgpu_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    W = product()
    L = W.lib()
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert syms <= set(W._SIGS), sorted(syms - set(W._SIGS))


def product_params(n, security_level=100, pow_bits=16, ff=4, ff0=None, rate=1, rs=3, soundness=2, svo=1,
                   leaf_width=16, digest_elems=8):
    W = product()
    return W.params(n, security_level=security_level, pow_bits=pow_bits, folding_factor=ff,
                    folding_factor_first=ff0, starting_log_inv_rate=rate, rs_domain_initial_reduction_factor=rs,
                    soundness=soundness, svo=bool(svo), leaf_hash_width=leaf_width, digest_elems=digest_elems)


@pytest.mark.parametrize("n,kw", CASES)
def test_schedule_matches_oracle(n, kw):
    W = product()
    cfg = W.WhirConfig(product_params(n, **kw), device=-1)
    out = (ctypes.c_int * 256)()
    k = oracle().orc_config(oracle_params(n, **kw), out, 256)
    want = list(out[:k])
    got = cfg.schedule()
    # generators: the oracle reports canonical values, the library canonical too
    assert got == [v & 0xFFFFFFFF for v in want]


@pytest.mark.parametrize("n,kw", CASES)
@pytest.mark.parametrize("digest", [8, 32])
def test_domain_separator_matches_oracle(n, kw, digest):
    W = product()
    cfg = W.WhirConfig(product_params(n, digest_elems=digest, **kw), device=-1)
    out = np.zeros(4096, np.uint32)
    k = oracle().orc_domain_separator(oracle_params(n, digest_elems=digest, **kw), u32p(out), 4096)
    assert np.array_equal(cfg.domain_separator(), out[:k])


def test_bad_params_are_reported():
    W = product()
    with pytest.raises(W.WhirError, match="error 1"):
        W.WhirConfig(product_params(8, ff=9), device=-1)
    with pytest.raises(W.WhirError, match="rate"):
        W.WhirConfig(product_params(12, ff=2, rs=3), device=-1)
    with pytest.raises(W.WhirError):
        W.WhirConfig(product_params(30, rate=1), device=-1)


@pytest.mark.parametrize("n,ff,ok", [(25, 5, True), (25, 4, True), (27, 5, True), (28, 6, True), (22, 8, True),
                                     (22, 9, False), (20, 10, False)])
def test_encoding_shapes_are_checked_at_creation(n, ff, ok):
    # every Reed-Solomon encoding of the schedule must be a shape the LDE
    # kernels cover (rows of at most 2^8 base elements, columns up to 2^24);
    # others are refused when the configuration is created, not at commit
    W = product()
    prm = product_params(n, ff=ff, security_level=60)
    if ok:
        W.WhirConfig(prm, device=-1).close()
    else:
        with pytest.raises(W.WhirError, match="Reed-Solomon encoding"):
            W.WhirConfig(prm, device=-1)


def test_host_only_context_refuses_device_work():
    W = product()
    cfg = W.WhirConfig(product_params(10, security_level=32, pow_bits=5, ff=2, rs=1), device=-1)
    with pytest.raises(W.WhirError, match="error 6"):
        W.InitialStatement(cfg, np.zeros(1 << 10, np.uint32))


def test_sample_field_matches_oracle():
    W = product()
    for seed in (0, 1, 42):
        assert np.array_equal(W.sample_field(seed, 5000), oracle_sample(seed, 5000))


@pytest.mark.parametrize("width", [16, 24])
def test_poseidon2_constants_and_permutation_match_oracle(width):
    W = product()
    L = oracle()
    perms = W.poseidon2_from_seed(1)
    ei = np.zeros(4 * width, np.uint32)
    et = np.zeros(4 * width, np.uint32)
    it = np.zeros(20 if width == 16 else 23, np.uint32)
    L.orc_poseidon2_constants(1, width, u32p(ei), u32p(et), u32p(it))
    pre = "w16_" if width == 16 else "w24_"
    assert list(getattr(perms, pre + "external_initial")) == list(ei)
    assert list(getattr(perms, pre + "external_terminal")) == list(et)
    assert list(getattr(perms, pre + "internal")) == list(it)
    rng = np.random.default_rng(width)
    # random states, then the edges of the lazy reductions' ranges: all p - 1,
    # all 0, alternating, and values next to p / 2
    states = [to_monty(rng.integers(0, 0x7F000001, width, dtype=np.uint64)) for _ in range(200)]
    states += [np.full(width, 0x7F000000, np.uint32), np.zeros(width, np.uint32),
               np.array([0x7F000000 if i % 2 else 0 for i in range(width)], np.uint32),
               np.array([0x3F800000 + (i % 3) for i in range(width)], np.uint32)]
    for st in states:
        a = st.copy()
        b = st.copy()
        W.check(W.lib().whirgpu_poseidon2_permute(ctypes.byref(perms), width, u32p(a)))
        L.orc_poseidon2_permute(1, width, u32p(b))
        assert np.array_equal(a, b)


def test_challenger_matches_oracle():
    W = product()
    L = oracle()
    cfg = W.WhirConfig(product_params(10, security_level=32, pow_bits=5, ff=2, rs=1), device=-1)
    ch = cfg.challenger(observe_domain_separator=False)
    rng = np.random.default_rng(5)
    ops, args, expect_kinds = [], [], []
    for step in range(60):
        k = int(rng.integers(0, 3))
        if k == 0:
            for _ in range(int(rng.integers(1, 12))):
                ops.append(0)
                args.append(int(to_monty([int(rng.integers(0, 0x7F000001))])[0]))
        else:
            for _ in range(int(rng.integers(1, 10))):
                ops.append(1)
                args.append(0)
    ops_a = np.array(ops, np.uint32)
    args_a = np.array(args, np.uint32)
    out = np.zeros(len(ops), np.uint32)
    cnt = L.orc_challenger_script(1, u32p(ops_a), u32p(args_a), len(ops), u32p(out))
    got = []
    for o, a in zip(ops, args):
        if o == 0:
            v = np.array([a], np.uint32)
            W.check(W.lib().whirgpu_challenger_observe(cfg.handle, ctypes.byref(ch), u32p(v), 1))
        else:
            v = np.zeros(1, np.uint32)
            W.check(W.lib().whirgpu_challenger_sample(cfg.handle, ctypes.byref(ch), u32p(v), 1))
            got.append(int(v[0]))
    assert got == [int(x) for x in out[:cnt]]


def test_noalloc_evaluation_rejects_sizes_beyond_its_scratch():
    # checked on the host before anything is enqueued (no device needed)
    W = product()
    z = np.zeros(4 * 25, np.uint32)
    dummy = ctypes.c_void_p(256)
    rc = W.lib().whirgpu_mle_eval_base_noalloc(dummy, 25, u32p(z), dummy, None)
    assert rc == 1
    assert b"[1, 24]" in W.lib().whirgpu_last_error()
    assert W.lib().whirgpu_mle_eval_base_noalloc(dummy, 0, u32p(z), dummy, None) == 1
    assert W.lib().whirgpu_mle_eval_base_noalloc(None, 8, u32p(z), dummy, None) == 1
