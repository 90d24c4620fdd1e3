"""The oracle against the known-answer vectors of tests/golden/golden.json."""
import json
import os

import numpy as np
import pytest

from helpers import P, from_monty, oracle, to_monty, u32p

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def m(x):
    return np.ascontiguousarray(to_monty(np.asarray(x, dtype=np.uint64).reshape(-1)))


def c(mont):
    return [int(v) for v in from_monty(mont)]


def test_ef_mul():
    L = oracle()
    for a, b, want in G["ef_mul"]:
        out = np.zeros(4, np.uint32)
        L.orc_ef_mul(u32p(m(a)), u32p(m(b)), u32p(out))
        assert c(out) == want
    out = np.zeros(4, np.uint32)
    x = m([0, 1, 0, 0])
    x2 = np.zeros(4, np.uint32)
    L.orc_ef_mul(u32p(x), u32p(x), u32p(x2))
    L.orc_ef_mul(u32p(x2), u32p(x2), u32p(out))
    assert c(out) == G["ef_x4"] == [3, 0, 0, 0]


def test_ef_inv():
    L = oracle()
    for a, want in G["ef_inv"]:
        out = np.zeros(4, np.uint32)
        L.orc_ef_inv(u32p(m(a)), u32p(out))
        assert c(out) == want


def test_two_adic_generators():
    L = oracle()
    for k, w in G["two_adic_generator"].items():
        assert c([L.orc_two_adic_generator(int(k))]) == [w]


@pytest.mark.parametrize("naive", [1, 0])
def test_dft(naive):
    L = oracle()
    for x, want in G["dft"]:
        a = m(x)
        L.orc_dft(u32p(a), len(x), naive)
        assert c(a) == want


def test_dft_radix2_matches_naive_large():
    L = oracle()
    rng = np.random.default_rng(3)
    x = to_monty(rng.integers(0, P, 1024, dtype=np.uint64))
    a, b = x.copy(), x.copy()
    L.orc_dft(u32p(a), 1024, 1)
    L.orc_dft(u32p(b), 1024, 0)
    assert np.array_equal(a, b)


def test_eq_table():
    L = oracle()
    for z, s, want in G["eq_table"]:
        n = len(z)
        out = np.zeros(4 << n, np.uint32)
        zm = m(z) if n else np.zeros(4, np.uint32)
        L.orc_eq_table(u32p(zm), n, u32p(m(s)), u32p(out))
        assert c(out) == [v for e in want for v in e]


def test_mle():
    L = oracle()
    for f, z, want in G["mle_base"]:
        out = np.zeros(4, np.uint32)
        L.orc_mle_base(u32p(m(f)), len(z), u32p(m(z)), u32p(out))
        assert c(out) == want
    for f, z, want in G["mle_ext"]:
        out = np.zeros(4, np.uint32)
        L.orc_mle_ext(u32p(m(f)), len(z), u32p(m(z)), u32p(out))
        assert c(out) == want


def test_sumcheck_coeffs():
    L = oracle()
    for e, w, h0, h2 in G["sumcheck_coeffs"]:
        out = np.zeros(8, np.uint32)
        L.orc_sumcheck_coeffs(u32p(m(e)), u32p(m(w)), len(e), u32p(out))
        assert c(out) == h0 + h2


def test_fold():
    L = oracle()
    for a, r, want in G["fold"]:
        x = m(a)
        L.orc_fold_ext(u32p(x), len(a), u32p(m(r)))
        assert c(x[: 4 * len(want)]) == [v for e in want for v in e]


def test_extrapolate():
    L = oracle()
    for e0, e1, e2, r, want in G["extrapolate_012"]:
        out = np.zeros(4, np.uint32)
        L.orc_extrapolate_012(u32p(m(e0)), u32p(m(e1)), u32p(m(e2)), u32p(m(r)), u32p(out))
        assert c(out) == want


def test_rng_known_answers():
    L = oracle()
    st = np.zeros(4, np.uint64)
    L.orc_rng_state(0, st.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_uint64)))
    assert [int(v) for v in st] == G["splitmix64_seed0"]
    import ctypes
    state = np.array([1, 2, 3, 4], np.uint64)
    out = np.zeros(10, np.uint64)
    L.orc_xoshiro_outputs(state.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 10)
    assert [int(v) for v in out] == G["xoshiro256pp_1234"]
