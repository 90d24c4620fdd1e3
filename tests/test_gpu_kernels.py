"""Device kernels through the C ABI against the oracle, bit for bit."""
import ctypes

import numpy as np
import pytest

from helpers import from_monty, oracle, oracle_params, oracle_sample, product, to_monty, u32p

pytestmark = pytest.mark.gpu

P = 0x7F000001


@pytest.fixture(scope="module")
def W():
    return product()


def ctx(W, n=16, **kw):
    prm = W.params(n, security_level=kw.pop("security_level", 32), pow_bits=kw.pop("pow_bits", 5),
                   folding_factor=kw.pop("ff", 2), rs_domain_initial_reduction_factor=kw.pop("rs", 1), **kw)
    return W.WhirConfig(prm, device=0)


def dev(a):
    return product().DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.uint32))


def empty(count):
    return product().DeviceBuffer(4 * count)


def host(b):
    return b.download(np.uint32)


def rand_monty(rng, n):
    return to_monty(rng.integers(0, P, n, dtype=np.uint64))


def oracle_lde_base(f, m, k, R):
    L = 1 << (m - k)
    Wd = 1 << k
    H = L << R
    mat = np.zeros((H, Wd), np.uint32)
    mat[:L, :] = f.reshape(Wd, L).T
    mat = np.ascontiguousarray(mat.reshape(-1))
    oracle().orc_dft_batch(u32p(mat), H, Wd)
    return mat


def oracle_lde_ext(f, m, k, R):
    # f: 2^m extension elements (4 u32 each); columns are coefficient-major
    L = 1 << (m - k)
    Wd = 1 << k
    H = L << R
    e = f.reshape(Wd, L, 4)                # [a][b][coef]
    mat = np.zeros((H, Wd, 4), np.uint32)
    mat[:L] = e.transpose(1, 0, 2)
    mat = np.ascontiguousarray(mat.reshape(-1))
    oracle().orc_dft_batch(u32p(mat), H, Wd * 4)
    return mat


@pytest.mark.parametrize("m,k,R", [(2, 2, 1), (6, 2, 1), (10, 4, 1), (14, 4, 1), (16, 4, 1), (18, 4, 1),
                                   (13, 1, 3), (12, 6, 2), (20, 4, 1), (22, 4, 1), (24, 4, 1)])
def test_lde_base(W, m, k, R):
    cfg = ctx(W)
    rng = np.random.default_rng(m * 100 + k * 10 + R)
    f = rand_monty(rng, 1 << m)
    d_in = dev(f)
    d_out = empty(1 << (m + R))
    W.check(W.lib().whirgpu_lde_base(cfg.handle, d_in.ptr, m, k, R, d_out.ptr))
    assert np.array_equal(host(d_out), oracle_lde_base(f, m, k, R))


@pytest.mark.parametrize("m,k,R", [(4, 4, 11), (8, 4, 2), (12, 4, 5), (16, 4, 2), (14, 4, 3), (9, 1, 2), (6, 6, 1),
                                   (16, 4, 5), (18, 4, 2), (20, 4, 2)])
def test_lde_ext(W, m, k, R):
    cfg = ctx(W)
    rng = np.random.default_rng(m * 100 + k * 10 + R + 1)
    f = rand_monty(rng, 4 << m)
    d_in = dev(f)
    d_out = empty(4 << (m + R))
    W.check(W.lib().whirgpu_lde_ext(cfg.handle, d_in.ptr, m, k, R, d_out.ptr))
    assert np.array_equal(host(d_out), oracle_lde_ext(f, m, k, R))


@pytest.mark.parametrize("leaf_width", [16, 24])
@pytest.mark.parametrize("log_h,row_len", [(0, 16), (1, 16), (5, 16), (7, 16), (8, 16), (10, 16), (12, 16),
                                           (9, 64), (3, 2), (4, 20), (14, 16), (15, 16), (16, 16), (18, 16),
                                           (19, 8)])
def test_merkle_root_and_openings(W, leaf_width, log_h, row_len):
    cfg = ctx(W, leaf_hash_width=leaf_width)
    rng = np.random.default_rng(log_h * 7 + row_len + leaf_width)
    H = 1 << log_h
    rows = rand_monty(rng, H * row_len)
    d_rows = dev(rows)
    d_tree = empty((2 * H - 1) * 8)
    root = np.zeros(8, np.uint32)
    W.check(W.lib().whirgpu_merkle_commit(cfg.handle, d_rows.ptr, H, row_len, d_tree.ptr, u32p(root)))
    want = np.zeros(8, np.uint32)
    idx = int(rng.integers(0, H))
    row = np.zeros(row_len, np.uint32)
    path = np.zeros(8 * max(log_h, 1), np.uint32)
    oracle().orc_merkle(1, leaf_width, u32p(rows), H, row_len, u32p(want), idx, u32p(row), u32p(path))
    assert np.array_equal(root, want)
    tree = host(d_tree).reshape(-1, 8)
    # sibling digests of the oracle's opening sit at the tree's level offsets
    off = 0
    for l in range(log_h):
        assert np.array_equal(tree[off + ((idx >> l) ^ 1)], path[8 * l: 8 * l + 8])
        off += H >> l


# row lengths: odd (a last lane holding one word), one lane, exactly 17 lanes
# (one absorption, no extra permutation), 18 lanes and 64 words (two
# absorptions)
@pytest.mark.parametrize("log_h,row_len", [(0, 16), (1, 3), (5, 16), (8, 1), (9, 64), (4, 34), (6, 35), (3, 36),
                                           (10, 20), (14, 16), (16, 16)])
def test_keccak_merkle_root_and_openings(W, log_h, row_len):
    cfg = ctx(W, mmcs="keccak")
    rng = np.random.default_rng(log_h * 11 + row_len)
    H = 1 << log_h
    rows = rand_monty(rng, H * row_len)
    d_rows = dev(rows)
    d_tree = empty((2 * H - 1) * 8)
    root = np.zeros(8, np.uint32)
    W.check(W.lib().whirgpu_merkle_commit(cfg.handle, d_rows.ptr, H, row_len, d_tree.ptr, u32p(root)))
    want = np.zeros(8, np.uint32)
    idx = int(rng.integers(0, H))
    row = np.zeros(row_len, np.uint32)
    path = np.zeros(8 * max(log_h, 1), np.uint32)
    oracle().orc_merkle_keccak(u32p(rows), H, row_len, u32p(want), idx, u32p(row), u32p(path))
    assert np.array_equal(root, want)
    tree = host(d_tree).reshape(-1, 8)
    off = 0
    for l in range(log_h):
        assert np.array_equal(tree[off + ((idx >> l) ^ 1)], path[8 * l: 8 * l + 8])
        off += H >> l
    cfg.close()


@pytest.mark.parametrize("bits", [1, 4, 9, 12, 13, 14, 17, 18])
@pytest.mark.parametrize("n_obs", [0, 3, 7, 8, 13])
def test_grind_smallest_witness(W, bits, n_obs):
    cfg = ctx(W)
    L = oracle()
    vals = to_monty(np.arange(1, n_obs + 1, dtype=np.uint64) * 7919 + bits)
    ch = cfg.challenger(observe_domain_separator=False)
    if n_obs:
        W.check(W.lib().whirgpu_challenger_observe(cfg.handle, ctypes.byref(ch), u32p(vals), n_obs))
    w = np.zeros(1, np.uint32)
    W.check(W.lib().whirgpu_grind(cfg.handle, ctypes.byref(ch), bits, u32p(w)))
    ops = np.array([0] * n_obs + [3, 1], np.uint32)
    args = np.array(list(vals) + [bits, 0], np.uint32)
    out = np.zeros(4, np.uint32)
    assert L.orc_challenger_script(1, u32p(ops), u32p(args), len(ops), u32p(out)) == 2
    assert int(from_monty(w)[0]) == int(out[0])
    nxt = np.zeros(1, np.uint32)
    W.check(W.lib().whirgpu_challenger_sample(cfg.handle, ctypes.byref(ch), u32p(nxt), 1))
    assert int(nxt[0]) == int(out[1])


@pytest.mark.parametrize("n", [0, 1, 5, 9, 14, 17])
def test_eq_expand(W, n):
    cfg = ctx(W, n=max(n, 8))
    rng = np.random.default_rng(n + 11)
    z = rand_monty(rng, 4 * max(n, 1))
    s = rand_monty(rng, 4)
    d_out = empty(4 << n)
    W.check(W.lib().whirgpu_eq_expand(cfg.handle, u32p(z), n, u32p(s), d_out.ptr))
    want = np.zeros(4 << n, np.uint32)
    oracle().orc_eq_table(u32p(z), n, u32p(s), u32p(want))
    assert np.array_equal(host(d_out), want)


@pytest.mark.parametrize("n", [1, 4, 8, 9, 13, 18])
def test_mle_eval(W, n):
    cfg = ctx(W, n=max(n, 8))
    rng = np.random.default_rng(n + 21)
    z = rand_monty(rng, 4 * n)
    fb = rand_monty(rng, 1 << n)
    fe = rand_monty(rng, 4 << n)
    out = np.zeros(4, np.uint32)
    want = np.zeros(4, np.uint32)
    W.check(W.lib().whirgpu_mle_eval_base(cfg.handle, dev(fb).ptr, n, u32p(z), u32p(out)))
    oracle().orc_mle_base(u32p(fb), n, u32p(z), u32p(want))
    assert np.array_equal(out, want)
    W.check(W.lib().whirgpu_mle_eval_ext(cfg.handle, dev(fe).ptr, n, u32p(z), u32p(out)))
    oracle().orc_mle_ext(u32p(fe), n, u32p(z), u32p(want))
    assert np.array_equal(out, want)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 17])
def test_sumcheck_rounds(W, n):
    cfg = ctx(W)
    L = oracle()
    rng = np.random.default_rng(n + 31)
    e = rand_monty(rng, 4 << n)
    w = rand_monty(rng, 4 << n)
    de, dw = dev(e), dev(w)
    cur_e, cur_w = e.copy(), w.copy()
    for m in range(n, 0, -1):
        got = np.zeros(8, np.uint32)
        W.check(W.lib().whirgpu_sumcheck_coeffs(cfg.handle, de.ptr, dw.ptr, m, u32p(got)))
        want = np.zeros(8, np.uint32)
        L.orc_sumcheck_coeffs(u32p(cur_e), u32p(cur_w), 1 << m, u32p(want))
        assert np.array_equal(got, want), m
        r = rand_monty(rng, 4)
        nxt = np.zeros(8, np.uint32)
        W.check(W.lib().whirgpu_sumcheck_fold(cfg.handle, de.ptr, dw.ptr, m, u32p(r), u32p(nxt)))
        L.orc_fold_ext(u32p(cur_e), 1 << m, u32p(r))
        L.orc_fold_ext(u32p(cur_w), 1 << m, u32p(r))
        cur_e = np.ascontiguousarray(cur_e[: 4 << (m - 1)])
        cur_w = np.ascontiguousarray(cur_w[: 4 << (m - 1)])
        assert np.array_equal(host(de)[: 4 << (m - 1)], cur_e)
        assert np.array_equal(host(dw)[: 4 << (m - 1)], cur_w)
        if m >= 2:
            want2 = np.zeros(8, np.uint32)
            L.orc_sumcheck_coeffs(u32p(cur_e), u32p(cur_w), 1 << (m - 1), u32p(want2))
            assert np.array_equal(nxt, want2)


@pytest.mark.parametrize("n,l0", [(1, 0), (3, 0), (6, 2), (12, 3), (12, 4), (16, 3), (20, 3), (22, 3)])
@pytest.mark.parametrize("mode", ["scripted", "challenger"])
def test_sumcheck_eq_round_polynomials(W, n, l0, mode):
    # the kernel-level sumcheck (SVO first rounds, then device folds) against
    # classic full-table rounds on the oracle
    cfg = ctx(W, n=max(n, 8))
    rng = np.random.default_rng(n * 10 + l0)
    f = rand_monty(rng, 1 << n)
    z = rand_monty(rng, 4 * n)
    scripted = rand_monty(rng, 4 * n) if mode == "scripted" else None
    polys = np.zeros(8 * n, np.uint32)
    fin = np.zeros(8, np.uint32)
    d_f = dev(f)
    if scripted is not None:
        W.check(W.lib().whirgpu_sumcheck_eq(cfg.handle, d_f.ptr, n, u32p(z), l0, u32p(scripted), None, u32p(polys),
                                            u32p(fin)))
    else:
        ch = cfg.challenger(observe_domain_separator=False)
        W.check(W.lib().whirgpu_sumcheck_eq(cfg.handle, d_f.ptr, n, u32p(z), l0, None, ctypes.byref(ch), u32p(polys),
                                            u32p(fin)))
    want = np.zeros(8 * n, np.uint32)
    wfin = np.zeros(8, np.uint32)
    rc = oracle().orc_sumcheck_eq(u32p(f), n, u32p(z), u32p(scripted) if scripted is not None else None, 1,
                                  u32p(want), u32p(wfin))
    assert rc == 0
    assert np.array_equal(polys, want)
    assert np.array_equal(fin, wfin)
