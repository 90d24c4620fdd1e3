"""The benchmark configurations at their own parameters, against the oracle.

- Kernel-level sumcheck + eq at 2^26 (1 GiB EF4 tables): too large for the
  oracle's full-table sumcheck, so checked through what the mathematics fixes
  whatever the implementation: the claim starts at f(z) (the oracle's
  multilinear evaluation), every round polynomial continues it (the claim
  after round i is h_i(r_i), with h_i(1) = claim - h_i(0)), the challenges are
  those of a host challenger observing the round polynomials, and the last
  claim equals f(r) * eq(z, r) with f(r) evaluated by the oracle and eq in
  closed form; a wrong round polynomial breaks the chain.  The eq table is
  checked by its sum (the scale), by marginal sums (scale * z_j over the
  half where variable j is 1), and at sampled entries against the product
  formula.
- Batched proving (64 polynomials of 2^16..2^22 in the benchmark; here a
  spread of sizes over the same range): folding 4, rate 1/2, rs-reduction 3,
  lambda 100, 16 PoW bits, one constraint each, every proof byte-identical to
  the oracle's and accepted by its verifier.
- Serving (2^16..2^24, 1-4 constraints, Poisson arrivals): the same check for
  every request of a mix that includes one 2^24 request.
"""
import ctypes
import time

import numpy as np
import pytest

from helpers import P, oracle, oracle_params, oracle_prove, oracle_sample, oracle_verify, product, u32p

pytestmark = pytest.mark.gpu

ONE = np.array([0x01FFFFFE, 0, 0, 0], np.uint32)   # 1 in Montgomery form


def ef_mul(a, b):
    out = np.zeros(4, np.uint32)
    oracle().orc_ef_mul(u32p(np.ascontiguousarray(a, np.uint32)), u32p(np.ascontiguousarray(b, np.uint32)),
                        u32p(out))
    return out


def ef_sub(a, b):
    # Montgomery form is linear: component-wise subtraction mod p
    return ((a.astype(np.int64) - b.astype(np.int64)) % P).astype(np.uint32)


def ef_add(a, b):
    return ((a.astype(np.uint64) + b.astype(np.uint64)) % P).astype(np.uint32)


def extrapolate(h0, h1, h2, r):
    out = np.zeros(4, np.uint32)
    oracle().orc_extrapolate_012(u32p(h0), u32p(h1), u32p(h2), u32p(np.ascontiguousarray(r)), u32p(out))
    return out


def eq_closed(z, x):
    """prod_i (z_i x_i + (1 - z_i)(1 - x_i)) over extension points of Montgomery words."""
    acc = ONE.copy()
    for zi, xi in zip(z, x):
        t = ef_add(ef_mul(zi, xi), ef_mul(ef_sub(ONE, zi), ef_sub(ONE, xi)))
        acc = ef_mul(acc, t)
    return acc


def replay_challenges(polys, n):
    """Challenges of a fresh duplex challenger (Poseidon2 seed 1) that observes
    each round's (h(0), h(2)) and samples an extension element."""
    ops, args = [], []
    for i in range(n):
        for w in polys[8 * i: 8 * i + 8]:
            ops.append(0)
            args.append(int(w))
        ops += [1, 1, 1, 1]
        args += [0, 0, 0, 0]
    out = np.zeros(4 * n, np.uint32)
    k = oracle().orc_challenger_script(1, u32p(np.array(ops, np.uint32)), u32p(np.array(args, np.uint32)), len(ops),
                                       u32p(out))
    assert k == 4 * n
    return out.reshape(n, 4)


def test_config3_sumcheck_2p26_claim_chain():
    W = product()
    n, l0 = 26, 3
    cfg = W.WhirConfig(W.params(n), device=0)
    f = W.sample_field(10, 1 << n)
    z = W.sample_field(11, 4 * n)
    d_f = W.DeviceBuffer.from_array(f)
    polys = np.zeros(8 * n, np.uint32)
    fin = np.zeros(8, np.uint32)
    ch = cfg.challenger(observe_domain_separator=False)
    W.check(W.lib().whirgpu_sumcheck_eq(cfg.handle, d_f.ptr, n, u32p(z), l0, None, ctypes.byref(ch), u32p(polys),
                                        u32p(fin)))
    d_f.free()
    cfg.close()
    rs = replay_challenges(polys, n)
    L = oracle()
    claim = np.zeros(4, np.uint32)
    L.orc_mle_base(u32p(f), n, u32p(z), u32p(claim))   # f(z) = sum_x f(x) eq(z, x)
    for i in range(n):
        h0, h2 = polys[8 * i: 8 * i + 4], polys[8 * i + 4: 8 * i + 8]
        claim = extrapolate(h0, ef_sub(claim, h0), h2, rs[i])
    f_r = np.zeros(4, np.uint32)
    L.orc_mle_base(u32p(f), n, u32p(np.ascontiguousarray(rs.reshape(-1))), u32p(f_r))
    eq_r = eq_closed(z.reshape(n, 4), rs)
    assert np.array_equal(fin[:4], f_r), "folded f differs from f(r)"
    assert np.array_equal(fin[4:], eq_r), "folded eq differs from eq(z, r)"
    assert np.array_equal(ef_mul(f_r, eq_r), claim), "round polynomials do not chain to f(r) eq(z, r)"


def test_config3_eq_expand_2p26_sums_and_entries():
    W = product()
    n = 26
    cfg = W.WhirConfig(W.params(n), device=0)
    z = W.sample_field(11, 4 * n)
    scale = W.sample_field(12, 4)
    d_eq = W.DeviceBuffer(16 << n)
    W.check(W.lib().whirgpu_eq_expand(cfg.handle, u32p(z), n, u32p(scale), d_eq.ptr))
    t = d_eq.download(np.uint32).reshape(1 << n, 4)
    d_eq.free()
    cfg.close()
    total = (t.sum(axis=0, dtype=np.uint64) % P).astype(np.uint32)
    assert np.array_equal(total, scale), "sum of eq(z, x) over the hypercube must be the scale"
    zz = z.reshape(n, 4)
    for j in (0, 1, 12, 25):
        # variable j is index bit n-1-j: the half where it is 1
        half = t.reshape(1 << j, 2, 1 << (n - 1 - j), 4)[:, 1]
        s = (half.sum(axis=(0, 1), dtype=np.uint64) % P).astype(np.uint32)
        assert np.array_equal(s, ef_mul(scale, zz[j])), j
    rng = np.random.default_rng(3)
    for x in rng.integers(0, 1 << n, 24):
        bits = [(int(x) >> (n - 1 - i)) & 1 for i in range(n)]
        want = scale.copy()
        for zi, b in zip(zz, bits):
            want = ef_mul(want, zi if b else ef_sub(ONE, zi))
        assert np.array_equal(t[int(x)], want), int(x)


BENCH_KW = dict(security_level=100, pow_bits=16, ff=4, rs=3)


def bench_params(W, n=0):
    return W.params(n, security_level=100, pow_bits=16, folding_factor=4, rs_domain_initial_reduction_factor=3)


def test_config4_batch_at_benchmark_parameters_matches_oracle():
    W = product()
    ns = [16, 22, 17, 21, 18, 20, 19, 16, 22, 18]
    reqs = []
    for i, n in enumerate(ns):
        poly = W.sample_field(2000 + i, 1 << n)
        pts = [W.sample_field(3000 + i, 4 * n).reshape(n, 4)]
        reqs.append((W.ProofRequest(poly, pts), poly, pts))
    assert W.prove_batch(bench_params(W), [r for r, _, _ in reqs], workers=6) == 0
    for r, poly, pts in reqs:
        assert r.job.status == 0, r.job.error
        want, evals, _ = oracle_prove(oracle_params(r.n, **BENCH_KW), poly, pts)
        assert np.array_equal(r.evals[:4].reshape(1, 4), evals)
        assert r.proof() == want, f"n={r.n}"
        rc, err = oracle_verify(oracle_params(r.n, **BENCH_KW), want, pts, evals)
        assert rc == 0, err


def test_config5_served_mix_at_benchmark_parameters_matches_oracle():
    W = product()
    rng = np.random.default_rng(7)
    specs = [(int(n), int(k)) for n, k in zip(np.rint(np.exp(rng.uniform(np.log(16), np.log(22), 9))),
                                              rng.integers(1, 5, 9))] + [(24, 2)]
    reqs = []
    for i, (n, k) in enumerate(specs):
        poly = W.sample_field(5000 + i, 1 << n)
        pts = [W.sample_field(6000 + 10 * i + j, 4 * n).reshape(n, 4) for j in range(k)]
        reqs.append((W.ProofRequest(poly, pts), poly, pts))
    srv = W.Server(bench_params(W), workers=4)
    gaps = rng.exponential(1.0 / 200.0, len(reqs))
    for (r, _, _), g in zip(reqs, gaps):
        srv.submit(r)
        time.sleep(float(g))
    for r, _, _ in reqs:
        assert srv.wait(r) == 0, r.job.error
    srv.close()
    for r, poly, pts in reqs:
        want, evals, _ = oracle_prove(oracle_params(r.n, **BENCH_KW), poly, pts)
        assert np.array_equal(r.evals[: 4 * len(pts)].reshape(len(pts), 4), evals)
        assert r.proof() == want, f"n={r.n}, {len(pts)} constraints"
        rc, err = oracle_verify(oracle_params(r.n, **BENCH_KW), want, pts, evals)
        assert rc == 0, err
