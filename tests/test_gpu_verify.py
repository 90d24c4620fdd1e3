"""The product's host verifier on proofs the device writes: the n=24 benchmark
proof, a group of proofs, and the batching server's self-check.  The
verifier's challenger must end where the prover's ended."""
import ctypes
import time

import numpy as np
import pytest

from helpers import oracle_params, oracle_verify, product, u32p

pytestmark = pytest.mark.gpu


def prove_keep_challenger(W, cfg, poly=None, pts=(), device_ptr=None):
    st = W.InitialStatement(cfg, poly, device_ptr=device_ptr)
    evals = [st.evaluate(p) for p in pts]
    ch = cfg.challenger()
    proof = W.WhirProof(cfg)
    pd = W.CommitmentWriter(cfg).commit(proof, ch, st)
    W.Prover(cfg).prove(proof, ch, st, pd)
    data = proof.to_bytes()
    proof.close()
    st.close()
    return data, np.array(evals, dtype=np.uint32).reshape(len(pts), 4), ch


def next_samples(W, cfg, ch, k=8):
    out = np.zeros(k, np.uint32)
    W.check(W.lib().whirgpu_challenger_sample(cfg.handle, ctypes.byref(ch), u32p(out), k))
    return out


def test_n24_benchmark_proof_verifies_and_transcripts_agree():
    W = product()
    n = 24
    cfg = W.WhirConfig(W.params(n), device=0)
    buf = W.DeviceBuffer.from_array(W.sample_field(0, 1 << n))
    pts = [W.sample_field(1 + j, 4 * n).reshape(n, 4) for j in range(2)]
    data, evals, pch = prove_keep_challenger(W, cfg, None, pts, device_ptr=buf.ptr)
    buf.free()
    vch = cfg.challenger()
    t = time.perf_counter()
    W.Verifier(cfg).verify(data, vch, pts, evals)
    print(f"n=24 verify: {(time.perf_counter() - t) * 1e3:.2f} ms, {len(data)} bytes")
    assert np.array_equal(next_samples(W, cfg, vch), next_samples(W, cfg, pch))
    bad = evals.copy()
    bad[1, 3] ^= 1
    assert not W.verify_bytes(cfg, data, pts, bad)
    cfg.close()


def test_group_proofs_verify():
    W = product()
    n = 16
    cfg = W.WhirConfig(W.params(n), device=0)
    polys = [W.sample_field(40 + i, 1 << n) for i in range(5)]
    pts = [[W.sample_field(50 + 10 * i + j, 4 * n).reshape(n, 4) for j in range(2)] for i in range(5)]
    got = W.prove_group(cfg, polys, pts)
    for (data, evals), p in zip(got, pts):
        assert W.verify_bytes(cfg, data, p, evals)
        rc, err = oracle_verify(oracle_params(n), data, p, evals)
        assert rc == 0, err
    # a proof does not verify against another proof's statement
    assert not W.verify_bytes(cfg, got[0][0], pts[1], got[1][1])
    cfg.close()


def test_server_self_check_publishes_verified_proofs():
    W = product()
    rng = np.random.default_rng(11)
    ns = [int(x) for x in rng.integers(14, 19, 12)]
    reqs = []
    for i, n in enumerate(ns):
        poly = W.sample_field(700 + i, 1 << n)
        pts = [W.sample_field(800 + 10 * i + j, 4 * n).reshape(n, 4) for j in range(1 + i % 3)]
        reqs.append((n, pts, W.ProofRequest(poly, pts)))
    srv = W.Server(W.params(0), workers=4, self_check=True)
    srv.submit_many([r for _, _, r in reqs])
    for _, _, r in reqs:
        assert srv.wait(r) == 0, r.job.error
    srv.close()
    for n, pts, r in reqs:
        cfg = W.WhirConfig(W.params(n), device=-1)
        ev = np.array(r.evals, dtype=np.uint32).reshape(len(pts), 4)
        assert W.verify_bytes(cfg, r.proof(), pts, ev)
        cfg.close()
