"""whirgpu_mle_eval_base_noalloc, the entry point that allocates nothing,
against the allocating whirgpu_mle_eval_base on the same inputs: bit-exact,
on two calls in a row (the second smaller, so stale scratch contents from the
first would show), then at its 24-variable limit."""
import numpy as np
import pytest

from helpers import oracle, product, u32p

pytestmark = pytest.mark.gpu


def evaluate_both(W, cfg, n, seed):
    f = W.sample_field(seed, 1 << n)
    z = W.sample_field(seed + 100, 4 * n)
    d_f = W.DeviceBuffer.from_array(f)
    d_out = W.DeviceBuffer(16)
    want = np.zeros(4, np.uint32)
    W.check(W.lib().whirgpu_mle_eval_base(cfg.handle, d_f.ptr, n, u32p(z), u32p(want)))
    W.check(W.lib().whirgpu_mle_eval_base_noalloc(d_f.ptr, n, u32p(z), d_out.ptr, None))
    got = d_out.download(np.uint32, 4)   # a copy on the null stream: after the kernels
    d_f.free()
    d_out.free()
    return f, z, got, want


def test_noalloc_matches_allocating_entry_point_twice():
    W = product()
    cfg = W.WhirConfig(W.params(20, security_level=32, pow_bits=5, folding_factor=2,
                                rs_domain_initial_reduction_factor=1), device=0)
    f, z, got, want = evaluate_both(W, cfg, 18, 1)
    assert np.array_equal(got, want)
    f, z, got, want = evaluate_both(W, cfg, 11, 2)   # smaller second call
    assert np.array_equal(got, want)
    ref = np.zeros(4, np.uint32)
    oracle().orc_mle_base(u32p(f), 11, u32p(z), u32p(ref))
    assert np.array_equal(got, ref)
    cfg.close()


def test_noalloc_at_its_limit():
    W = product()
    cfg = W.WhirConfig(W.params(24), device=0)
    _, _, got, want = evaluate_both(W, cfg, 24, 3)
    assert np.array_equal(got, want)
    cfg.close()
