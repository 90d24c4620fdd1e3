"""Launch shapes on the device: every tuning knob gives the proof bytes of the
built-in shapes (and the oracle's), and the autotuner's sweeps run and keep
that property."""
import pytest

from helpers import oracle_params, oracle_prove, oracle_sample, product

pytestmark = pytest.mark.gpu

# One non-default value per knob and a few combinations; each must be accepted
# for the configurations below and leave the proof unchanged.
VARIANTS = [
    {"lde_lp2_base": 8}, {"lde_lp2_base": 9, "lde_tb_log_base": 2}, {"lde_tb_log_base": 8},
    {"lde_lp2_ext": 5}, {"lde_lp2_ext": 8, "lde_tb_log_ext": 1}, {"lde_tb_log_ext": 6},
    {"tree_lane_log": 12}, {"tree_lane_log": 17, "tree_s128_log": 17}, {"tree_s128_log": 13},
    {"grind_wide_bits": 14, "grind_lock_bits": 15}, {"grind_wide_bits": 8}, {"grind_lock_bits": 20},
    {"grind_blocks": 256}, {"grind_blocks": 4096, "tree_lane_log": 4},
]


@pytest.mark.parametrize("n,kw", [(18, {}), (24, {}), (14, dict(security_level=80, pow_bits=12, ff=3, rs=2))])
def test_every_launch_shape_gives_the_same_proof(n, kw):
    W = product()
    okw = dict(kw)
    prm = W.params(n, security_level=okw.get("security_level", 100), pow_bits=okw.get("pow_bits", 16),
                   folding_factor=okw.get("ff", 4), rs_domain_initial_reduction_factor=okw.get("rs", 3))
    cfg = W.WhirConfig(prm, device=0)
    poly = oracle_sample(n, 1 << n)
    pts = [oracle_sample(n + 1, 4 * n).reshape(n, 4)]
    buf = W.DeviceBuffer.from_array(poly)
    ref, _ = W.prove_once(cfg, None, pts, device_ptr=buf.ptr)
    if n <= 18:
        want, _, _ = oracle_prove(oracle_params(n, **kw), poly, pts)
        assert ref == want
    default = cfg.tuning().as_dict()
    applied = 0
    for v in VARIANTS:
        t = dict(default)
        t.update(v)
        try:
            cfg.set_tuning(t)
        except W.WhirError:
            continue   # no launch plan for this shape under it (checked before anything runs)
        applied += 1
        got, _ = W.prove_once(cfg, None, pts, device_ptr=buf.ptr)
        assert got == ref, v
    assert applied >= len(VARIANTS) - 3
    buf.free()
    cfg.close()


def test_kernel_autotuner_runs_and_keeps_the_proof(tmp_path):
    W = product()
    path = str(tmp_path / "cache.json")
    entry = W.autotune.tune_kernels(W.params(16), proofs=1, cache_path=path)
    assert set(entry["tuning"]) == set(W.Tuning().as_dict())
    assert W.autotune.cached_tuning(W.params(16), cache_path=path) == entry["tuning"]
    measured = [c for c in entry["candidates"] if c["ms"] is not None]
    assert len(measured) >= 30
    cfg = W.WhirConfig(W.params(16), device=0)
    assert W.autotune.apply_cached(cfg, cache_path=path) == entry["tuning"]
    cfg.close()


def test_schedule_autotuner_picks_a_valid_schedule(tmp_path):
    W = product()
    path = str(tmp_path / "cache.json")
    entry = W.autotune.tune_schedule(14, security_level=80, pow_bits=12,
                                     candidates=[(3, None, 1), (4, None, 3), (4, 3, 2), (9, None, 1)], proofs=1,
                                     cache_path=path)
    valid = [r for r in entry["candidates"] if r["valid"]]
    assert len(valid) == 3 and not [r for r in entry["candidates"] if r["folding_factor"] == 9][0]["valid"]
    assert entry["best"]["prove_ms"] == min(r["prove_ms"] for r in valid)
