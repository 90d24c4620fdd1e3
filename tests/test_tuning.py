"""Launch-shape tuning on the host: the ranges and launch plans a tuning must
satisfy, its round trip through a context, and the autotuner's cache and
candidate enumeration (the sweeps themselves run in the GPU tests)."""
import pytest

from helpers import product


def test_tuning_round_trip_and_validation():
    W = product()
    cfg = W.WhirConfig(W.params(24), device=-1)
    t = cfg.tuning().as_dict()
    assert t == {"lde_lp2_base": 0, "lde_lp2_ext": 0, "lde_tb_log_base": -1, "lde_tb_log_ext": -1,
                 "tree_lane_log": 14, "tree_s128_log": 15, "grind_wide_bits": 12, "grind_lock_bits": 17,
                 "grind_blocks": 0}
    for field, value in [("lde_lp2_base", 8), ("lde_tb_log_ext", 3), ("tree_lane_log", 12), ("grind_blocks", 512)]:
        u = dict(t)
        u[field] = value
        cfg.set_tuning(u)
        assert cfg.tuning().as_dict() == u
    cfg.set_tuning(t)
    for field, value in [("lde_lp2_base", 11), ("lde_tb_log_base", 9), ("tree_lane_log", 3), ("tree_s128_log", 6),
                         ("grind_lock_bits", 0), ("grind_blocks", -1)]:
        u = dict(t)
        u[field] = value
        with pytest.raises(W.WhirError, match="out of range"):
            cfg.set_tuning(u)
        assert cfg.tuning().as_dict() == t   # unchanged after a refusal
    # a single-pass length too short for rows of 2^4 base elements has no launch plan
    u = dict(t)
    u["lde_lp2_base"] = 3
    with pytest.raises(W.WhirError, match="Reed-Solomon encoding"):
        cfg.set_tuning(u)
    cfg.close()


def test_cache_round_trip(tmp_path):
    W = product()
    A = W.autotune
    path = str(tmp_path / "cache.json")
    assert A.load_cache(path) == {"version": A.CACHE_VERSION, "devices": {}}
    prm = W.params(20)
    cache = A.load_cache(path)
    cache["devices"]["dev"] = {"kernels": {A.params_key(prm): {"tuning": {"tree_lane_log": 13}}}, "schedules": {}}
    A.save_cache(cache, path)
    assert A.load_cache(path) == cache
    assert A.params_key(prm) != A.params_key(W.params(21))


def test_schedule_candidates():
    W = product()
    c = W.autotune.schedule_candidates(20, folding=(3, 4), first=(None, 5), rs=(1, 3))
    assert (4, None, 3) in c and (3, 5, 1) in c
    assert len(c) == 8 and len(set(c)) == 8
    assert all(k0 != k for k, k0, _ in c)


def test_committed_cache_keys_match_params():
    """Every launch-shape entry of the committed cache is keyed by the current
    parameter fields, so that a configuration can find its entry (a field
    added to the parameters orphans every older key)."""
    W = product()
    A = W.autotune
    c = A.load_cache(A.DEFAULT_CACHE)
    assert c["devices"], "committed tuning cache is empty"
    fields = [f for f, _ in W.params(20)._fields_]
    for dev in c["devices"].values():
        for key in dev["kernels"]:
            assert [kv.split("=")[0] for kv in key.split(",")] == fields, key
    prm = W.params(24)
    hits = [d for d in c["devices"].values() if A.params_key(prm) in d["kernels"]]
    assert hits, "no cached launch shapes for the n=24 benchmark parameters"
