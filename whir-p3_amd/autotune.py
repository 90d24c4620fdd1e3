"""Autotuner for the prover's launch shapes and folding schedule, with a JSON
cache per device.

Launch shapes (`tune_kernels`).  For one configuration, each group of launch
parameters of `Tuning` is swept in turn with the others held: the LDE's
single-pass length and pass-1 lines per block (base field, then extension
field), the Merkle cut-overs, and the grinding thresholds and block count.
Each candidate is timed with the library's per-stage HIP-event timers over
real commit + prove runs on synthetic inputs, and must produce the same proof
bytes as the built-in shapes (every candidate is the same computation in
another launch shape; a difference raises).  The fastest value of each group
is kept when it beats the current one by more than 2% (so timer noise does not
move it), and the tuned shapes are measured again at the end against the
defaults.

Folding schedule (`tune_schedule`).  The reference lets a caller choose, at a
fixed security level, the folding factor (Constant(k) or
ConstantFromSecondRound(k0, k), src/parameters/mod.rs:31-143) and the initial
Reed-Solomon domain reduction (src/whir/parameters.rs:81-234).  Every valid
combination in the given ranges is proved and timed (median wall time of a
few proofs, proof size, host verification time), and the one with the least
cost is kept: prove time by default, or prove time plus a weight per KiB of
proof.  A schedule changes the proof, so it is the caller's choice to apply;
the launch shapes never change it.

Cache: {"version": 1, "devices": {name: {"kernels": {key: entry},
"schedules": {key: entry}}}}, keyed by the device's name and architecture and
by the configuration's parameters.
"""
import json
import os
import time

import numpy as np

if __name__ == "__main__":   # run as a script: load the package from this directory
    import importlib.util
    import sys

    _spec = importlib.util.spec_from_file_location(
        "whir_p3_amd", os.path.join(os.path.dirname(os.path.abspath(__file__)), "__init__.py"))
    _pkg = importlib.util.module_from_spec(_spec)
    sys.modules["whir_p3_amd"] = _pkg
    _spec.loader.exec_module(_pkg)

import whir_p3_amd as W  # noqa: E402

CACHE_VERSION = 1
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "cache.json")

# A candidate replaces the current value only if its stage is this much faster
# (timer noise between runs of the same shapes is about 1%).
MARGIN = 0.02

# Candidate values of each group of launch parameters (0 / -1: built-in rule).
SWEEPS = [
    ("lde_base", [("lde_lp2_base", [0, 6, 7, 8, 9, 10])]),
    ("lde_base", [("lde_tb_log_base", [-1, 2, 3, 4, 5, 6, 7, 8])]),
    ("lde_ext", [("lde_lp2_ext", [0, 2, 3, 4, 5, 6, 7, 8])]),
    ("lde_ext", [("lde_tb_log_ext", [-1, 1, 2, 3, 4, 5, 6])]),
    ("merkle_tree", [("tree_lane_log", [12, 13, 14, 15, 16, 17])]),
    ("merkle_tree", [("tree_s128_log", [13, 14, 15, 16, 17])]),
    ("grind", [("grind_wide_bits", [10, 11, 12, 13, 14])]),
    ("grind", [("grind_lock_bits", [15, 16, 17, 18, 19])]),
    ("grind", [("grind_blocks", [0, 256, 512, 1024, 2048])]),
]


def params_key(prm):
    return ",".join(f"{f}={int(getattr(prm, f))}" for f, _ in prm._fields_)


def load_cache(path=DEFAULT_CACHE):
    if not os.path.exists(path):
        return {"version": CACHE_VERSION, "devices": {}}
    with open(path) as f:
        c = json.load(f)
    if c.get("version") != CACHE_VERSION:
        return {"version": CACHE_VERSION, "devices": {}}
    return c


def save_cache(cache, path=DEFAULT_CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


class _Workload:
    """Synthetic inputs for one configuration, resident in HBM, and one proof step."""

    def __init__(self, cfg, n, npoints=1, seed=0):
        self.cfg = cfg
        self.buf = W.DeviceBuffer.from_array(W.sample_field(seed, 1 << n))
        self.pts = [W.sample_field(seed + 1 + j, 4 * n).reshape(n, 4) for j in range(npoints)]

    def prove(self):
        return W.prove_once(self.cfg, None, self.pts, device_ptr=self.buf.ptr)[0]

    def stage_ms(self, proofs):
        """Median per-proof device time of every stage over `proofs` proofs, and the proof bytes."""
        per = {s: [] for s in self.cfg.STAGES}
        data = None
        self.cfg.profile(True)
        try:
            for _ in range(proofs):
                data = self.prove()
                for s, v in self.cfg.profile_read().items():
                    per[s].append(v["ms"])
        finally:
            self.cfg.profile(False)
        return {s: float(np.median(v)) for s, v in per.items()}, data

    def close(self):
        self.buf.free()


def tune_kernels(prm, proofs=3, device=0, cache_path=DEFAULT_CACHE, save=True, log=None):
    """Sweeps the launch shapes for configuration `prm`; returns the cache entry
    {"tuning", "stage_ms", "default_stage_ms", "candidates"} and stores it."""
    n = int(prm.num_variables)
    cfg = W.WhirConfig(prm, device=device)
    work = _Workload(cfg, n)
    try:
        work.prove()   # warm-up: pools, kernels, code objects
        default_ms, ref = work.stage_ms(proofs)
        best = cfg.tuning().as_dict()
        best_stage = dict(default_ms)
        tried = []
        for stage, groups in SWEEPS:
            for field, values in groups:
                for v in values:
                    cand = dict(best)
                    cand[field] = v
                    if cand == best and v != values[0]:
                        continue
                    try:
                        cfg.set_tuning(cand)
                    except W.WhirError:
                        tried.append({"field": field, "value": v, "stage": stage, "ms": None})
                        continue
                    ms, data = work.stage_ms(proofs)
                    if data != ref:
                        raise RuntimeError(f"launch shape {field}={v} changed the proof bytes")
                    tried.append({"field": field, "value": v, "stage": stage, "ms": round(ms[stage], 4)})
                    if log:
                        log(f"{field}={v}: {stage} {ms[stage]:.4f} ms")
                    if ms[stage] < best_stage[stage] * (1 - MARGIN):
                        best, best_stage = cand, ms
                cfg.set_tuning(best)
        final_ms, data = work.stage_ms(proofs)
        if data != ref:
            raise RuntimeError("the tuned launch shapes changed the proof bytes")
        entry = {"tuning": best, "stage_ms": {k: round(v, 4) for k, v in final_ms.items()},
                 "default_stage_ms": {k: round(v, 4) for k, v in default_ms.items()}, "candidates": tried,
                 "proofs_per_candidate": proofs}
    finally:
        work.close()
        cfg.close()
    if save:
        cache = load_cache(cache_path)
        dev = cache["devices"].setdefault(W.device_name(device), {"kernels": {}, "schedules": {}})
        dev["kernels"][params_key(prm)] = entry
        save_cache(cache, cache_path)
    return entry


def cached_tuning(prm, device=0, cache_path=DEFAULT_CACHE):
    """The cached launch shapes for `prm` on this device, or None."""
    dev = load_cache(cache_path)["devices"].get(W.device_name(device))
    if not dev:
        return None
    e = dev["kernels"].get(params_key(prm))
    return e["tuning"] if e else None


def cached_tunings_by_size(prm, device=0, cache_path=DEFAULT_CACHE):
    """{num_variables: tuning} of every cached entry whose parameters equal
    `prm`'s except the number of variables (for a server's per-size contexts)."""
    dev = load_cache(cache_path)["devices"].get(W.device_name(device))
    if not dev:
        return {}
    out = {}
    q = type(prm).from_buffer_copy(prm)
    for n in range(1, 31):
        q.num_variables = n
        e = dev["kernels"].get(params_key(q))
        if e:
            out[n] = e["tuning"]
    return out


def apply_cached(cfg, device=0, cache_path=DEFAULT_CACHE):
    """Applies the cached launch shapes to `cfg` if there are any; returns them."""
    t = cached_tuning(cfg.params, device, cache_path)
    if t is not None:
        cfg.set_tuning(t)
    return t


def schedule_candidates(n, folding=(2, 3, 4, 5, 6), first=(None,), rs=(1, 2, 3)):
    """(folding_factor, folding_factor_first, rs_reduction) triples to try; first None = Constant."""
    out = []
    for k in folding:
        for k0 in first:
            for r in rs:
                if k0 is not None and k0 == k:
                    continue
                out.append((k, k0, r))
    return out


def tune_schedule(n, security_level=100, pow_bits=16, soundness="CapacityBound", starting_log_inv_rate=1,
                  candidates=None, proofs=3, kib_weight_ms=0.0, device=0, cache_path=DEFAULT_CACHE, save=True,
                  log=None):
    """Proves every valid schedule among `candidates` and keeps the one with the
    least prove_ms + kib_weight_ms * proof KiB.  Returns the cache entry."""
    candidates = candidates or schedule_candidates(n)
    rows = []
    for k, k0, r in candidates:
        prm = W.params(n, security_level=security_level, pow_bits=pow_bits, folding_factor=k,
                       folding_factor_first=k0, starting_log_inv_rate=starting_log_inv_rate,
                       rs_domain_initial_reduction_factor=r, soundness=soundness)
        try:
            cfg = W.WhirConfig(prm, device=device)
        except W.WhirError as e:
            rows.append({"folding_factor": k, "folding_factor_first": k0, "rs_reduction": r, "valid": False,
                         "reason": str(e)[:120]})
            continue
        work = _Workload(cfg, n)
        try:
            work.prove()
            times = []
            for _ in range(proofs):
                t = time.perf_counter()
                data = work.prove()
                times.append((time.perf_counter() - t) * 1e3)
            st = W.InitialStatement(cfg, None, device_ptr=work.buf.ptr)
            ev = np.array([st.evaluate(p) for p in work.pts], dtype=np.uint32).reshape(len(work.pts), 4)
            st.close()
            t = time.perf_counter()
            ok = W.verify_bytes(cfg, data, work.pts, ev)
            verify_ms = (time.perf_counter() - t) * 1e3
            if not ok:
                raise RuntimeError(f"schedule {(k, k0, r)}: the proof does not verify")
            row = {"folding_factor": k, "folding_factor_first": k0, "rs_reduction": r, "valid": True,
                   "prove_ms": round(float(np.median(times)), 3), "proof_kib": round(len(data) / 1024, 1),
                   "verify_ms": round(verify_ms, 3), "rounds": cfg.schedule()[2]}
            row["cost"] = round(row["prove_ms"] + kib_weight_ms * row["proof_kib"], 3)
            rows.append(row)
            if log:
                log(f"ff={k} ff0={k0} rs={r}: {row['prove_ms']} ms, {row['proof_kib']} KiB")
        finally:
            work.close()
            cfg.close()
    valid = [r for r in rows if r["valid"]]
    if not valid:
        raise RuntimeError("no valid schedule among the candidates")
    best = min(valid, key=lambda r: r["cost"])
    entry = {"best": best, "candidates": rows, "objective": f"prove_ms + {kib_weight_ms} * proof_kib",
             "security_level": security_level, "pow_bits": pow_bits, "soundness": soundness,
             "starting_log_inv_rate": starting_log_inv_rate}
    if save:
        cache = load_cache(cache_path)
        dev = cache["devices"].setdefault(W.device_name(device), {"kernels": {}, "schedules": {}})
        dev["schedules"][f"n={n},lambda={security_level},pow={pow_bits},{soundness},rate={starting_log_inv_rate}"] = entry
        save_cache(cache, cache_path)
    return entry


def _main():
    import argparse

    ap = argparse.ArgumentParser(description="Tune launch shapes (and optionally folding schedules) on this GPU.")
    ap.add_argument("--kernels", default="16-24", help="sizes whose launch shapes to tune, e.g. 16-24 or 20,24")
    ap.add_argument("--schedule", default="", help="sizes whose folding schedule to tune, e.g. 24,20")
    ap.add_argument("--proofs", type=int, default=3, help="proofs per candidate")
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    args = ap.parse_args()

    def sizes(spec):
        out = []
        for part in filter(None, spec.split(",")):
            a, _, b = part.partition("-")
            out += list(range(int(a), int(b or a) + 1))
        return out

    for n in sorted(sizes(args.kernels), reverse=True):
        e = tune_kernels(W.params(n), proofs=args.proofs, cache_path=args.cache)
        d, f = e["default_stage_ms"], e["stage_ms"]
        print(f"n={n} tuning {e['tuning']}", flush=True)
        for s in ("lde_base", "lde_ext", "merkle_tree", "grind"):
            print(f"  {s}: {d[s]:.4f} -> {f[s]:.4f} ms", flush=True)
    for n in sizes(args.schedule):
        cands = schedule_candidates(n, folding=(3, 4, 5, 6)) + [(4, 5, 3), (4, 6, 3), (5, 4, 3), (4, 3, 3)]
        e = tune_schedule(n, proofs=args.proofs, cache_path=args.cache, candidates=cands)
        print(f"schedule n={n}: best {e['best']}", flush=True)


if __name__ == "__main__":
    _main()
