"""Shared test helpers: the oracle (CPU checker) and the product bindings."""
import ctypes
import importlib.util
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = 0x7F000001
R_MOD_P = (1 << 32) % P
RINV = pow(R_MOD_P, P - 2, P)

U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)


def to_monty(x):
    a = np.asarray(x, dtype=np.uint64)
    return ((a * np.uint64(R_MOD_P)) % np.uint64(P)).astype(np.uint32)


def from_monty(m):
    a = np.asarray(m, dtype=np.uint64)
    return ((a * np.uint64(RINV)) % np.uint64(P)).astype(np.uint32)


def u32p(a):
    return a.ctypes.data_as(U32P)


_oracle = None


def oracle():
    """The CPU oracle (oracle/liboracle.so), built on first use if absent."""
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        sz = ctypes.c_size_t
        sigs = {
            "orc_config": (ctypes.c_int, [ip, ip, ctypes.c_int]),
            "orc_domain_separator": (ctypes.c_int, [ip, U32P, ctypes.c_int]),
            "orc_sample_field": (None, [ctypes.c_uint64, U32P, sz]),
            "orc_poseidon2_constants": (None, [ctypes.c_uint64, ctypes.c_int, U32P, U32P, U32P]),
            "orc_poseidon2_permute": (None, [ctypes.c_uint64, ctypes.c_int, U32P]),
            "orc_ef_mul": (None, [U32P, U32P, U32P]),
            "orc_ef_inv": (None, [U32P, U32P]),
            "orc_two_adic_generator": (ctypes.c_uint32, [ctypes.c_int]),
            "orc_dft": (None, [U32P, sz, ctypes.c_int]),
            "orc_dft_batch": (None, [U32P, sz, sz]),
            "orc_sponge": (None, [ctypes.c_uint64, ctypes.c_int, U32P, sz, U32P]),
            "orc_merkle": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, U32P, sz, sz, U32P, sz, U32P, U32P]),
            "orc_eq_table": (None, [U32P, ctypes.c_int, U32P, U32P]),
            "orc_mle_base": (None, [U32P, ctypes.c_int, U32P, U32P]),
            "orc_mle_ext": (None, [U32P, ctypes.c_int, U32P, U32P]),
            "orc_sumcheck_coeffs": (None, [U32P, U32P, sz, U32P]),
            "orc_fold_ext": (None, [U32P, sz, U32P]),
            "orc_extrapolate_012": (None, [U32P, U32P, U32P, U32P, U32P]),
            "orc_challenger_script": (ctypes.c_int, [ctypes.c_uint64, U32P, U32P, ctypes.c_int, U32P]),
            "orc_prove": (ctypes.c_int, [ip, ctypes.c_uint64, U32P, U32P, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8),
                                         sz, ctypes.POINTER(sz), U32P, ctypes.POINTER(ctypes.c_double)]),
            "orc_verify": (ctypes.c_int, [ip, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8), sz, U32P, U32P,
                                          ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
            "orc_rng_state": (None, [ctypes.c_uint64, U64P]),
            "orc_sumcheck_eq": (ctypes.c_int, [U32P, ctypes.c_int, U32P, U32P, ctypes.c_uint64, U32P, U32P]),
            "orc_xoshiro_outputs": (None, [U64P, U64P, sz]),
            "orc_keccak_f1600": (None, [U64P]),
            "orc_merkle_keccak": (ctypes.c_int, [U32P, sz, sz, U32P, sz, U32P, U32P]),
            "orc_keccak_grind_bytes": (ctypes.c_uint32, [ctypes.POINTER(ctypes.c_uint8), sz, ctypes.c_int, U32P]),
            "orc_keccak256": (None, [ctypes.POINTER(ctypes.c_uint8), sz, ctypes.c_uint8, ctypes.POINTER(ctypes.c_uint8)]),
            "orc_keccak_challenger_script": (ctypes.c_int, [U32P, U32P, ctypes.c_int, U32P]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _oracle = L
    return _oracle


def oracle_params(n, *, security_level=100, pow_bits=16, ff=4, ff0=None, rate=1, rs=3, soundness=2, svo=1,
                  leaf_width=16, digest_elems=None, mmcs=0):
    """mmcs 0: Poseidon2 Merkle tree and duplex challenger; 1: the Keccak tree and byte challenger (whose domain
    separator counts 4 digest elements unless told otherwise)."""
    if digest_elems is None:
        digest_elems = 4 if mmcs == 1 else 8
    return (ctypes.c_int * 12)(n, security_level, pow_bits, ff if ff0 is None else ff0, ff, rate, rs, soundness, svo,
                               leaf_width, digest_elems, mmcs)


def oracle_prove(prm, poly_monty, points_monty, perm_seed=1):
    """Commit + prove on the oracle; returns (proof bytes, evals (npts x 4), [commit_ms, prove_ms, *stage ms]),
    the stages named by ORACLE_SPANS."""
    L = oracle()
    npts = len(points_monty)
    pts = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.uint32).reshape(-1) for p in points_monty])
                               if npts else np.zeros(4, dtype=np.uint32))
    poly = np.ascontiguousarray(poly_monty, dtype=np.uint32)
    cap = 1 << 26
    buf = (ctypes.c_uint8 * cap)()
    ln = ctypes.c_size_t()
    ev = np.zeros(4 * max(npts, 1), dtype=np.uint32)
    ms = (ctypes.c_double * 8)()
    rc = L.orc_prove(prm, perm_seed, u32p(poly), u32p(pts), npts, buf, cap, ctypes.byref(ln), u32p(ev), ms)
    assert rc == 0, f"oracle prove failed: {rc}"
    return bytes(buf[:ln.value]), ev[:4 * npts].reshape(npts, 4), list(ms)


# Names of the oracle's stage times (the reference's tracing spans), in the
# order oracle_prove returns them after commit_ms and prove_ms.
ORACLE_SPANS = ("dft", "commit_matrix", "ood evaluation", "initial sumcheck (svo rounds)", "rounds", "final_round")


def oracle_verify(prm, proof, points_monty, evals_monty, perm_seed=1):
    L = oracle()
    npts = len(points_monty)
    pts = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.uint32).reshape(-1) for p in points_monty])
                               if npts else np.zeros(4, dtype=np.uint32))
    ev = np.ascontiguousarray(np.asarray(evals_monty, dtype=np.uint32).reshape(-1) if npts
                              else np.zeros(4, dtype=np.uint32))
    buf = (ctypes.c_uint8 * len(proof)).from_buffer_copy(proof)
    err = ctypes.create_string_buffer(256)
    rc = L.orc_verify(prm, perm_seed, buf, len(proof), u32p(pts), u32p(ev), npts, err, 256)
    return rc, err.value.decode()


def oracle_sample(seed, count):
    out = np.empty(count, dtype=np.uint32)
    oracle().orc_sample_field(seed, u32p(out), count)
    return out


_product = None


def product():
    """The product bindings (package directory whir-p3_amd, imported as whir_p3_amd)."""
    global _product
    if _product is None:
        if "whir_p3_amd" in sys.modules:
            _product = sys.modules["whir_p3_amd"]
        else:
            spec = importlib.util.spec_from_file_location(
                "whir_p3_amd", os.path.join(ROOT, "whir-p3_amd", "__init__.py"))
            mod = importlib.util.module_from_spec(spec)
            sys.modules["whir_p3_amd"] = mod
            spec.loader.exec_module(mod)
            _product = mod
    return _product
