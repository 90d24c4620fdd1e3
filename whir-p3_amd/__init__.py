"""Python bindings (ctypes) for the MI355X WHIR prover library.

The prover is native: HIP kernels plus a C++ host behind the C ABI declared in
include/whir_gpu.h.  This module only loads that library and wraps its entry
points with the names of the reference crate's API (tcoratger/whir-p3):
``WhirConfig``, ``DomainSeparator``, ``DuplexChallenger``, ``InitialStatement``,
``CommitmentWriter.commit`` and ``Prover.prove``.  There is no fallback: if
libwhirgpu.so is missing, importing the library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwhirgpu.so")
P = 0x7F000001

U32P = ctypes.POINTER(ctypes.c_uint32)


class Params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in (
        "num_variables", "security_level", "pow_bits", "folding_factor_first", "folding_factor",
        "starting_log_inv_rate", "rs_domain_initial_reduction_factor", "soundness_type",
        "sumcheck_strategy", "leaf_hash_width", "digest_elems", "mmcs")]


class Poseidon2Constants(ctypes.Structure):
    _fields_ = [
        ("w16_external_initial", ctypes.c_uint32 * 64),
        ("w16_external_terminal", ctypes.c_uint32 * 64),
        ("w16_internal", ctypes.c_uint32 * 20),
        ("w24_external_initial", ctypes.c_uint32 * 96),
        ("w24_external_terminal", ctypes.c_uint32 * 96),
        ("w24_internal", ctypes.c_uint32 * 23),
    ]


class KeccakChallengerState(ctypes.Structure):
    """whirgpu_keccak_challenger: SerializingChallenger32<F, HashChallenger<u8, Keccak256Hash, 32>>."""
    _fields_ = [("input", ctypes.c_uint8 * 4096), ("input_len", ctypes.c_uint32), ("output", ctypes.c_uint8 * 32),
                ("output_len", ctypes.c_uint32)]


MMCS = {"poseidon2": 0, "keccak": 1}


class ChallengerState(ctypes.Structure):
    _fields_ = [("sponge", ctypes.c_uint32 * 16), ("input", ctypes.c_uint32 * 8), ("output", ctypes.c_uint32 * 8),
                ("input_len", ctypes.c_uint32), ("output_len", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    """whirgpu_tuning: launch shapes of a context's kernels (0 / -1: built-in rule)."""
    _fields_ = [(f, ctypes.c_int32) for f in ("lde_lp2_base", "lde_lp2_ext", "lde_tb_log_base", "lde_tb_log_ext",
                                             "tree_lane_log", "tree_s128_log", "grind_wide_bits",
                                             "grind_lock_bits", "grind_blocks")]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for f, _ in cls._fields_:
            setattr(t, f, int(d[f]))
        return t


class Job(ctypes.Structure):
    """whirgpu_job: one commit + prove handled by a server."""
    _fields_ = [("num_variables", ctypes.c_uint32), ("evals", ctypes.c_void_p), ("evals_on_device", ctypes.c_int),
                ("points", ctypes.c_void_p), ("npoints", ctypes.c_uint32), ("evals_out", ctypes.c_void_p),
                ("proof_out", ctypes.c_void_p), ("proof_cap", ctypes.c_size_t), ("proof_len", ctypes.c_size_t),
                ("status", ctypes.c_int), ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("error", ctypes.c_char * 160)]


PENDING = -1
SOUNDNESS = {"UniqueDecoding": 0, "JohnsonBound": 1, "CapacityBound": 2}

# Function signatures of include/whir_gpu.h: name -> (restype, argtypes)
_VP = ctypes.c_void_p
_SIGS = {
    "whirgpu_last_error": (ctypes.c_char_p, []),
    "whirgpu_poseidon2_from_seed": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(Poseidon2Constants)]),
    "whirgpu_sample_field": (ctypes.c_int, [ctypes.c_uint64, U32P, ctypes.c_size_t]),
    "whirgpu_poseidon2_permute": (ctypes.c_int, [ctypes.POINTER(Poseidon2Constants), ctypes.c_uint32, U32P]),
    "whirgpu_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(Poseidon2Constants),
                                      ctypes.POINTER(_VP)]),
    "whirgpu_destroy": (ctypes.c_int, [_VP]),
    "whirgpu_config": (ctypes.c_int, [_VP, U32P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_domain_separator": (ctypes.c_int, [_VP, U32P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_challenger_new": (ctypes.c_int, [ctypes.POINTER(ChallengerState)]),
    "whirgpu_challenger_observe": (ctypes.c_int, [_VP, ctypes.POINTER(ChallengerState), U32P, ctypes.c_size_t]),
    "whirgpu_challenger_observe_domain_separator": (ctypes.c_int, [_VP, ctypes.POINTER(ChallengerState)]),
    "whirgpu_challenger_sample": (ctypes.c_int, [_VP, ctypes.POINTER(ChallengerState), U32P, ctypes.c_size_t]),
    "whirgpu_statement_create": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.POINTER(_VP)]),
    "whirgpu_statement_evaluate": (ctypes.c_int, [_VP, U32P, U32P]),
    "whirgpu_statement_clone": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_void_p)]),
    "whirgpu_statement_destroy": (ctypes.c_int, [_VP]),
    "whirgpu_proof_create": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
    "whirgpu_proof_destroy": (ctypes.c_int, [_VP]),
    "whirgpu_proof_serialize": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_proof_deserialize": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t,
                                                 ctypes.POINTER(_VP)]),
    "whirgpu_verify": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ChallengerState), ctypes.c_uint32, U32P, U32P]),
    "whirgpu_commit": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ChallengerState), _VP, ctypes.POINTER(_VP)]),
    "whirgpu_commit_keccak": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(KeccakChallengerState), _VP, ctypes.POINTER(_VP)]),
    "whirgpu_prove_keccak": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(KeccakChallengerState), _VP, _VP]),
    "whirgpu_verify_keccak": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(KeccakChallengerState), ctypes.c_uint32, U32P,
                                             U32P]),
    "whirgpu_keccak_challenger_new": (ctypes.c_int, [ctypes.POINTER(KeccakChallengerState)]),
    "whirgpu_keccak_challenger_observe": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState), U32P,
                                                         ctypes.c_size_t]),
    "whirgpu_keccak_challenger_sample": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState), U32P,
                                                        ctypes.c_size_t]),
    "whirgpu_keccak_challenger_observe_domain_separator": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState)]),
    "whirgpu_keccak_f1600": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64)]),
    "whirgpu_keccak_grind": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState), ctypes.c_uint32, U32P]),
    "whirgpu_keccak_challenger_observe_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState),
                                                               ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t]),
    "whirgpu_keccak_challenger_sample_bits": (ctypes.c_int, [_VP, ctypes.POINTER(KeccakChallengerState),
                                                             ctypes.c_uint32, U32P]),
    "whirgpu_prove": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ChallengerState), _VP, _VP]),
    "whirgpu_prover_data_destroy": (ctypes.c_int, [_VP]),
    "whirgpu_prove_group": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                           ctypes.POINTER(ChallengerState)]),
    "whirgpu_statement_evaluate_group": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(_VP), U32P, U32P]),
    "whirgpu_last_timings": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double)]),
    "whirgpu_profile_enable": (ctypes.c_int, [_VP, ctypes.c_int]),
    "whirgpu_profile_records": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_profile_read": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_sumcheck_eq": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, U32P, ctypes.c_uint32, U32P,
                                           ctypes.POINTER(ChallengerState), U32P, U32P]),
    "whirgpu_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(Poseidon2Constants),
                                             ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "whirgpu_server_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Job)]),
    "whirgpu_server_wait": (ctypes.c_int, [_VP, ctypes.POINTER(Job), ctypes.c_int]),
    "whirgpu_server_submit_many": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.POINTER(Job)), ctypes.c_size_t]),
    "whirgpu_server_set_group_max": (ctypes.c_int, [_VP, ctypes.c_int]),
    "whirgpu_server_set_self_check": (ctypes.c_int, [_VP, ctypes.c_int]),
    "whirgpu_server_set_tuning": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(Tuning)]),
    "whirgpu_get_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    "whirgpu_set_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    "whirgpu_device_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "whirgpu_server_stats": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "whirgpu_server_destroy": (ctypes.c_int, [_VP]),
    "whirgpu_prove_batch": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(Poseidon2Constants),
                                           ctypes.POINTER(Job), ctypes.c_size_t, ctypes.c_int]),
    "whirgpu_device_alloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "whirgpu_device_free": (ctypes.c_int, [ctypes.c_int, _VP]),
    "whirgpu_memcpy": (ctypes.c_int, [ctypes.c_int, _VP, _VP, ctypes.c_size_t, ctypes.c_int]),
    "whirgpu_lde_base": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP]),
    "whirgpu_lde_ext": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP]),
    "whirgpu_merkle_commit": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64, ctypes.c_uint32, _VP, U32P]),
    "whirgpu_grind": (ctypes.c_int, [_VP, ctypes.POINTER(ChallengerState), ctypes.c_uint32, U32P]),
    "whirgpu_eq_expand": (ctypes.c_int, [_VP, U32P, ctypes.c_uint32, U32P, _VP]),
    "whirgpu_mle_eval_base": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, U32P, U32P]),
    "whirgpu_mle_eval_ext": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, U32P, U32P]),
    "whirgpu_mle_eval_base_noalloc": (ctypes.c_int, [_VP, ctypes.c_uint32, U32P, _VP, _VP]),
    "whirgpu_sumcheck_coeffs": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, U32P]),
    "whirgpu_sumcheck_fold": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, U32P, U32P]),
}

_lib = None


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (run whir-p3_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class WhirError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise WhirError(f"whirgpu error {rc}: {lib().whirgpu_last_error().decode()}")


def u32p(a):
    return a.ctypes.data_as(U32P)


def params(num_variables, *, security_level=100, pow_bits=16, folding_factor=4, folding_factor_first=None,
           starting_log_inv_rate=1, rs_domain_initial_reduction_factor=3, soundness="CapacityBound",
           svo=True, leaf_hash_width=16, digest_elems=None, mmcs="poseidon2"):
    """ProtocolParameters + num_variables, with the 2^24 benchmark's defaults.  mmcs="keccak" selects the
    Keccak variant (src/whir/mod.rs:236-383), whose domain separator counts 4 digest elements."""
    m = MMCS.get(mmcs, mmcs)
    if digest_elems is None:
        digest_elems = 4 if m == 1 else 8
    return Params(num_variables, security_level, pow_bits,
                  folding_factor if folding_factor_first is None else folding_factor_first, folding_factor,
                  starting_log_inv_rate, rs_domain_initial_reduction_factor,
                  SOUNDNESS.get(soundness, soundness), 1 if svo else 0, leaf_hash_width, digest_elems, m)


def device_name(device=0):
    buf = ctypes.create_string_buffer(256)
    check(lib().whirgpu_device_name(device, buf, 256))
    return buf.value.decode()


def poseidon2_from_seed(seed=1):
    c = Poseidon2Constants()
    check(lib().whirgpu_poseidon2_from_seed(seed, ctypes.byref(c)))
    return c


def sample_field(seed, count):
    """`count` draws of rng.random::<KoalaBear>() from SmallRng::seed_from_u64(seed), Montgomery form."""
    out = np.empty(count, dtype=np.uint32)
    check(lib().whirgpu_sample_field(seed, u32p(out), count))
    return out


class DeviceBuffer:
    """A device allocation owned by Python (inputs kept resident in HBM)."""

    def __init__(self, nbytes, device=0):
        self.device = device
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        check(lib().whirgpu_device_alloc(device, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a, device=0):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, device)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().whirgpu_memcpy(self.device, ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                   a.nbytes, 0))

    def download(self, dtype=np.uint32, count=None):
        dt = np.dtype(dtype)
        count = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(count, dtype=dt)
        check(lib().whirgpu_memcpy(self.device, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                   out.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            lib().whirgpu_device_free(self.device, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class WhirConfig:
    """WhirConfig::new + the device context.  device=-1 gives a host-only context
    (schedule and transcript only)."""

    def __init__(self, prm, perm_seed=1, device=0):
        self.params = prm
        self.perms = poseidon2_from_seed(perm_seed)
        h = ctypes.c_void_p()
        check(lib().whirgpu_create(device, ctypes.byref(prm), ctypes.byref(self.perms), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().whirgpu_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def schedule(self):
        out = (ctypes.c_uint32 * 512)()
        n = ctypes.c_size_t()
        check(lib().whirgpu_config(self.handle, out, 512, ctypes.byref(n)))
        return list(out[:n.value])

    def domain_separator(self):
        out = (ctypes.c_uint32 * 4096)()
        n = ctypes.c_size_t()
        check(lib().whirgpu_domain_separator(self.handle, out, 4096, ctypes.byref(n)))
        return np.array(out[:n.value], dtype=np.uint32)

    def challenger(self, observe_domain_separator=True):
        """The context's challenger (duplex Poseidon2, or Keccak for a Keccak context)."""
        if self.params.mmcs == 1:
            ch = KeccakChallengerState()
            check(lib().whirgpu_keccak_challenger_new(ctypes.byref(ch)))
            if observe_domain_separator:
                check(lib().whirgpu_keccak_challenger_observe_domain_separator(self.handle, ctypes.byref(ch)))
            return ch
        ch = ChallengerState()
        check(lib().whirgpu_challenger_new(ctypes.byref(ch)))
        if observe_domain_separator:
            check(lib().whirgpu_challenger_observe_domain_separator(self.handle, ctypes.byref(ch)))
        return ch

    STAGES = ("lde_base", "lde_ext", "merkle_leaves", "merkle_tree", "grind", "sumcheck", "eq_dot", "combine",
              "svo_fold", "eq_expand")

    def tuning(self):
        t = Tuning()
        check(lib().whirgpu_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def set_tuning(self, t):
        """Applies a Tuning (or a dict of its fields); refused if a launch plan does not exist under it."""
        if isinstance(t, dict):
            t = Tuning.from_dict(t)
        check(lib().whirgpu_set_tuning(self.handle, ctypes.byref(t)))

    def profile(self, on=True):
        check(lib().whirgpu_profile_enable(self.handle, 1 if on else 0))

    def profile_read(self):
        out = (ctypes.c_double * 64)()
        n = ctypes.c_size_t()
        check(lib().whirgpu_profile_read(self.handle, out, 64, ctypes.byref(n)))
        v = list(out[:n.value])
        return {s: {"ms": v[3 * i], "launches": int(v[3 * i + 1]), "work": v[3 * i + 2]}
                for i, s in enumerate(self.STAGES)}

    def profile_records(self):
        """Per launch group, without resetting: [(stage, ms, work)]."""
        n = ctypes.c_size_t()
        lib().whirgpu_profile_records(self.handle, None, 0, ctypes.byref(n))
        out = (ctypes.c_double * max(1, n.value))()
        check(lib().whirgpu_profile_records(self.handle, out, n.value, ctypes.byref(n)))
        v = list(out[:n.value])
        return [(self.STAGES[int(v[i])], v[i + 1], v[i + 2]) for i in range(0, len(v), 3)]

    def timings(self):
        out = (ctypes.c_double * 7)()
        check(lib().whirgpu_last_timings(self.handle, out))
        return list(out)


class InitialStatement:
    """WhirConfig::initial_statement + InitialStatement::evaluate."""

    def __init__(self, cfg, evals=None, device_ptr=None):
        self.cfg = cfg
        h = ctypes.c_void_p()
        if device_ptr is not None:
            check(lib().whirgpu_statement_create(cfg.handle, ctypes.c_void_p(device_ptr), 1, ctypes.byref(h)))
        else:
            self._host = np.ascontiguousarray(evals, dtype=np.uint32)
            check(lib().whirgpu_statement_create(cfg.handle, self._host.ctypes.data_as(ctypes.c_void_p), 0,
                                                 ctypes.byref(h)))
        self.handle = h

    def evaluate(self, point):
        pt = np.ascontiguousarray(point, dtype=np.uint32).reshape(-1)
        out = np.empty(4, dtype=np.uint32)
        check(lib().whirgpu_statement_evaluate(self.handle, u32p(pt), u32p(out)))
        return out

    def clone(self):
        """InitialStatement::clone: the constraints so far (a borrowed device buffer stays borrowed)."""
        h = ctypes.c_void_p()
        check(lib().whirgpu_statement_clone(self.handle, ctypes.byref(h)))
        st = InitialStatement.__new__(InitialStatement)
        st.cfg = self.cfg
        st.handle = h
        if hasattr(self, "_host"):
            st._host = self._host
        return st

    def close(self):
        if self.handle:
            lib().whirgpu_statement_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WhirProof:
    def __init__(self, cfg, handle=None):
        if handle is None:
            handle = ctypes.c_void_p()
            check(lib().whirgpu_proof_create(cfg.handle, ctypes.byref(handle)))
        self.handle = handle

    @classmethod
    def from_bytes(cls, cfg, data):
        """Decodes proof bytes (serde deserialization of WhirProof)."""
        data = bytes(data)
        buf = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
        h = ctypes.c_void_p()
        check(lib().whirgpu_proof_deserialize(cfg.handle, buf, len(data), ctypes.byref(h)))
        return cls(cfg, handle=h)

    def to_bytes(self):
        n = ctypes.c_size_t()
        rc = lib().whirgpu_proof_serialize(self.handle, None, 0, ctypes.byref(n))
        buf = (ctypes.c_uint8 * n.value)()
        check(lib().whirgpu_proof_serialize(self.handle, buf, n.value, ctypes.byref(n)))
        del rc
        return bytes(buf)

    def close(self):
        if self.handle:
            lib().whirgpu_proof_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CommitmentWriter:
    def __init__(self, cfg):
        self.cfg = cfg

    def commit(self, proof, challenger, statement):
        pd = ctypes.c_void_p()
        fn = lib().whirgpu_commit_keccak if isinstance(challenger, KeccakChallengerState) else lib().whirgpu_commit
        check(fn(self.cfg.handle, proof.handle, ctypes.byref(challenger), statement.handle, ctypes.byref(pd)))
        return pd


class Prover:
    def __init__(self, cfg):
        self.cfg = cfg

    def prove(self, proof, challenger, statement, prover_data):
        fn = lib().whirgpu_prove_keccak if isinstance(challenger, KeccakChallengerState) else lib().whirgpu_prove
        check(fn(self.cfg.handle, proof.handle, ctypes.byref(challenger), statement.handle, prover_data))


class Verifier:
    """Verifier::verify (src/whir/verifier/mod.rs:58-200) on the host, after
    CommitmentReader::parse_commitment."""

    def __init__(self, cfg):
        self.cfg = cfg

    def verify(self, proof, challenger, points, evals):
        """Raises WhirError (code 7, the failed check in the message) when the
        proof is rejected; `challenger` is the verifier's, after the domain separator."""
        if isinstance(proof, (bytes, bytearray)):
            proof = WhirProof.from_bytes(self.cfg, proof)
        npts = len(points)
        pts = np.ascontiguousarray(np.concatenate([np.asarray(p, np.uint32).reshape(-1) for p in points])
                                   if npts else np.zeros(4, np.uint32))
        evs = np.ascontiguousarray(np.asarray(evals, np.uint32).reshape(-1) if npts else np.zeros(4, np.uint32))
        fn = lib().whirgpu_verify_keccak if isinstance(challenger, KeccakChallengerState) else lib().whirgpu_verify
        check(fn(self.cfg.handle, proof.handle, ctypes.byref(challenger), npts, u32p(pts), u32p(evs)))


def verify_bytes(cfg, proof_bytes, points, evals):
    """True if the proof bytes verify for {points -> evals} from a fresh challenger."""
    try:
        Verifier(cfg).verify(bytes(proof_bytes), cfg.challenger(), points, evals)
        return True
    except WhirError:
        return False


def prove_once(cfg, poly, points=(), device_ptr=None):
    """Commit + prove as benches/whir.rs does; returns (proof bytes, evaluations)."""
    st = InitialStatement(cfg, poly, device_ptr=device_ptr)
    evals = [st.evaluate(p) for p in points]
    ch = cfg.challenger()
    proof = WhirProof(cfg)
    pd = CommitmentWriter(cfg).commit(proof, ch, st)
    Prover(cfg).prove(proof, ch, st, pd)
    out = proof.to_bytes()
    proof.close()
    st.close()
    return out, evals


def prove_group(cfg, polys, points_per_proof, device_ptrs=None):
    """Commit + prove of several polynomials of cfg's size in one group
    (whirgpu_prove_group), each with its own constraints (the same number for
    all); returns [(proof bytes, evaluations)] in input order."""
    if cfg.params.mmcs != MMCS["poseidon2"]:
        raise WhirError("group proving runs the duplex challenger's transcript on the device: "
                        "a Keccak context proves one proof at a time")
    count = len(polys) if device_ptrs is None else len(device_ptrs)
    sts = []
    for i in range(count):
        if device_ptrs is None:
            sts.append(InitialStatement(cfg, polys[i]))
        else:
            sts.append(InitialStatement(cfg, None, device_ptr=device_ptrs[i]))
    handles = (ctypes.c_void_p * count)(*[s.handle.value for s in sts])
    npts = len(points_per_proof[0]) if count else 0
    evals = [[] for _ in range(count)]
    for k in range(npts):
        pts = np.ascontiguousarray(np.concatenate([np.asarray(points_per_proof[i][k], dtype=np.uint32).reshape(-1)
                                                   for i in range(count)]))
        out = np.zeros(4 * count, np.uint32)
        check(lib().whirgpu_statement_evaluate_group(count, handles, u32p(pts), u32p(out)))
        for i in range(count):
            evals[i].append(out[4 * i: 4 * i + 4].copy())
    proofs = [WhirProof(cfg) for _ in range(count)]
    phandles = (ctypes.c_void_p * count)(*[p.handle.value for p in proofs])
    chs = (ChallengerState * count)()
    for i in range(count):
        chs[i] = cfg.challenger()
    check(lib().whirgpu_prove_group(cfg.handle, count, handles, phandles, chs))
    out = [(p.to_bytes(), ev) for p, ev in zip(proofs, evals)]
    for p in proofs:
        p.close()
    for s in sts:
        s.close()
    return out


class ProofRequest:
    """Host-side holder of one job's inputs and output buffer (kept alive while the job runs)."""

    def __init__(self, poly=None, points=(), device_ptr=None, num_variables=None, proof_cap=1 << 22):
        if device_ptr is None:
            self.poly = np.ascontiguousarray(poly, dtype=np.uint32)
            n = int(self.poly.size).bit_length() - 1
        else:
            self.poly = None
            n = num_variables
        self.n = n
        pts = [np.ascontiguousarray(p, dtype=np.uint32).reshape(-1) for p in points]
        self.points = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros(4, np.uint32))
        self.evals = np.zeros(4 * max(len(pts), 1), np.uint32)
        self.buf = (ctypes.c_uint8 * proof_cap)()
        self.job = Job()
        j = self.job
        j.num_variables = n
        j.evals = device_ptr if device_ptr is not None else self.poly.ctypes.data
        j.evals_on_device = 1 if device_ptr is not None else 0
        j.points = self.points.ctypes.data
        j.npoints = len(pts)
        j.evals_out = self.evals.ctypes.data
        j.proof_out = ctypes.addressof(self.buf)
        j.proof_cap = proof_cap
        j.status = PENDING

    def proof(self):
        return bytes(self.buf[: self.job.proof_len])


class Server:
    """whirgpu_server: continuous proving with admission by device memory."""

    def __init__(self, prm, perm_seed=1, device=0, workers=8, hbm_budget=0, group_max=None, self_check=False,
                 tuning_cache=None):
        self.perms = poseidon2_from_seed(perm_seed)
        self.prm = prm
        h = ctypes.c_void_p()
        check(lib().whirgpu_server_create(device, ctypes.byref(prm), ctypes.byref(self.perms), workers, hbm_budget,
                                          ctypes.byref(h)))
        self.handle = h
        if group_max is not None:
            self.set_group_max(group_max)
        if self_check:
            self.set_self_check(True)
        if tuning_cache is not None:
            for n, t in autotune.cached_tunings_by_size(prm, device, tuning_cache).items():
                self.set_tuning(n, t)

    def submit(self, req):
        check(lib().whirgpu_server_submit(self.handle, ctypes.byref(req.job)))

    def submit_many(self, reqs):
        """Submits all requests in one queue update (same-shape jobs can share a group)."""
        arr = (ctypes.POINTER(Job) * len(reqs))(*[ctypes.pointer(r.job) for r in reqs])
        check(lib().whirgpu_server_submit_many(self.handle, arr, len(reqs)))

    def set_group_max(self, g):
        check(lib().whirgpu_server_set_group_max(self.handle, g))

    def set_tuning(self, num_variables, t):
        """Launch shapes for the contexts workers create for `num_variables`."""
        if isinstance(t, dict):
            t = Tuning.from_dict(t)
        check(lib().whirgpu_server_set_tuning(self.handle, num_variables, ctypes.byref(t)))

    def set_self_check(self, on=True):
        """Verify every proof on the host before it is published."""
        check(lib().whirgpu_server_set_self_check(self.handle, 1 if on else 0))

    def wait(self, req, timeout_ms=-1):
        return lib().whirgpu_server_wait(self.handle, ctypes.byref(req.job), timeout_ms)

    def stats(self):
        q, r, b = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().whirgpu_server_stats(self.handle, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b)))
        return q.value, r.value, b.value

    def close(self):
        if self.handle:
            lib().whirgpu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prove_batch(prm, requests, workers=8, perm_seed=1, device=0):
    """whirgpu_prove_batch over ProofRequest objects; returns the first error code (0 if all succeeded)."""
    perms = poseidon2_from_seed(perm_seed)
    arr = (Job * len(requests))()
    for i, r in enumerate(requests):
        arr[i] = r.job
    rc = lib().whirgpu_prove_batch(device, ctypes.byref(prm), ctypes.byref(perms), arr, len(requests), workers)
    for i, r in enumerate(requests):
        r.job = arr[i]
    return rc


from . import autotune  # noqa: E402  (launch-shape and schedule autotuner; uses the bindings above)
