"""ctypes view of the CPU oracle (oracle/libck_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libck_oracle.so"
REF_SHIM_PATH = HERE / "_ref" / "libref_shim.so"

BGH25, DRRT18 = 1, 2
DERIVED = 0xFFFFFFFF
MODES = {"bgh25": BGH25, "drrt18": DRRT18}


class Key(C.Structure):
    _fields_ = [("w", C.c_uint32 * 44)]

    def words(self):
        return list(self.w)


class Params(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("mode", "k", "pnum", "buckets", "kick", "stash",
                                            "ell_i", "input_mode", "half_bits", "ell_d")] + \
               [(n, C.c_uint64) for n in ("bucket_size", "domain", "mask", "range_mask")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Trial(C.Structure):
    _fields_ = [("key", Key), ("walk", C.c_uint64), ("randomness", C.c_uint64),
                ("shift", C.c_uint64), ("step", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("probes", C.c_uint64), ("evals", C.c_uint64),
                ("stashed", C.c_uint64), ("placed", C.c_uint64)]


def _load():
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is not built; run `make -C oracle`")
    lib = C.CDLL(str(LIB_PATH))
    P, U32, U64 = C.POINTER, C.c_uint32, C.c_uint64
    sig = {
        "ock_bit_width": (C.c_uint, [U64]),
        "ock_original_log2": (C.c_float, [U64]),
        "ock_parameters": (C.c_int, [U32] * 8 + [P(Params)]),
        "ock_sbox": (None, [P(C.c_uint8)]),
        "ock_expand_key": (None, [U64, U64, P(Key)]),
        "ock_expand_words": (None, [P(U32), P(Key)]),
        "ock_aes_bytewise": (None, [P(Key), P(U32), P(U32)]),
        "ock_aes": (None, [P(Key), P(U32), P(U32)]),
        "ock_has_aesni": (C.c_int, []),
        "ock_batch_key": (None, [U64, U64, P(Key)]),
        "ock_splitmix": (U64, [P(U64)]),
        "ock_trial_setup": (None, [P(Key), U32, P(Params), P(Trial)]),
        "ock_bucket_direct": (U32, [P(Trial), P(Params), U32, U32]),
        "ock_build_rounds": (None, [P(Trial), P(Params), P(C.c_uint16)]),
        "ock_bucket_cached": (U32, [P(Trial), P(Params), P(C.c_uint16), U32, U32]),
        "ock_bucket_map": (None, [P(Key), U32, P(Params), P(U32)]),
        "ock_trial_run": (C.c_int, [P(Key), U32, P(Params), P(U32), P(C.c_uint16), P(Stats)]),
        "ock_run_range": (U32, [P(Key), U32, U32, P(Params), C.c_int, C.c_int, P(U32), P(U32), P(Stats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def u32(*vals):
    return (C.c_uint32 * len(vals))(*vals)


def parameters(mode, pnum, ell_i, k=3, buckets=DERIVED, kick=DERIVED, stash=0, input_mode=0) -> Params:
    if isinstance(mode, str):
        mode = MODES[mode]
    p = Params()
    rc = LIB.ock_parameters(mode, pnum, ell_i, k, buckets, kick, stash, input_mode, C.byref(p))
    if rc != 0:
        raise ValueError(f"invalid parameters (code {rc})")
    return p


def batch_key(seed: int, batch: int) -> Key:
    k = Key()
    LIB.ock_batch_key(seed, batch, C.byref(k))
    return k


def expand_key(lo: int, hi: int) -> Key:
    k = Key()
    LIB.ock_expand_key(lo, hi, C.byref(k))
    return k


def expand_words(words) -> Key:
    k = Key()
    LIB.ock_expand_words(u32(*words), C.byref(k))
    return k


def aes(key: Key, block, bytewise=False):
    out = (C.c_uint32 * 4)()
    fn = LIB.ock_aes_bytewise if bytewise else LIB.ock_aes
    fn(C.byref(key), u32(*block), out)
    return list(out)


def sbox():
    out = (C.c_uint8 * 256)()
    LIB.ock_sbox(out)
    return bytes(out)


def splitmix(state: int):
    s = C.c_uint64(state)
    v = LIB.ock_splitmix(C.byref(s))
    return s.value, v


def trial_setup(key: Key, offset: int, p: Params) -> Trial:
    t = Trial()
    LIB.ock_trial_setup(C.byref(key), offset, C.byref(p), C.byref(t))
    return t


def bucket_direct(t: Trial, p: Params, ordinal, which):
    return LIB.ock_bucket_direct(C.byref(t), C.byref(p), ordinal, which)


def build_rounds(t: Trial, p: Params) -> np.ndarray:
    r = np.zeros(8 << p.half_bits, dtype=np.uint16)
    LIB.ock_build_rounds(C.byref(t), C.byref(p), r.ctypes.data_as(C.POINTER(C.c_uint16)))
    return r


def bucket_cached(t: Trial, p: Params, rounds: np.ndarray, ordinal, which):
    return LIB.ock_bucket_cached(C.byref(t), C.byref(p), rounds.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 ordinal, which)


def bucket_map(key: Key, offset: int, p: Params) -> np.ndarray:
    m = np.zeros(p.pnum * p.k, dtype=np.uint32)
    LIB.ock_bucket_map(C.byref(key), offset, C.byref(p), _u32p(m))
    return m


def trial_run(key: Key, offset: int, p: Params, cached=True):
    bins = np.zeros(p.buckets, dtype=np.uint32)
    rounds = np.zeros(8 << p.half_bits, dtype=np.uint16) if (cached and p.half_bits <= 16) else None
    st = Stats()
    failed = LIB.ock_trial_run(C.byref(key), offset, C.byref(p), _u32p(bins),
                               rounds.ctypes.data_as(C.POINTER(C.c_uint16)) if rounds is not None else None,
                               C.byref(st))
    return bool(failed), bins, st


def run_range(key: Key, first: int, count: int, p: Params, threads=None, cached=True, probes=False):
    """Failing offsets (sorted), per-trial probe counts (or None) and totals."""
    threads = threads or min(8, os.cpu_count() or 1)
    failed = np.zeros(max(count, 1), dtype=np.uint32)
    pr = np.zeros(max(count, 1), dtype=np.uint32) if probes else None
    st = Stats()
    n = LIB.ock_run_range(C.byref(key), first, count, C.byref(p), threads, 1 if cached else 0,
                          _u32p(failed), _u32p(pr) if pr is not None else None, C.byref(st))
    return failed[:n].copy(), (pr[:count].copy() if pr is not None else None), st


def ref_shim():
    """The reference's own host functions (oracle/_ref), or None when not built."""
    if not REF_SHIM_PATH.exists():
        return None
    lib = C.CDLL(str(REF_SHIM_PATH))
    P, U32, U64 = C.POINTER, C.c_uint32, C.c_uint64
    lib.ref_parameters.argtypes = [C.c_int, U32, U32, P(U32), P(U32), P(U64), P(U64), P(U32), P(U32)]
    lib.ref_parameters.restype = C.c_int
    lib.ref_original_log2.argtypes = [U64]
    lib.ref_original_log2.restype = C.c_float
    lib.ref_batch_key.argtypes = [U64, U64, P(U32)]
    lib.ref_expand_key.argtypes = [U64, U64, P(U32)]
    lib.ref_host_aes.argtypes = [P(U32), P(U32), P(U32)]
    lib.ref_cpu_aes.argtypes = [P(U32), P(U32), P(U32)]
    lib.ref_bucket.argtypes = [C.c_int, C.c_int, U32, U32, U64, U64, U32, U32, U32]
    lib.ref_bucket.restype = U32
    lib.ref_trial.argtypes = [C.c_int, C.c_int, U32, U32, U64, U64, U32, C.c_int, P(U32)]
    lib.ref_trial.restype = C.c_int
    return lib


# ----------------------------------------------------------- table digest

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def table_digest(bins: np.ndarray, bmap: np.ndarray, k: int) -> int:
    """Order-independent digest of a final table (entry = ordinal | choice << 30):
    the wrapping sum over occupied buckets b of mix64(b*golden ^ H ^ choice << 61),
    H = sum over w < k of mix64(w << 32 | bucket(occupant, w))."""
    with np.errstate(over="ignore"):
        b = np.nonzero(bins != 0xFFFFFFFF)[0].astype(np.uint64)
        e = bins[b.astype(np.int64)].astype(np.uint64)
        ordinal = (e & np.uint64(0x3FFFFFFF)).astype(np.int64)
        choice = e >> np.uint64(30)
        h = np.zeros(len(b), dtype=np.uint64)
        for w in range(k):
            cand = bmap[ordinal * k + w].astype(np.uint64)
            h += _mix64((np.uint64(w) << np.uint64(32)) | cand)
        terms = _mix64((b * _GOLD) ^ h ^ (choice << np.uint64(61)))
        return int(np.sum(terms, dtype=np.uint64))
