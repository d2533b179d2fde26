"""Python view of libcuckoo_mc.so (the HIP library) through its C ABI.

This is the product path for tests and bench.py: every call runs the gfx950
kernels.  There is no CPU fallback; loading fails loudly when the library is
not built.  The ABI is declared in include/cuckoo_mc.h.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libcuckoo_mc.so"

CK_OK = 0
BGH25, DRRT18 = 1, 2
DERIVED = 0xFFFFFFFF
TIER_AUTO, TIER_HBM, TIER_LDS, TIER_LANE = 0, 1, 2, 3
PENDING = 1
SCHED_BALANCE, SCHED_FIFO, SCHED_SRPT = 0, 1, 2
LAYOUT_AUTO, LAYOUT_MAP, LAYOUT_DESC, LAYOUT_MAP_XCHG, LAYOUT_DESC_PLAIN, LAYOUT_RESIDUAL = 0, 1, 2, 3, 4, 5
MODES = {"bgh25": BGH25, "drrt18": DRRT18}


class Params(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("mode", "k", "pnum", "buckets", "kick", "stash",
                                            "ell_i", "input_mode", "half_bits", "ell_d")] + \
               [(n, C.c_uint64) for n in ("bucket_size", "domain", "mask", "range_mask")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("trials", C.c_uint64), ("failed", C.c_uint64), ("probes", C.c_uint64),
                ("tier", C.c_uint32), ("chunk", C.c_uint32), ("build_ms", C.c_double),
                ("chain_ms", C.c_double), ("fused_ms", C.c_double), ("launches", C.c_uint32),
                ("graph", C.c_uint32), ("span_ms", C.c_double)]


class Job(C.Structure):
    _fields_ = [("p", Params), ("seed", C.c_uint64), ("trials", C.c_uint64), ("bt", C.c_uint32),
                ("keep_failures", C.c_uint32)]


class JobResult(C.Structure):
    _fields_ = [("trials", C.c_uint64), ("failed", C.c_uint64), ("probes", C.c_uint64),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class EngineConfig(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("buffers", C.c_uint32), ("max_slots", C.c_uint32),
                ("max_jobs", C.c_uint32), ("policy", C.c_uint32), ("window", C.c_uint32), ("spread", C.c_uint32)]


class EngineStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("passes", "segments", "slots", "jobs_submitted", "jobs_done", "trials",
                                          "failed", "probes")] + \
               [(n, C.c_double) for n in ("build_ms", "chain_ms", "wall_ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TuneEntry(C.Structure):
    _fields_ = [("tier", C.c_uint32), ("layout", C.c_uint32), ("chunk", C.c_uint32), ("pad", C.c_uint32),
                ("trials_per_s", C.c_double)]


PRP_AUTO, PRP_CACHE, PRP_DIRECT = 0, 1, 2
LDS_AUTO, LDS_MAP, LDS_ROUND = 0, 1, 2


class Placement(C.Structure):
    _fields_ = [("tier", C.c_uint32), ("layout", C.c_uint32), ("prp", C.c_uint32), ("lds_variant", C.c_uint32),
                ("budget_bytes", C.c_uint64)]


class Plan(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("tier", "layout", "round_cache", "lds_variant", "trials_per_pass",
                                            "buffers", "windows", "chain_waves", "chain_lds_bytes",
                                            "build_lds_bytes")] + [("workspace_bytes", C.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TrialInputs(C.Structure):
    _fields_ = [("prp_key", C.c_uint32 * 44), ("walk", C.c_uint64), ("randomness", C.c_uint64),
                ("shift", C.c_uint64), ("step", C.c_uint64)]


class CkError(RuntimeError):
    pass


def _load():
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is not built: run `make -C {HERE}` (or __graft_entry__.build()); "
            "there is no CPU fallback for this path")
    lib = C.CDLL(str(LIB_PATH))
    P, U32, U64 = C.POINTER, C.c_uint32, C.c_uint64
    vp = C.c_void_p
    sig = {
        "ck_params_derive": (C.c_int, [U32] * 8 + [P(Params)]),
        "ck_batch_key": (None, [U64, U64, P(U32)]),
        "ck_host_aes": (None, [P(U32), P(U32), P(U32)]),
        "ck_ctx_create": (C.c_int, [C.c_int, P(vp)]),
        "ck_ctx_destroy": (None, [vp]),
        "ck_ctx_set_tier": (C.c_int, [vp, U32]),
        "ck_ctx_set_chunk": (C.c_int, [vp, U32]),
        "ck_ctx_set_window": (C.c_int, [vp, U32]),
        "ck_ctx_set_place_search": (C.c_int, [vp, U32]),
        "ck_ctx_place_report": (C.c_int, [vp, P(C.c_float), U32, P(U32), P(C.c_int32)]),
        "ck_ctx_place_regions": (C.c_int, [vp, P(C.c_int32), U32, P(U32), P(C.c_int32)]),
        "ck_ctx_set_layout": (C.c_int, [vp, U32]),
        "ck_ctx_set_graph": (C.c_int, [vp, U32]),
        "ck_launch_digest": (C.c_int, [vp, P(Params), P(U32), U32, U32, P(U32), P(U32), P(U32), P(U64),
                                       P(Stats)]),
        "ck_aes_self_test": (C.c_int, [vp]),
        "ck_launch": (C.c_int, [vp, P(Params), P(U32), U32, U32, P(U32), P(U32), P(Stats)]),
        "ck_launch_static": (C.c_int, [vp, P(Params), P(U32), U32, U32, P(U32), P(U32), P(Stats)]),
        "ck_launch_debug": (C.c_int, [vp, P(Params), P(U32), U32, U32, P(U32), P(U32), P(U32), P(U32),
                                      P(Stats)]),
        "ck_trial_map": (C.c_int, [vp, P(Params), P(U32), U32, P(U32), P(U64), P(U64), P(U32)]),
        "ck_last_error": (C.c_char_p, []),
        "ck_ctx_tune": (C.c_int, [vp, P(Params), P(U32), U32, P(TuneEntry), U32, P(U32)]),
        "ck_engine_create": (C.c_int, [C.c_int, P(EngineConfig), P(vp)]),
        "ck_engine_destroy": (None, [vp]),
        "ck_engine_submit": (C.c_int, [vp, P(Job), P(U64)]),
        "ck_engine_poll": (C.c_int, [vp, U64, P(JobResult)]),
        "ck_engine_wait": (C.c_int, [vp, U64, P(JobResult)]),
        "ck_engine_failures": (C.c_int, [vp, U64, P(U64), U64, P(U64)]),
        "ck_engine_release": (C.c_int, [vp, U64]),
        "ck_engine_stats_get": (C.c_int, [vp, P(EngineStats)]),
        "ck_engine_arena": (C.c_int, [vp, P(U32)]),
        "ck_trial_inputs_of": (C.c_int, [P(Params), P(U32), U32, P(TrialInputs)]),
        "ck_round_cache": (C.c_int, [P(Params), P(U32), P(C.c_uint16)]),
        "ck_ctx_set_placement": (C.c_int, [vp, P(Placement)]),
        "ck_ctx_plan": (C.c_int, [vp, P(Params), U32, P(Plan)]),
        "ck_plan_describe": (C.c_int, [P(Params), P(Placement), U32, P(Plan)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()

# Every symbol declared in include/cuckoo_mc.h.
EXPORTS = ["ck_params_derive", "ck_batch_key", "ck_host_aes", "ck_ctx_create", "ck_ctx_destroy",
           "ck_ctx_set_tier", "ck_ctx_set_chunk", "ck_ctx_set_window", "ck_ctx_set_layout", "ck_ctx_set_graph",
           "ck_aes_self_test", "ck_launch",
           "ck_launch_debug", "ck_launch_digest", "ck_trial_map", "ck_last_error", "ck_ctx_tune", "ck_engine_create",
           "ck_engine_destroy", "ck_engine_submit", "ck_engine_poll", "ck_engine_wait", "ck_engine_failures",
           "ck_engine_release", "ck_engine_stats_get", "ck_trial_inputs_of", "ck_round_cache",
           "ck_ctx_set_placement", "ck_ctx_plan", "ck_plan_describe", "ck_launch_static",
           "ck_ctx_set_place_search", "ck_ctx_place_report", "ck_ctx_place_regions", "ck_engine_arena"]


def _check(rc):
    if rc != CK_OK:
        raise CkError(f"code {rc}: {LIB.ck_last_error().decode()}")


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def params(mode, pnum, ell_i, k=3, buckets=DERIVED, kick=DERIVED, stash=0, input_mode=0) -> Params:
    if isinstance(mode, str):
        mode = MODES[mode]
    p = Params()
    _check(LIB.ck_params_derive(mode, pnum, ell_i, k, buckets, kick, stash, input_mode, C.byref(p)))
    return p


def batch_key(seed: int, batch: int) -> np.ndarray:
    rk = np.zeros(44, dtype=np.uint32)
    LIB.ck_batch_key(seed, batch, _u32p(rk))
    return rk


def host_aes(rk: np.ndarray, block) -> list:
    out = (C.c_uint32 * 4)()
    LIB.ck_host_aes(_u32p(np.ascontiguousarray(rk, dtype=np.uint32)), (C.c_uint32 * 4)(*block), out)
    return list(out)


def trial_inputs(p: Params, rk: np.ndarray, offset: int) -> TrialInputs:
    """Host derivation of one trial's inputs for the trial API."""
    ti = TrialInputs()
    _check(LIB.ck_trial_inputs_of(C.byref(p), _u32p(np.ascontiguousarray(rk, dtype=np.uint32)), offset,
                                  C.byref(ti)))
    return ti


def round_cache(p: Params, prp_key) -> np.ndarray:
    """The trial's round cache, 8 << half_bits u16 entries."""
    key = np.ascontiguousarray(np.asarray(prp_key, dtype=np.uint32))
    out = np.zeros(8 << p.half_bits, dtype=np.uint16)
    _check(LIB.ck_round_cache(C.byref(p), _u32p(key), out.ctypes.data_as(C.POINTER(C.c_uint16))))
    return out


def plan_describe(p: Params, count: int, tier=TIER_AUTO, layout=LAYOUT_AUTO, prp=PRP_AUTO, lds_variant=LDS_AUTO,
                  budget_bytes=0) -> Plan:
    """The plan of a launch without a device."""
    out = Plan()
    want = Placement(tier, layout, prp, lds_variant, budget_bytes)
    _check(LIB.ck_plan_describe(C.byref(p), C.byref(want), count, C.byref(out)))
    return out


class Context:
    """One device context (one host thread), as gpu_worker holds per GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(LIB.ck_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            LIB.ck_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tier(self, tier: int):
        _check(LIB.ck_ctx_set_tier(self.h, tier))

    def set_chunk(self, trials: int):
        _check(LIB.ck_ctx_set_chunk(self.h, trials))

    def set_window(self, items: int):
        """Residual-layout item window (0 = automatic)."""
        _check(LIB.ck_ctx_set_window(self.h, items))

    def set_layout(self, layout: int):
        _check(LIB.ck_ctx_set_layout(self.h, layout))

    def set_place_search(self, candidates: int):
        """Table placement on (candidates of the fallback search, 2..64) or off (0)."""
        _check(LIB.ck_ctx_set_place_search(self.h, candidates))

    def place_report(self):
        """The last placement search: (walk ms per candidate, kept indices)."""
        ms = (C.c_float * 256)()
        n = C.c_uint32()
        kept = (C.c_int32 * 2)()
        _check(LIB.ck_ctx_place_report(self.h, ms, 256, C.byref(n), kept))
        return [float(x) for x in ms[:min(n.value, 256)]], (kept[0], kept[1])

    def place_regions(self):
        """The last region placement: (region of each piece, pieces of buffer 0,
        pieces of buffer 1); ([], (-1,)*3, (-1,)*3) when none."""
        reg = (C.c_int32 * 256)()
        n = C.c_uint32()
        used = (C.c_int32 * 6)()
        _check(LIB.ck_ctx_place_regions(self.h, reg, 256, C.byref(n), used))
        return [int(x) for x in reg[:min(n.value, 256)]], tuple(used[0:3]), tuple(used[3:6])

    def set_graph(self, on: bool):
        _check(LIB.ck_ctx_set_graph(self.h, 1 if on else 0))

    def set_placement(self, tier=TIER_AUTO, layout=LAYOUT_AUTO, prp=PRP_AUTO, lds_variant=LDS_AUTO,
                      budget_bytes=0):
        want = Placement(tier, layout, prp, lds_variant, budget_bytes)
        _check(LIB.ck_ctx_set_placement(self.h, C.byref(want)))

    def plan(self, p: Params, count: int) -> Plan:
        out = Plan()
        _check(LIB.ck_ctx_plan(self.h, C.byref(p), count, C.byref(out)))
        return out

    def self_test(self):
        _check(LIB.ck_aes_self_test(self.h))

    def tune(self, p: Params, rk: np.ndarray, sample: int):
        """Times every applicable placement on trials [0, sample) and keeps the
        fastest; returns [{tier, layout, chunk, trials_per_s}] in run order."""
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        ents = (TuneEntry * 32)()
        n = C.c_uint32()
        _check(LIB.ck_ctx_tune(self.h, C.byref(p), _u32p(rk), sample, ents, 32, C.byref(n)))
        return [{"tier": e.tier, "layout": e.layout, "chunk": e.chunk, "trials_per_s": e.trials_per_s}
                for e in ents[:min(n.value, 32)]]

    def launch(self, p: Params, rk: np.ndarray, first: int, count: int):
        """Failing offsets (sorted) of trials [first, first+count) and stats."""
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        out = np.zeros(max(count, 1), dtype=np.uint32)
        n = C.c_uint32()
        st = Stats()
        _check(LIB.ck_launch(self.h, C.byref(p), _u32p(rk), first, count, _u32p(out), C.byref(n), C.byref(st)))
        return out[:n.value].copy(), st

    def launch_static(self, p: Params, rk: np.ndarray, first: int, count: int):
        """ck_launch on the library's static device array (no allocation)."""
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        out = np.zeros(max(count, 1), dtype=np.uint32)
        n = C.c_uint32()
        st = Stats()
        _check(LIB.ck_launch_static(self.h, C.byref(p), _u32p(rk), first, count, _u32p(out), C.byref(n),
                                    C.byref(st)))
        return out[:n.value].copy(), st

    def launch_debug(self, p: Params, rk: np.ndarray, first: int, count: int, tables=False):
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        out = np.zeros(max(count, 1), dtype=np.uint32)
        probes = np.zeros(max(count, 1), dtype=np.uint32)
        tab = np.zeros(count * p.buckets, dtype=np.uint32) if tables else None
        n = C.c_uint32()
        st = Stats()
        _check(LIB.ck_launch_debug(self.h, C.byref(p), _u32p(rk), first, count, _u32p(out), C.byref(n),
                                   _u32p(probes), _u32p(tab) if tab is not None else None, C.byref(st)))
        return (out[:n.value].copy(), probes[:count].copy(),
                tab.reshape(count, p.buckets) if tab is not None else None, st)

    def launch_digest(self, p: Params, rk: np.ndarray, first: int, count: int):
        """Failing offsets, per-trial probe counts and per-trial table digests."""
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        out = np.zeros(max(count, 1), dtype=np.uint32)
        probes = np.zeros(max(count, 1), dtype=np.uint32)
        dig = np.zeros(max(count, 1), dtype=np.uint64)
        n = C.c_uint32()
        st = Stats()
        _check(LIB.ck_launch_digest(self.h, C.byref(p), _u32p(rk), first, count, _u32p(out), C.byref(n),
                                    _u32p(probes), dig.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(st)))
        return out[:n.value].copy(), probes[:count].copy(), dig[:count].copy(), st

    def trial_map(self, p: Params, rk: np.ndarray, offset: int):
        rk = np.ascontiguousarray(rk, dtype=np.uint32)
        key = np.zeros(44, dtype=np.uint32)
        walk, rnd = C.c_uint64(), C.c_uint64()
        m = np.zeros(p.pnum * p.k, dtype=np.uint32)
        _check(LIB.ck_trial_map(self.h, C.byref(p), _u32p(rk), offset, _u32p(key), C.byref(walk),
                                C.byref(rnd), _u32p(m)))
        return key, walk.value, rnd.value, m


class Engine:
    """Batched / served sweeps: jobs of any shape share device passes; results
    are the same trial-for-trial as Context.launch over the same ranges."""

    def __init__(self, device: int = 0, arena_bytes: int = 0, buffers: int = 0, max_slots: int = 0,
                 max_jobs: int = 0, policy: int = SCHED_SRPT, window: int = 0, spread: bool = False):
        cfg = EngineConfig(arena_bytes, buffers, max_slots, max_jobs, policy, window, int(spread))
        h = C.c_void_p()
        _check(LIB.ck_engine_create(device, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            LIB.ck_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, p: Params, seed: int, trials: int, bt: int = 20, keep_failures: bool = False) -> int:
        job = Job(p, seed, trials, bt, 1 if keep_failures else 0)
        jid = C.c_uint64()
        _check(LIB.ck_engine_submit(self.h, C.byref(job), C.byref(jid)))
        return jid.value

    def poll(self, job_id: int):
        """The JobResult when the job has finished, else None."""
        r = JobResult()
        rc = LIB.ck_engine_poll(self.h, job_id, C.byref(r))
        if rc == PENDING:
            return None
        _check(rc)
        return r

    def wait(self, job_id: int) -> JobResult:
        r = JobResult()
        _check(LIB.ck_engine_wait(self.h, job_id, C.byref(r)))
        return r

    def failures(self, job_id: int) -> np.ndarray:
        n = C.c_uint64()
        _check(LIB.ck_engine_failures(self.h, job_id, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(LIB.ck_engine_failures(self.h, job_id, out.ctypes.data_as(C.POINTER(C.c_uint64)), n.value,
                                      C.byref(n)))
        return out[:n.value].copy()

    def release(self, job_id: int):
        _check(LIB.ck_engine_release(self.h, job_id))

    def arena_regions(self):
        """Pieces of the arena per memory region (0, 1, 2, unclear); all zero
        when the arena is one allocation."""
        counts = (C.c_uint32 * 4)()
        _check(LIB.ck_engine_arena(self.h, counts))
        return tuple(counts)

    def stats(self) -> EngineStats:
        s = EngineStats()
        _check(LIB.ck_engine_stats_get(self.h, C.byref(s)))
        return s
