"""Kernel-schedule autotuner.

Each GEMM shape (transposes, m, n, k, batch, leading dimensions) is timed once
per process over the kernel variants of pypde_dgemm_batched (LDS-tiled
128x128, 64x64 and 32x32, split-K register-direct tiles of 16/32 rows and columns over 4 or 8
waves); the fastest variant is cached and written into the descriptor that the
fused step replays.  Banded applies (stencil maps, sparse plan right-hand
sides) are tuned the same way over the rows each thread (axis 0) or block
(axis 1) covers.

Timing uses HIP events on the launch stream, `reps` back-to-back launches per
candidate after one warm launch.  Nothing is tuned while a stream is being
captured into a graph: captured launches use the cached choice or the size
heuristic of the library (variant 0).

PYPDE_AUTOTUNE=0 in the environment switches tuning off (heuristics only).

Decisions measured once on the target hardware can be shipped as a database
(pypde_amd/tune_db.json, or the file PYPDE_TUNE_DB names; PYPDE_TUNE_DB=""
ignores it): entries for the current device architecture are loaded on first
use and taken as cached choices, so a fresh process runs the same kernels as
the run that wrote them.  `python bench.py --tune-db PATH` writes one.
"""
import ast
import json
import os

import torch

ROWS_CANDIDATES = (1, 2, 4, 8, 18, 20, 24)
# along rows also the coefficient-resident form (33..36: four rows in flight,
# 4..32 rows per block; 37..40: eight rows in flight, 8..64 rows per block)
ROWS_CANDIDATES_AXIS1 = ROWS_CANDIDATES + (33, 34, 35, 36, 37, 38, 39, 40)
FOLD_CANDIDATES = (1, 2, 3, 4, 5)
# output folds (modes 1, 2) also run split over waves
OUTPUT_FOLD_CANDIDATES = FOLD_CANDIDATES + (6, 7)
GEMM_CANDIDATES = tuple(range(1, 43))

_enabled = os.environ.get("PYPDE_AUTOTUNE", "1") != "0"
_cache = {}
_timings = {}


def rows_candidates(axis):
    """Row hints of the banded / stencil applies worth timing along `axis`."""
    return ROWS_CANDIDATES_AXIS1 if axis == 1 else ROWS_CANDIDATES


def enabled():
    return _enabled


def set_enabled(flag):
    global _enabled
    _enabled = bool(flag)


def cache():
    """{key: chosen variant} of this process."""
    if torch.cuda.is_available():
        _ensure_db()
    return dict(_cache)


def timings():
    """{key: {variant: microseconds per launch}} of this process."""
    return {k: dict(v) for k, v in _timings.items()}


def _capturing():
    try:
        return torch.cuda.is_current_stream_capturing()
    except Exception:
        return False


_delay_bufs = None


def queue_delay():
    """Enqueue about a millisecond of GPU work (an FP64 matrix product) so that
    the launches queued behind it while it runs execute back to back: the events
    around them then time the GPU, not the host's launch rate, which for kernels
    of a few microseconds is the larger of the two."""
    global _delay_bufs
    if _delay_bufs is None:
        a = torch.ones((2560, 2560), dtype=torch.float64, device="cuda") / 2560.0
        _delay_bufs = (a, torch.empty_like(a))
    a, c = _delay_bufs
    torch.matmul(a, a, out=c)


def time_candidates(run, candidates, reps=5, rounds=3):
    """{candidate: us per launch}; run(c) enqueues one launch on the current stream.
    Each candidate's time is the best of `rounds` rounds that each time every
    candidate once, so that a slow moment of the device (clock ramp, another
    queue) does not decide between candidates a few percent apart."""
    cands = list(candidates)
    for c in cands:
        run(c)  # first launch of a kernel loads its code object
    out = {c: float("inf") for c in cands}
    for _ in range(rounds):
        for c in cands:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            queue_delay()
            e0.record()
            for _ in range(reps):
                run(c)
            e1.record()
            e1.synchronize()
            out[c] = min(out[c], e0.elapsed_time(e1) * 1e3 / reps)
    return out


_DB_DEFAULT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_db.json")
_db_loaded = False


def _arch():
    try:
        return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
    except Exception:
        return None


def load_db(path=None):
    """Load the tuning decisions stored for this device architecture (returns
    the number of entries taken)."""
    path = os.environ.get("PYPDE_TUNE_DB", _DB_DEFAULT) if path is None else path
    if not path or not os.path.exists(path):
        return 0
    with open(path) as f:
        db = json.load(f)
    if db.get("arch") != _arch():
        return 0
    n = 0
    for k, v in db.get("entries", {}).items():
        key = ast.literal_eval(k)
        if key not in _cache:
            # JSON stores tuple decisions (form, variant) as lists
            _cache[key] = tuple(v) if isinstance(v, list) else v
            n += 1
    return n


def save_db(path):
    """Write this process's decisions, for the current device architecture."""
    entries = {repr(k): v for k, v in sorted(_cache.items(), key=lambda kv: repr(kv[0]))}
    with open(path, "w") as f:
        json.dump({"arch": _arch(), "entries": entries}, f, indent=0, sort_keys=True)
    return len(entries)


def _ensure_db():
    global _db_loaded
    if not _db_loaded:
        _db_loaded = True
        load_db()


def choose(key, run, candidates, default=0):
    """Cached best candidate for `key`, timing the candidates on first use."""
    _ensure_db()
    v = _cache.get(key)
    if v is not None and v in candidates:
        return v
    if not _enabled or _capturing():
        return default
    t = time_candidates(run, candidates)
    v = min(t, key=t.get)
    _cache[key] = v
    _timings[key] = t
    return v


def gemm_key(d):
    return ("gemm", d.transA, d.transB, d.m, d.n, d.k, d.batch, d.lda, d.ldb, d.ldc,
            d.D is not None and d.beta != 0.0, d.fold, d.csn, d.ska)


def aliased(d):
    """True when the descriptor's output is also one of its inputs."""
    return d.C in (d.A, d.B) or (d.D is not None and d.D == d.C and d.beta != 0.0) or \
        (d.E is not None and d.E == d.C) or (d.A1 is not None and d.C == d.A1) or \
        (d.B1 is not None and d.C == d.B1)


def cached(key, default=0):
    """The stored decision for key (database or this process), else default."""
    _ensure_db()
    return _cache.get(key, default)


def tune_gemm_desc(d, launch, allow_alias=False):
    """Fill d.variant; launch(d) enqueues the GEMM with the descriptor's variant.
    A descriptor whose output is also an input (C aliases A, B or D) is not
    timed unless the caller restores the output afterwards (allow_alias),
    since repeated launches change its result; it keeps the size heuristic."""
    if aliased(d) and not allow_alias:
        _ensure_db()
        d.variant = _cache.get(gemm_key(d), 0)
        return d.variant

    def run(v):
        d.variant = v
        launch(d)
    cands = (OUTPUT_FOLD_CANDIDATES if d.fold in (1, 2) else FOLD_CANDIDATES) if d.fold \
        else GEMM_CANDIDATES
    d.variant = choose(gemm_key(d), run, cands)
    return d.variant
