"""Autotuner of kernel launch parameters, with a persistent cache.

A tuning problem is a key (the kernel and a coarse signature of the workload
and device), a list of candidate parameter sets and a function that enqueues
the work for one candidate.  Every candidate is timed with HIP events on the
launch stream (warmup + median of a few runs); the fastest wins and is stored
in a JSON cache so later runs of the same signature skip the search.

Cache file: $DWBA_AUTOTUNE_CACHE, else ~/.cache/dwba_amd/autotune.json.
Setting DWBA_AUTOTUNE=off makes every lookup miss, and ScanBatch.tune()
then times nothing and keeps the library defaults.

Every parameter tuned here changes only the order of FP64 operations (how
the grid is cut into lanes, how many multipoles share registers, which
contraction kernel runs, which streams the stages run on), never the
quadrature, so tuned and default runs agree to rounding.
"""
from __future__ import annotations

import json
import math
import os
import threading
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

_lock = threading.Lock()


def cache_path() -> str:
    p = os.environ.get("DWBA_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "dwba_amd", "autotune.json")


def enabled() -> bool:
    return os.environ.get("DWBA_AUTOTUNE", "on").lower() not in ("0", "off", "false", "no")


def bucket(x: float) -> int:
    """Power-of-two bucket of a size (signatures stay stable across nearby sizes)."""
    return 0 if x <= 0 else 1 << int(round(math.log2(x)))


def signature(kernel: str, device: str, **fields) -> str:
    parts = [kernel, device] + [f"{k}={fields[k]}" for k in sorted(fields)]
    return "|".join(parts)


@dataclass
class Choice:
    params: Dict[str, int]
    ms: float
    timings: List[dict] = field(default_factory=list)
    cached: bool = False


class Autotuner:
    def __init__(self, path: Optional[str] = None, persist: bool = True):
        self.path = path or cache_path()
        self.persist = persist
        self._table: Dict[str, dict] = {}
        self._loaded = False

    # ------------------------------------------------------------ cache
    def _load(self):
        if self._loaded:
            return
        self._loaded = True
        try:
            with open(self.path) as f:
                data = json.load(f)
            if isinstance(data, dict):
                self._table = data
        except (OSError, ValueError):
            self._table = {}

    def lookup(self, key: str) -> Optional[Choice]:
        if not enabled():
            return None
        with _lock:
            self._load()
            e = self._table.get(key)
        if not e:
            return None
        return Choice(params=dict(e["params"]), ms=float(e["ms"]), timings=list(e.get("timings", [])), cached=True)

    def store(self, key: str, choice: Choice):
        with _lock:
            self._load()
            self._table[key] = {"params": choice.params, "ms": choice.ms, "timings": choice.timings}
            if not self.persist:
                return
            try:
                os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
                tmp = self.path + f".tmp{os.getpid()}"
                with open(tmp, "w") as f:
                    json.dump(self._table, f, indent=1, sort_keys=True)
                os.replace(tmp, self.path)
            except OSError:
                pass          # an unwritable cache only costs a re-tune next time

    # ------------------------------------------------------------ search
    def tune(self, key: str, candidates: Sequence[Dict[str, int]], run: Callable[[Dict[str, int]], None],
             timer: Optional[Callable[[Callable[[], None], int], float]] = None, steps: int = 3,
             warmup: int = 1, margin: float = 0.0) -> Choice:
        """Fastest candidate for `key` (from the cache when present).
        timer(fn, steps) -> ms per call; defaults to HIP events on torch's
        current stream.  margin > 0: the earliest candidate (in the given
        order) within (1 + margin) of the fastest wins, so timing noise does
        not displace a preferred first candidate."""
        hit = self.lookup(key)
        if hit is not None:
            return hit
        if not candidates:
            raise ValueError("autotune: no candidates")
        timer = timer or _event_timer
        timings = []
        best = None
        for cand in candidates:
            for _ in range(warmup):
                run(cand)
            runs = sorted(timer(lambda: run(cand), 1) for _ in range(max(1, steps)))
            ms = runs[len(runs) // 2]
            timings.append({"params": dict(cand), "ms": ms})
            if best is None or ms < best[1]:
                best = (dict(cand), ms)
        if margin > 0:
            for t in timings:
                if t["ms"] <= best[1] * (1.0 + margin):
                    best = (dict(t["params"]), t["ms"])
                    break
        choice = Choice(params=best[0], ms=best[1], timings=timings)
        self.store(key, choice)
        return choice


def _event_timer(fn: Callable[[], None], steps: int) -> float:
    import torch
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


_default: Optional[Autotuner] = None


def default_tuner() -> Autotuner:
    global _default
    if _default is None:
        _default = Autotuner()
    return _default


# ---------------------------------------------------------------------------
# semiseparable contraction: (seg, lb)

SEMISEP_SEGS = (0, 16, 32, 64, 128, 256)
SEMISEP_LBS = (4, 8)


def semisep_candidates(L_max: int = None) -> List[Dict[str, int]]:
    """(seg, lb) of the wavefront-scan kernel, plus the lane-serial kernel
    (one lane per job, family and block of lb multipoles; seg unused), whole
    range per lane or split between two wavefronts (split = the occupancy,
    2 or 3 waves per SIMD, that kernel is compiled for; 4 = the LDS-staged
    split kernel, lb 4 from L_max 4 and lb 8 from L_max 16); order 1: the lane kernel's jobs grouped
    by energy instead of sorted by length alone."""
    stage = [{"seg": 0, "lb": b, "lanes": 1, "split": 4, "order": 1} for b in SEMISEP_LBS
             if L_max is None or (16 if b == 8 else 4) <= L_max <= 63]
    return [{"seg": s, "lb": b} for b in SEMISEP_LBS for s in SEMISEP_SEGS] + [
        {"seg": 0, "lb": b, "lanes": 1} for b in SEMISEP_LBS] + [
        {"seg": 0, "lb": b, "lanes": 1, "split": w} for b in SEMISEP_LBS for w in (2, 3)] + [
        {"seg": 0, "lb": 8, "lanes": 1, "split": w, "order": 1} for w in (0, 2)] + stage


def device_id(dev) -> str:
    """Architecture and compute-unit count of a device: stable across
    processes and profilers (the marketing name can read empty under a
    profiler's interception layer)."""
    import torch
    p = torch.cuda.get_device_properties(dev)
    arch = getattr(p, "gcnArchName", "") or p.name
    return f"{arch.split(':')[0]}x{p.multi_processor_count}"
