"""Run-time autotuning of the M2L contraction's launch shape, with an on-disk
cache (the analogue of sumpy's ``code_cache``, ``sumpy/__init__.py:69-70``,
``tools.py:474-514``, for tuning results instead of generated code).

The class-batched M2L contraction (``sumpyhip_m2l_apply``) runs one block per
(tile, split) item.  Two shapes are free choices: the tile width (64 or 32
target boxes per tile: narrower tiles hold fewer empty columns on sparse
levels, wider ones reuse each gathered class fragment across more columns),
and how many splits a tile's class slots are cut into (any choice gives the
same sums up to the order of the split additions).  The default (64 columns;
:func:`sumpy_amd.e2e.choose_nsplit`: about 4 items per CU, at least 12 slots
per split) suits dense trees; sparse trees and other orders may prefer
others.  :class:`M2LSplitTuner` times the contraction of a pass under every
tile width and split policy and keeps the fastest, per

    (library digest, kernel, expansion order, precision, GPU architecture,
     schedule digest)

so that a second run on the same kind of tree reads the choice from disk.
The cache lives in ``$SUMPY_AMD_CACHE_DIR`` (default ``~/.cache/sumpy_amd``)
as one JSON file; a missing or unreadable file means "not tuned yet".
"""

from __future__ import annotations

import hashlib
import json
import os

#: (items per CU, minimum class slots per split) candidates
SPLIT_POLICIES = [(f, m) for f in (1, 2, 4, 8) for m in (6, 12, 24)]
DEFAULT_POLICY = (4, 12)
_FILE = "autotune.json"


def cache_dir() -> str:
    return os.environ.get("SUMPY_AMD_CACHE_DIR",
                          os.path.join(os.path.expanduser("~"), ".cache", "sumpy_amd"))


class TuneCache:
    """A JSON dictionary on disk, read once and rewritten on every store."""

    def __init__(self, directory=None):
        self.path = os.path.join(directory or cache_dir(), _FILE)
        self._data = None

    def _load(self):
        if self._data is None:
            try:
                with open(self.path) as f:
                    self._data = json.load(f)
            except (OSError, ValueError):
                self._data = {}
        return self._data

    def get(self, key):
        return self._load().get(key)

    def put(self, key, value):
        data = self._load()
        data[key] = value
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + f".{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # read-only home: the choice is kept for this process only


_CODE_VERSION = None


def code_version() -> str:
    """Digest of the built HIP library (``libsumpyhip.so``): a rebuilt or
    changed contraction kernel gets new keys, so choices timed on old code are
    never reused (the role of ``KERNEL_VERSION`` in the reference's cache
    keys, ``sumpy/tools.py:483-488``)."""
    global _CODE_VERSION
    if _CODE_VERSION is None:
        from sumpy_amd import _capi
        h = hashlib.sha1()
        with open(_capi.LIB_PATH, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        _CODE_VERSION = h.hexdigest()[:12]
    return _CODE_VERSION


def make_key(kernel, order, precision, arch, digest, version=None):
    version = code_version() if version is None else version
    return f"m2l_apply|{version}|{kernel}|{order}|{precision}|{arch}|{digest}"


def make_shape_key(kernel, order, precision, arch, digest, version=None):
    """The key of a (tile width, split policy) choice (:meth:`M2LSplitTuner.choose`)."""
    return make_key(kernel, order, precision, arch, digest, version) + "|tile_width"


def schedule_digest(entries):
    """Digest of the work shape of a batch of levels: tiles, slots and pairs
    per level, rounded to two significant digits so that trees of the same
    kind and size share an entry."""
    def r2(x):
        x = int(x)
        if x < 100:
            return x
        s = len(str(x)) - 2
        return round(x, -s)
    shape = sorted((r2(e["sched"]["ntiles"]), r2(e["sched"]["nslots"]),
                    r2(e["sched"]["npairs"])) for e in entries)
    return hashlib.sha1(repr(shape).encode()).hexdigest()[:16]


def device_arch(device) -> str:
    import torch
    if torch.device(device).type != "cuda":
        return "cpu"
    props = torch.cuda.get_device_properties(device)
    return getattr(props, "gcnArchName", props.name).split(":")[0]


class M2LSplitTuner:
    """Chooses the split policy of :func:`sumpy_amd.e2e.m2l_apply_levels`
    for a batch of levels by timing it (HIP events on the current stream),
    memoised in memory and on disk."""

    def __init__(self, cache=None, repeats=3):
        self.cache = cache if cache is not None else TuneCache()
        self.repeats = repeats
        self._memo = {}

    def policy(self, actx, entries, kernel, order, precision):
        import torch

        from sumpy_amd.e2e import m2l_apply_levels
        entries = [e for e in entries if e is not None]
        if not entries:
            return DEFAULT_POLICY
        key = make_key(kernel, order, precision, device_arch(actx.device),
                       schedule_digest(entries))
        got = self._memo.get(key)
        if got is None:
            got = self.cache.get(key)
            if got is not None:
                got = tuple(got["policy"])
        if got is not None:
            self._memo[key] = got
            return got
        if torch.cuda.is_current_stream_capturing():
            return DEFAULT_POLICY  # never time inside a graph capture
        stream = torch.cuda.current_stream(actx.device)
        times = {}
        for pol in SPLIT_POLICIES:
            m2l_apply_levels(actx, entries, policy=pol)  # warm-up (and LDS limits)
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record(stream)
            for _ in range(self.repeats):
                m2l_apply_levels(actx, entries, policy=pol)
            stop.record(stream)
            stop.synchronize()
            times[pol] = start.elapsed_time(stop)/self.repeats
        best = min(times, key=times.get)
        self._memo[key] = best
        self.cache.put(key, {"policy": list(best), "ms": round(times[best], 4),
                             "default_ms": round(times[DEFAULT_POLICY], 4)})
        return best

    def _time(self, actx, entries, pol):
        import torch

        from sumpy_amd.e2e import m2l_apply_levels
        stream = torch.cuda.current_stream(actx.device)
        m2l_apply_levels(actx, entries, policy=pol)  # warm-up (and LDS limits)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(self.repeats):
            m2l_apply_levels(actx, entries, policy=pol)
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop)/self.repeats

    def choose(self, actx, make_entries, kernel, order, precision):
        """(tile width, split policy) for a batch of levels: *make_entries*
        (tile width) gives the batch's level entries for that width; every
        width of :data:`sumpy_amd.e2e.TILE_WIDTHS` is timed under every split
        policy.  Keyed by the default (64-column) schedule's digest."""
        import torch

        from sumpy_amd.e2e import TILE, TILE_WIDTHS
        base = [e for e in make_entries(TILE) if e is not None]
        if not base:
            return TILE, DEFAULT_POLICY
        key = make_shape_key(kernel, order, precision, device_arch(actx.device),
                             schedule_digest(base))
        got = self._memo.get(key)
        if got is None:
            stored = self.cache.get(key)
            if stored is not None:
                got = (int(stored["tile_width"]), tuple(stored["policy"]))
        if got is not None:
            self._memo[key] = got
            return got
        if torch.cuda.is_current_stream_capturing():
            return TILE, DEFAULT_POLICY  # never time inside a graph capture
        times = {}
        for tw in TILE_WIDTHS:
            ents = base if tw == TILE else [e for e in make_entries(tw) if e is not None]
            for pol in SPLIT_POLICIES:
                times[tw, pol] = self._time(actx, ents, pol)
        best = min(times, key=times.get)
        self._memo[key] = best
        by_width = {str(tw): round(min(t for (w, _), t in times.items() if w == tw), 4)
                    for tw in TILE_WIDTHS}
        self.cache.put(key, {"tile_width": best[0], "policy": list(best[1]),
                             "ms": round(times[best], 4),
                             "default_ms": round(times[TILE, DEFAULT_POLICY], 4),
                             "best_ms_by_tile_width": by_width})
        return best


__all__ = ["DEFAULT_POLICY", "M2LSplitTuner", "SPLIT_POLICIES", "TuneCache", "cache_dir",
           "code_version", "make_key", "make_shape_key", "schedule_digest"]
