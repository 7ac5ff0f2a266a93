"""Tile-shape autotuner for the back-projection kernels.

A configuration is (slice_batch, tile_rows): slices per workgroup and image
rows per 64-column tile.  Every configuration computes the same bits (the
per-pixel arithmetic and angle order do not depend on tiling), so tuning
only changes speed.

Lookup order for a key (arch, kernel, mode, dtype, width bucket, stack
bucket): the in-process table, the user cache (``$ALGOTOM_AMD_TUNE_CACHE`` or
``~/.cache/algotom_amd/tuning.json``), the measured defaults shipped in
``tuned_defaults.json``, then the library's built-in default.  ``tune()``
times candidates with HIP events on the caller's workload and records the
winner in the user cache.
"""

import json
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(_HERE, "tuned_defaults.json")

MODE_NAMES = {0: "f64", 1: "f32seq", 2: "f32"}
CANDIDATE_ROWS = (16, 32, 64)

_lock = threading.Lock()
_table = None


def cache_path():
    return os.environ.get(
        "ALGOTOM_AMD_TUNE_CACHE",
        os.path.join(os.path.expanduser("~"), ".cache", "algotom_amd",
                     "tuning.json"))


def _arch():
    try:
        import torch
        if torch.cuda.is_available():
            props = torch.cuda.get_device_properties(0)
            return getattr(props, "gcnArchName", "gpu").split(":")[0]
    except Exception:  # pragma: no cover
        pass
    return "gfx950"


def width_bucket(ncol):
    """Power-of-two bucket of the detector width (1024, 2048, 4096, ...)."""
    b = 256
    while b < ncol:
        b *= 2
    return b


def stack_bucket(nslice):
    return 1 if nslice == 1 else (2 if nslice == 2 else 4)


def make_key(kernel, mode, dtype, ncol, nslice, arch=None):
    return "%s|%s|%s|%s|C%d|S%d" % (arch or _arch(), kernel,
                                    MODE_NAMES.get(mode, mode), dtype,
                                    width_bucket(ncol), stack_bucket(nslice))


def _load():
    global _table
    if _table is None:
        table = {}
        for path in (DEFAULTS_PATH, cache_path()):
            try:
                with open(path) as f:
                    table.update(json.load(f))
            except (OSError, ValueError):
                pass
        _table = table
    return _table


def compiled_configs(kernel, wide=False):
    """Every (slice_batch, tile_rows) the library compiles for a kernel
    ("stack": ath_back_projection_tiled, "batch":
    ath_back_projection_batched); wide = 16-bit storage."""
    if kernel == "batch":
        return {(1, 16), (1, 32), (1, 64), (4, 16), (4, 32)}
    out = {(sb, rows) for sb in (1, 2, 4) for rows in CANDIDATE_ROWS}
    if wide:
        out |= {(8, 16), (8, 32)}
    return out


def lookup(key, valid=None):
    """(slice_batch, tile_rows) for key, or None.  With ``valid`` (a set of
    configurations), an entry outside it -- a stale or hand-edited cache
    entry the library does not compile -- is dropped from the in-process
    table and None is returned."""
    with _lock:
        table = _load()
        hit = table.get(key)
        if not hit:
            return None
        try:
            cfg = tuple(int(v) for v in hit)
        except (TypeError, ValueError):
            cfg = None
        if cfg is None or len(cfg) != 2 or (valid is not None
                                            and cfg not in valid):
            del table[key]
            return None
    return cfg


def forget(key):
    """Drop key from the in-process table (a configuration that failed)."""
    with _lock:
        _load().pop(key, None)


def record(key, config, persist=True):
    with _lock:
        _load()[key] = list(config)
        if not persist:
            return
        path = cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            try:
                with open(path) as f:
                    saved = json.load(f)
            except (OSError, ValueError):
                saved = {}
            saved[key] = list(config)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(saved, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass  # a read-only home only loses persistence


def candidates(nslice, wide=False):
    """Configurations compiled for the kernels (see backproject.hip)."""
    if nslice == 1:
        sbs = (1,)
    elif nslice == 2:
        sbs = (2, 1)
    else:
        sbs = (4, 2, 8) if wide else (4, 2)
    out = []
    for sb in sbs:
        for rows in CANDIDATE_ROWS:
            if sb == 8 and rows == 64:
                continue
            out.append((sb, rows))
    return out


def time_config(run, reps=2):
    """Mean milliseconds of run() over reps after one warm-up (HIP events on
    the current stream)."""
    import torch
    run()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        run()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune(key, configs, run, persist=True):
    """Times run(config) for every config; records and returns the fastest
    with the table of timings."""
    timings = {}
    for cfg in configs:
        try:
            timings[tuple(cfg)] = time_config(lambda: run(cfg))
        except (ValueError, RuntimeError):
            continue  # configuration not compiled for this mode
    if not timings:
        raise RuntimeError("no tile configuration ran for %s" % key)
    best = min(timings, key=timings.get)
    record(key, best, persist)
    return best, timings


def reset():
    """Forget the in-process table (tests)."""
    global _table
    with _lock:
        _table = None
