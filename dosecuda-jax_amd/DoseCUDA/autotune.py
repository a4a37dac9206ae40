"""Autotuner for the batched schedules (dh_batch_tune).

What is tuned, per batch shape:

* ``ray_shape``: the ray-march's voxel patch per wave (tile shape) -- 4x4x4
  (one brick), 8x4x2, 16x4x1 (four bricks of a row, one z-layer), 16x2x2,
  16x1x4 or 8x2x4.
  The patch decides how many cache lines each of the march's gathers touches
  and how often its lanes reuse their last cell together.  Results do not
  depend on it.
* ``ray_blocks`` / ``fwd_blocks``: how bricks reach the ray-march and forward
  kernels -- one wave per brick through the hardware dispatcher (0) or a
  persistent grid of n 256-thread blocks striding over the bricks of all
  beams.  Results do not depend on it.
* ``adj_cs``, ``adj_wg_target``, ``adj_max_slabs``, ``adj_min_bricks``: the
  adjoint's spot-to-workgroup schedule -- spots per workgroup (32, 64, or 128
  packed) and how the bricks of a beam are cut into slabs (about wg_target
  workgroups per beam, at most max_slabs slabs, at least min_bricks bricks
  per slab).  Only
  the order of partial sums changes.

Search: coordinate descent, one parameter at a time in the order above.  A
candidate is timed on the phase it affects (WET, forward or adjoint): one
warm-up call, then `reps` calls back to back between two events, so the GPU
stays busy and clocked as in a real step.  The fastest candidate is kept
before the next parameter is tried.

Cache: a JSON file (``$DOSEHIP_TUNE_CACHE``, else
``~/.cache/dosecuda_hip/tune.json``) keyed by the device architecture and a
bucketed signature of the batch (beams, bricks, spots, spots per layer,
cutoff, WET mode).  Reads and writes are serialised by a process-wide lock and
the file is replaced atomically.
"""

import ctypes
import json
import math
import os
import tempfile
import threading

import torch

from . import _native as N

_LOCK = threading.Lock()

RAY_SHAPES = (0, 1, 2, 3, 4, 5)
RAY_BLOCKS = (0, 2048, 8192)
FWD_BLOCKS = (0, 2048, 8192)
ADJ_CS = (32, 64, 128)
ADJ_WG_TARGET = (16384, 32768, 65536, 131072, 262144)
ADJ_MAX_SLABS = (512, 1024, 2048, 4096)
ADJ_MIN_BRICKS = (16, 32, 64, 128)


def cache_path():
    p = os.environ.get("DOSEHIP_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "dosecuda_hip", "tune.json")


def _bucket(x):
    return int(round(math.log2(max(1.0, float(x))) * 2))   # half-octave buckets


def signature(engines, cutoff_sigma, wet_mode, grad="synthetic"):
    dev = engines[0].device
    arch = torch.cuda.get_device_properties(dev).gcnArchName.split(":")[0]
    beams = [b for e in engines for b in e.beams]
    bricks = sum(((e.grid.nx + 3) // 4) * ((e.grid.ny + 3) // 4) * ((e.grid.nz + 3) // 4) * len(e.beams)
                 for e in engines)
    spots = sum(b.spots.n_spots for b in beams)
    layers = sum(b.spots.n_layers for b in beams if b.spots.n_spots)
    return {"arch": arch, "beams": len(beams), "bricks": _bucket(bricks), "spots": _bucket(spots),
            "spots_per_layer": _bucket(spots / max(layers, 1)), "cull": bool(cutoff_sigma > 0),
            "wet": wet_mode, "grad": grad}


def _key(sig):
    return "|".join("%s=%s" % (k, sig[k]) for k in sorted(sig))


def load_cache(path=None):
    path = path or cache_path()
    with _LOCK:
        try:
            with open(path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}


def store(sig, tune, timings, path=None):
    path = path or cache_path()
    with _LOCK:
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        data[_key(sig)] = {"signature": sig, "tune": tune, "timings_ms": timings}
        d = os.path.dirname(path) or "."
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".tune-", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def lookup(sig, path=None):
    e = load_cache(path).get(_key(sig))
    return None if e is None else e["tune"]


def _phase_ms(fn, device, reps):
    """Mean milliseconds of fn() over `reps` back-to-back calls (after one warm-up)."""
    fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize(device)
    return e0.elapsed_time(e1) / reps


def tune(engines, cutoff_sigma=None, wet_mode="reach", reps=5, grad_doses=None, log=None, grad_fn=None):
    """Search the schedules for this batch; returns (tune dict, timings).
    The adjoint candidates run on dL/dD = grad_doses, else grad_fn(plan doses)
    (the caller's loss gradient), else a fixed pseudo-random field over the
    voxels with dose.

    The candidate batches write the engines' own WET volumes.  In "reach" mode
    those then hold WET only where the current spots reach (1e30 elsewhere),
    so every engine beam's cached WET is marked stale on return: a later
    engine call recomputes the full WET instead of reading that volume."""
    try:
        return _search(engines, cutoff_sigma, wet_mode, reps, grad_doses, log, grad_fn)
    finally:
        if wet_mode == "reach":
            for e in engines:
                for b in e.beams:
                    b.wet_valid = False


def _search(engines, cutoff_sigma, wet_mode, reps, grad_doses, log, grad_fn):
    from .batch import BatchedDose, make_tune
    cutoff = float(engines[0].cutoff_sigma if cutoff_sigma is None else cutoff_sigma)
    say = log or (lambda m: None)
    best = {"adj_cs": 0, "adj_wg_target": 0, "adj_min_bricks": 0, "adj_max_slabs": 0, "fwd_blocks": 0,
            "ray_blocks": 0, "ray_shape": 0}
    timings = {}
    dev = engines[0].device
    bd = BatchedDose(engines, cutoff_sigma=cutoff, tune=make_tune(**best), wet_mode=wet_mode)
    ray_flags = N.DH_BATCH_WET_REACH if wet_mode == "reach" else 0

    def ray():
        N.check(bd.lib.dh_batch_wet(ctypes.byref(bd.batch), ray_flags, bd._stream()), "dh_batch_wet")

    def fwd():
        N.check(bd.lib.dh_batch_forward(ctypes.byref(bd.batch), 0, bd._stream()), "dh_batch_forward")

    for field, cands, fn in (("ray_shape", RAY_SHAPES, ray), ("ray_blocks", RAY_BLOCKS, ray),
                             ("fwd_blocks", FWD_BLOCKS, fwd)):
        res = {}
        for c in cands:
            setattr(bd.batch, field, c)
            res[c] = _phase_ms(fn, dev, reps)
        best[field] = min(res, key=res.get)
        setattr(bd.batch, field, best[field])
        timings[field] = {str(c): v for c, v in res.items()}
        say("%s: %s -> %d" % (field, {c: round(v, 3) for c, v in res.items()}, best[field]))
    # the adjoint needs dL/dD: a fixed pseudo-random field over the voxels with dose
    ray()
    fwd()
    if grad_doses is None and grad_fn is not None:
        grad_doses = [g.detach().clone() for g in grad_fn([d.clone() for d in bd.doses])]
    if grad_doses is None:
        gen = torch.Generator(device=dev)
        gen.manual_seed(1234)
        grad_doses = [torch.randn(d.shape, generator=gen, device=dev) * (d > 0) for d in bd.doses]
    del bd

    def adj_time(t):
        b2 = BatchedDose(engines, cutoff_sigma=cutoff, tune=make_tune(**t), wet_mode=wet_mode)
        N.check(b2.lib.dh_batch_wet(ctypes.byref(b2.batch), ray_flags, b2._stream()), "dh_batch_wet")
        for g, gd in zip(b2.grad_doses, grad_doses):
            g.copy_(gd)

        # a fresh batch: the first (warm-up) call prepares the spot records, the timed ones reuse them
        def run():
            flags = N.DH_BATCH_REUSE_PREP if b2.batch.prepared else 0
            N.check(b2.lib.dh_batch_adjoint(ctypes.byref(b2.batch), flags, b2._stream()), "dh_batch_adjoint")
        return _phase_ms(run, dev, reps)

    for field, cands in (("adj_cs", ADJ_CS), ("adj_wg_target", ADJ_WG_TARGET), ("adj_max_slabs", ADJ_MAX_SLABS),
                         ("adj_min_bricks", ADJ_MIN_BRICKS)):
        res = {}
        for c in cands:
            t = dict(best)
            t[field] = c
            res[c] = adj_time(t)
        best[field] = min(res, key=res.get)
        timings[field] = {str(c): v for c, v in res.items()}
        say("%s: %s -> %d" % (field, {c: round(v, 3) for c, v in res.items()}, best[field]))
    return best, timings


def get_tune(engines, cutoff_sigma=None, wet_mode="reach", path=None, log=None, retune=False, grad_fn=None,
             grad_key="synthetic"):
    """Cached schedule for this batch shape, tuning it on a cache miss.
    grad_fn: the caller's dL/dD as a function of the plan doses (adjoint
    candidates are timed on it); grad_key names it in the cache key."""
    from .batch import make_tune
    cutoff = float(engines[0].cutoff_sigma if cutoff_sigma is None else cutoff_sigma)
    if grad_fn is not None and grad_key == "synthetic":
        grad_key = "given"
    sig = signature(engines, cutoff, wet_mode, grad_key)
    t = None if retune else lookup(sig, path)
    if t is None:
        t, timings = tune(engines, cutoff, wet_mode, log=log, grad_fn=grad_fn)
        store(sig, t, timings, path)
    return make_tune(**t), t
