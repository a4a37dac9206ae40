"""Batched plans: many (dose grid, plan) pairs computed together.

``BatchedDose([engine_0, engine_1, ...])`` gathers the device state of several
``IMPTDoseEngine`` objects (plans of any grid size, beam count and spot count)
into one batch of the C ABI (dh_batch_setup).  Each phase then runs as a
single launch over the work items of every beam of every plan:

* ``compute_wet()``    ray-march + smoothing of every beam (dh_batch_wet)
* ``forward()``        dose of every plan (dh_batch_forward)
* ``adjoint(grads)``   dL/dweight and dL/d(x, y) of every beam (dh_batch_adjoint)

WET modes: ``"full"`` computes every voxel's WET once and keeps it (positions
may then move freely); ``"reach"`` recomputes WET inside every ``forward()``,
only in the bricks the current spots can reach (dh_batch_wet with
DH_BATCH_WET_REACH), which gives the same dose and gradients bit for bit at a
fraction of the ray-march work.

The reference computes one beam per extension call and one plan per
``computeIMPTPlan`` (plan_impt.py:170-206); per plan the results here are the
same as the engine's (beam doses added in beam order, times n_fractions).

The phase calls only enqueue kernels that read tables prepared at set-up, so
``GraphedStep`` can capture a whole optimisation step (forward, a loss
gradient written with torch element-wise ops, adjoint) in one HIP graph and
replay it; spot weights and positions are read from device tensors at every
replay.

``StreamedBatches`` splits the plans into a few sub-batches whose phases run
concurrently on streams of their own (same results, bit for bit).
"""

import ctypes

import numpy as np
import torch

from . import _native as N
from . import device as D


TUNE_FIELDS = ("adj_cs", "adj_wg_target", "adj_min_bricks", "adj_max_slabs", "fwd_blocks", "ray_blocks",
               "ray_shape")


def make_tune(adj_cs=0, adj_wg_target=0, adj_min_bricks=0, adj_max_slabs=0, fwd_blocks=0, ray_blocks=0,
              ray_shape=0):
    """dh_batch_tune; zero fields take the library defaults (0, 32768, 64, 1024, 0, 0, 0)."""
    return N.dh_batch_tune(int(adj_cs), int(adj_wg_target), int(adj_min_bricks), int(adj_max_slabs),
                           int(fwd_blocks), int(ray_blocks), int(ray_shape))


def tune_dict(t):
    return {f: int(getattr(t, f)) for f in TUNE_FIELDS}


class BatchedDose:
    """Device state of a batch of plans and the three batched phases.

    engines: IMPTDoseEngine objects on one device (their HU -> RLSP volumes,
    beam models and sorted spots are used as they are).
    keep_beam_doses: give every beam a dose volume of its own (``beam_doses``),
    as the reference's ``IMPTDoseGrid.beam_doses``; otherwise beams of
    multi-beam plans write workspace scratch.
    tune: dh_batch_tune, None for the defaults, or "auto" for the autotuner's
    schedule for this batch shape (DoseCUDA.autotune; tuned on a cache miss).
    tune_grad_fn: with tune="auto", the caller's dL/dD as a function of the
    list of plan doses, so the adjoint schedule is tuned on the gradient field
    the caller will use (default: a pseudo-random field where there is dose);
    tune_grad_key names that function in the tuning cache.
    wet_mode: "full" or "reach" (see the module notes).
    """

    def __init__(self, engines, cutoff_sigma=None, tune=None, keep_beam_doses=False, wet_mode="full",
                 tune_grad_fn=None, tune_grad_key="synthetic"):
        if wet_mode not in ("full", "reach"):
            raise ValueError("wet_mode must be 'full' or 'reach'")
        self.wet_mode = wet_mode
        if not engines:
            raise ValueError("BatchedDose needs at least one plan")
        self.lib = N.load()
        self.engines = list(engines)
        self.device = self.engines[0].device
        for e in self.engines:
            if e.device != self.device:
                raise ValueError("all plans of a batch must live on one device")
        self.cutoff_sigma = float(self.engines[0].cutoff_sigma if cutoff_sigma is None else cutoff_sigma)
        if isinstance(tune, str):
            if tune != "auto":
                raise ValueError("tune must be a dh_batch_tune, None or 'auto'")
            from .autotune import get_tune
            tune, _ = get_tune(self.engines, self.cutoff_sigma, wet_mode, grad_fn=tune_grad_fn,
                               grad_key=tune_grad_key)
        self.tune = tune if tune is not None else make_tune()
        dev = self.device
        self.doses, self.grad_doses, self.beam_doses = [], [], []
        self.beam_grads = []   # per beam (grad_mu [S], grad_xy [S, 2]) in sorted spot order
        plans = (N.dh_batch_plan * len(self.engines))()
        n_beams = sum(len(e.beams) for e in self.engines)
        beams = (N.dh_batch_beam * max(n_beams, 1))()
        k = 0
        with torch.cuda.device(dev):
            for p, e in enumerate(self.engines):
                dose = torch.zeros(e.shape, dtype=torch.float32, device=dev)
                gdose = torch.zeros(e.shape, dtype=torch.float32, device=dev)
                self.doses.append(dose)
                self.grad_doses.append(gdose)
                plans[p].grid = e.grid
                plans[p].rlsp_padded = N.ptr(e.rlsp_padded).value
                plans[p].dose = N.ptr(dose).value
                plans[p].grad_dose = N.ptr(gdose).value
                plans[p].beam_start = k
                plans[p].n_beams = len(e.beams)
                for b in e.beams:
                    if b.wet is None:
                        b.wet = torch.zeros(e.shape, dtype=torch.float32, device=dev)
                    S = b.spots.n_spots
                    gmu = torch.zeros(max(S, 1), dtype=torch.float32, device=dev)[:S]
                    gxy = torch.zeros((max(S, 1), 2), dtype=torch.float32, device=dev)[:S]
                    self.beam_grads.append((gmu, gxy))
                    bd = None
                    if keep_beam_doses:
                        bd = torch.zeros(e.shape, dtype=torch.float32, device=dev)
                    self.beam_doses.append(bd)
                    B = beams[k]
                    B.geom = b.geo
                    B.model = b.dmodel.c
                    B.spots = b.spots.c
                    B.wet = N.ptr(b.wet).value
                    B.wet_raw = N.ptr(b.wet_raw).value if b.wet_raw is not None else None
                    B.dose = N.ptr(bd).value if bd is not None else None
                    B.grad_mu = N.ptr(gmu).value if S else None
                    B.grad_xy = N.ptr(gxy).value if S else None
                    k += 1
            self._plans, self._beams = plans, beams
            nbytes = int(self.lib.dh_batch_workspace_size(plans, len(self.engines), beams, n_beams,
                                                          ctypes.byref(self.tune)))
            if nbytes < 0:
                N.check(nbytes, "dh_batch_workspace_size")
            self.ws_bytes = nbytes
            self.ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
            self.batch = N.dh_batch()
            rc = self.lib.dh_batch_setup(ctypes.byref(self.batch), plans, len(self.engines), beams, n_beams,
                                         ctypes.byref(self.tune), self.cutoff_sigma, N.ptr(self.ws), nbytes,
                                         N.stream_handle(dev))
            N.check(rc, "dh_batch_setup")
        self.n_beams = n_beams
        self._wet_done = False

    # -- phases -------------------------------------------------------------------
    def _stream(self):
        return N.stream_handle(self.device)

    def compute_wet(self, force=False):
        """Full WET of every beam (written into each engine beam's ``wet``)."""
        if self._wet_done and not force:
            return
        N.check(self.lib.dh_batch_wet(ctypes.byref(self.batch), 0, self._stream()), "dh_batch_wet")
        self._wet_done = True
        for e in self.engines:
            for b in e.beams:
                b.wet_valid = True

    def _reach_wet(self):
        # prepares the spot / layer records, then WET only where they reach
        N.check(self.lib.dh_batch_wet(ctypes.byref(self.batch), N.DH_BATCH_WET_REACH, self._stream()),
                "dh_batch_wet")
        self._wet_done = False
        for e in self.engines:
            for b in e.beams:
                b.wet_valid = False   # not the full WET any more

    def set_spots(self, mus=None, xys=None):
        """Per-beam weights / positions (lists over all beams of the batch, in
        each beam's spot_list row order); None entries keep the current values."""
        k = 0
        for e in self.engines:
            for b in e.beams:
                mu = None if mus is None else mus[k]
                xy = None if xys is None else xys[k]
                if mu is not None or xy is not None:
                    b.spots.set_weights_positions(mu, xy)
                k += 1

    def forward(self, mus=None, xys=None):
        """Plan doses (list of (z, y, x) tensors, times n_fractions)."""
        self.set_spots(mus, xys)
        if self.wet_mode == "reach":
            self._reach_wet()
            flags = N.DH_BATCH_REUSE_PREP
        else:
            self.compute_wet()
            flags = 0
        N.check(self.lib.dh_batch_forward(ctypes.byref(self.batch), flags, self._stream()), "dh_batch_forward")
        for e, d in zip(self.engines, self.doses):
            if e.n_fractions != 1:
                d.mul_(float(e.n_fractions))
        return self.doses

    def adjoint(self, grad_doses=None, reuse_prep=False, input_order=True):
        """Spot gradients of every beam for dL/dD = grad_doses[p] of plan p
        (None: the batch's own ``grad_doses`` tensors, filled anew before each
        call; plans with n_fractions != 1 scale them in place).
        Returns [(dL/dmu [S], dL/d(x, y) [S, 2])] per beam.  In "reach" mode
        the WET is the one the last forward() computed for its spot positions:
        call adjoint after the forward of the same spots."""
        if self.wet_mode == "full":
            self.compute_wet()
        for p, e in enumerate(self.engines):
            g = self.grad_doses[p]
            if grad_doses is not None:
                g.copy_(grad_doses[p])
            if e.n_fractions != 1:
                g.mul_(float(e.n_fractions))
        flags = N.DH_BATCH_REUSE_PREP if reuse_prep else 0
        N.check(self.lib.dh_batch_adjoint(ctypes.byref(self.batch), flags, self._stream()), "dh_batch_adjoint")
        if not input_order:
            return self.beam_grads
        out = []
        k = 0
        for e in self.engines:
            for b in e.beams:
                gmu, gxy = self.beam_grads[k]
                out.append((b.spots.to_input_order(gmu), b.spots.to_input_order(gxy)))
                k += 1
        return out

    @property
    def smooth_tiled(self):
        """True when this batch smooths WET with the tiled (LDS) kernel; a beam
        whose stencil halo box exceeds 64 KB puts the batch on the gather kernel."""
        return bool(int(self.batch.tab[N.DH_BATCH_TAB_FLAGS]) & N.DH_BATCH_FLAG_SMOOTH_TILE)

    def poison_scratch(self):
        """Fill the workspace scratch (everything after the set-up tables) with
        NaN bit patterns: a phase that read scratch it had not written would
        then show up as NaN results (test aid)."""
        start = int(self.batch.tab[N.DH_BATCH_TAB_HEADER_END])
        self.ws[start:].fill_(255)

    def item_counts(self):
        """Work items of each phase (from the set-up tables)."""
        b = self.batch
        return {"bricks": b.n_bricks, "smooth_blocks": b.n_smooth_blocks, "spots": b.n_spots,
                "layers": b.n_layers, "sum_blocks": b.n_sum_blocks, "partials": b.n_partials,
                "adjoint_wg_32": b.n_adj32, "adjoint_wg_64": b.n_adj64, "adjoint_wg_128": b.n_adj128, "reduce_blocks": b.n_reduce_blocks,
                "workspace_bytes": self.ws_bytes}


def split_contiguous(costs, n):
    """Split items with the given costs into min(n, len(costs)) contiguous,
    non-empty groups of near-equal total cost: a group closes at the first item
    whose running sum reaches its share (j + 1) / n of the total, or when every
    later item is needed to give each remaining group one.  Returns a list of
    index lists."""
    costs = [max(float(c), 0.0) for c in costs]
    m = len(costs)
    n = max(1, min(int(n), m))
    if sum(costs) <= 0.0:
        costs = [1.0] * m
    total = sum(costs)
    groups, cur, run = [], [], 0.0
    for i, c in enumerate(costs):
        cur.append(i)
        run += c
        need = n - len(groups) - 1   # groups still to open after this one
        if need > 0 and (run >= total * (len(groups) + 1) / n or m - i - 1 == need):
            groups.append(cur)
            cur = []
    groups.append(cur)
    return groups


class StreamedBatches:
    """A batch of plans as several ``BatchedDose`` sub-batches whose phases run
    concurrently, each on a stream of its own.

    The plans are split into ``n_streams`` contiguous groups of near-equal work
    (voxels × spots of each plan).  Sub-batch 0 runs on the caller's current
    stream, sub-batch j > 0 on a side stream that first waits for the current
    stream, and the current stream waits for every side stream before a call
    returns: calls order with the caller's other work, and ``GraphedStep``
    captures them as one graph with parallel branches.  Kernels of different
    sub-batches then share the GPU (a ray-march, bound by its texture-path
    gathers, beside a dose kernel, bound by the VALU; each kernel's last waves
    beside the next kernel of another sub-batch).  Each sub-batch computes what
    one ``BatchedDose`` over all the plans computes for its plans, bit for bit.

    grad_fn(p, dose, out): the caller's dL/dD of plan p (index into
    ``engines``) written into out, element-wise device work; ``step`` uses it,
    and with tune="auto" (and no tune_grad_fn) each sub-batch's adjoint
    schedule is tuned on it.  Other keyword arguments go to every
    ``BatchedDose`` (with tune="auto" each sub-batch takes the autotuner's
    schedule for its own shape).  ``doses``, ``grad_doses`` and ``beam_grads``
    list every plan and beam in the order of ``engines``.
    """

    def __init__(self, engines, n_streams=2, grad_fn=None, **kw):
        engines = list(engines)
        if not engines:
            raise ValueError("StreamedBatches needs at least one plan")
        cost = [float(np.prod(e.shape)) * sum(b.spots.n_spots for b in e.beams) for e in engines]
        self.groups = split_contiguous(cost, n_streams)
        self.grad_fn = grad_fn
        self.subs = []
        for g in self.groups:
            kwg = dict(kw)
            if grad_fn is not None and kwg.get("tune_grad_fn") is None:
                kwg["tune_grad_fn"] = lambda doses, first=g[0]: self._grads_of(doses, first)
            self.subs.append(BatchedDose([engines[i] for i in g], **kwg))
        self.engines = engines
        self.device = self.subs[0].device
        self.streams = [None] + [torch.cuda.Stream(self.device) for _ in self.subs[1:]]
        self.doses = [d for s in self.subs for d in s.doses]
        self.grad_doses = [g for s in self.subs for g in s.grad_doses]
        self.beam_grads = [bg for s in self.subs for bg in s.beam_grads]
        self.n_beams = sum(s.n_beams for s in self.subs)
        self._beam_off = np.cumsum([0] + [s.n_beams for s in self.subs]).tolist()
        self._plan_off = np.cumsum([0] + [len(g) for g in self.groups]).tolist()

    def _grads_of(self, doses, first):
        out = [torch.empty_like(d) for d in doses]
        for i, (d, o) in enumerate(zip(doses, out)):
            self.grad_fn(first + i, d, o)
        return out

    def _run(self, fn):
        """fn(j, sub-batch) for every sub-batch, each on its stream; joined on the current stream."""
        cur = torch.cuda.current_stream(self.device)
        for j in range(1, len(self.subs)):
            s = self.streams[j]
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                fn(j, self.subs[j])
        fn(0, self.subs[0])
        for s in self.streams[1:]:
            cur.wait_stream(s)

    def _beams(self, values, j):
        return None if values is None else values[self._beam_off[j]:self._beam_off[j + 1]]

    def _input_order(self):
        out = []
        k = 0
        for e in self.engines:
            for b in e.beams:
                gmu, gxy = self.beam_grads[k]
                out.append((b.spots.to_input_order(gmu), b.spots.to_input_order(gxy)))
                k += 1
        return out

    def forward(self, mus=None, xys=None):
        """Plan doses of every plan (as BatchedDose.forward)."""
        self._run(lambda j, s: s.forward(self._beams(mus, j), self._beams(xys, j)))
        return self.doses

    def adjoint(self, grad_doses=None, reuse_prep=False, input_order=True):
        """Spot gradients of every beam (as BatchedDose.adjoint)."""
        def one(j, s):
            g = None if grad_doses is None else grad_doses[self._plan_off[j]:self._plan_off[j + 1]]
            s.adjoint(g, reuse_prep=reuse_prep, input_order=False)
        self._run(one)
        return self._input_order() if input_order else self.beam_grads

    def step(self, mus=None, xys=None, grad_fn=None, input_order=True):
        """One optimisation step, each sub-batch on its stream from start to end:
        forward, dL/dD of every plan p by grad_fn(p, dose, out) (default: the
        constructor's; element-wise device work that writes out =
        grad_doses[p]; neither given: grad_doses as they are), adjoint.
        Returns (doses, beam gradients) as forward / adjoint."""
        grad_fn = self.grad_fn if grad_fn is None else grad_fn

        def one(j, s):
            doses = s.forward(self._beams(mus, j), self._beams(xys, j))
            if grad_fn is not None:
                for i, (d, g) in enumerate(zip(doses, s.grad_doses)):
                    grad_fn(self._plan_off[j] + i, d, g)
            s.adjoint(None, reuse_prep=True, input_order=False)
        self._run(one)
        return self.doses, (self._input_order() if input_order else self.beam_grads)

    def item_counts(self):
        """Work items of each phase, summed over the sub-batches."""
        tot = {}
        for s in self.subs:
            for k, v in s.item_counts().items():
                tot[k] = tot.get(k, 0) + v
        return tot


class GraphedStep:
    """Capture ``fn()`` (device work only: batched phases and torch ops, no
    host synchronisation) in a HIP graph; ``replay()`` reruns it and returns
    what ``fn`` returned at capture (tensors refreshed in place)."""

    def __init__(self, fn, device, warmup=2):
        self.device = device
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = fn()
        torch.cuda.synchronize(device)

    def replay(self):
        self.graph.replay()
        return self.out
