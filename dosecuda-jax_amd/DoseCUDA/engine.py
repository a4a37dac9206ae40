"""Device-resident proton dose engine and its differentiable entry point.

``IMPTDoseEngine(dose_grid, plan)`` keeps everything on one HIP device:

* HU -> RLSP in one fused kernel (dh_rlsp_from_hu) straight into the padded
  layout the ray-march reads;
* WET per beam computed once and cached (the WET-cache pattern of
  computeIMPTWET / computeIMPTPlanFromWET, plan_impt.py:208-257);
* spot weights / positions as device tensors; the forward dose and the
  hand-written adjoint (dL/dweight, dL/d(x, y)) per beam.

``ProtonDose.apply(engine, mu_0, xy_0, mu_1, xy_1, ...)`` is a
torch.autograd.Function over the plan dose (sum of beams times n_fractions):
its backward calls the adjoint kernels, which recompute the WET-derived
per-voxel terms instead of storing per-spot kernels.
"""

import os

import numpy as np
import pandas as pd
import torch

from . import _native as N
from . import device as D
from .plan_impt import _model_for_beam, _data_path

_native = N.load()


def _hu_knots(machine_name):
    t = pd.read_csv(_data_path("lookuptables", "protons", machine_name, "HU_RLSP.csv"))
    return t["HU"].to_numpy(np.float64), t["RLSP"].to_numpy(np.float64)


class BeamState:
    """Per-beam device state: geometry, model tables, spots, WET, workspace."""

    def __init__(self, engine, beam, model, dmodel):
        self.beam = beam
        self.model = model
        self.dmodel = dmodel
        self.geo = D.make_geom(D.beam_geometry(beam.iso, np.asarray(engine.origin, np.float32),
                                               beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY))
        sl = np.asarray(beam.spot_list, dtype=np.float32).reshape(-1, 4)
        self.spots = D.DeviceSpots(sl, model.divergence_params, engine.device, spatial=engine.spatial_sort)
        self._grid = engine.grid
        self._ws = None
        self.mu0 = torch.from_numpy(sl[:, 2].copy()).to(engine.device)
        self.xy0 = torch.from_numpy(sl[:, :2].copy()).to(engine.device)
        self.wet_raw = None
        self.wet = None
        self.wet_valid = False   # wet holds this beam's WET (set by whoever computed it)

    def _workspace(self):
        # single-beam pencil-beam workspace, allocated on first use (batches bring their own)
        if self._ws is None:
            self._ws = D.alloc_workspace(self.spots.n_spots, self.spots.n_layers, self._grid, self.spots.device)
        return self._ws

    @property
    def ws(self):
        return self._workspace()[0]

    @property
    def ws_bytes(self):
        return self._workspace()[1]


class IMPTDoseEngine:
    def __init__(self, dose_grid, plan, gpu_id=0, cutoff_sigma=D.DEFAULT_CUTOFF_SIGMA, spatial_sort=True):
        if dose_grid.spacing[0] != dose_grid.spacing[1] or dose_grid.spacing[0] != dose_grid.spacing[2]:
            raise Exception("Spacing must be isotropic for IMPT dose calculation - consider resampling CT")
        self.device = torch.device("cuda", int(gpu_id))
        self.cutoff_sigma = float(cutoff_sigma)
        self.spatial_sort = spatial_sort
        self.plan = plan
        self.n_fractions = plan.n_fractions
        hu = np.ascontiguousarray(dose_grid.HU, dtype=np.float32)
        self.shape = hu.shape
        self.spacing = float(np.float32(dose_grid.spacing[0]))
        self.origin = np.asarray(dose_grid.origin, np.float32)
        self.grid = D.make_grid(hu.shape, self.spacing)
        with torch.cuda.device(self.device):
            self.hu = torch.from_numpy(hu).to(self.device)
            hk, rk = _hu_knots(plan.machine_name)
            self.rlsp_padded = D.rlsp_from_hu(self.hu, self.grid, hk, rk, self.device)
            dmodels = {}
            self.beams = []
            for beam in plan.beam_list:
                model = _model_for_beam(plan, beam)
                if id(model) not in dmodels:
                    dmodels[id(model)] = D.DeviceBeamModel(model, self.device)
                self.beams.append(BeamState(self, beam, model, dmodels[id(model)]))

    # -- WET -----------------------------------------------------------------
    def compute_wet(self, force=False):
        for b in self.beams:
            if not b.wet_valid or force:
                b.wet_raw, b.wet = D.wet_from_padded(self.rlsp_padded, self.grid, b.geo, self.device,
                                                     raw_out=b.wet_raw, out=b.wet)
                b.wet_valid = True
        return [b.wet for b in self.beams]

    # -- forward / adjoint on explicit tensors --------------------------------
    def _set_spots(self, b, mu, xy):
        b.spots.set_weights_positions(mu if mu is not None else b.mu0, xy if xy is not None else b.xy0)

    def beam_dose(self, k, mu=None, xy=None, out=None, accumulate=False):
        """Dose of beam k (not times n_fractions)."""
        b = self.beams[k]
        self.compute_wet()
        self._set_spots(b, mu, xy)
        dose = out if out is not None else torch.empty(self.shape, dtype=torch.float32, device=self.device)
        D.pb_forward(b.wet, dose, self.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                     cutoff_sigma=self.cutoff_sigma, accumulate=accumulate)
        return dose

    def dose(self, mus=None, xys=None, out=None):
        """Plan dose: sum of beams times n_fractions."""
        out = out if out is not None else torch.empty(self.shape, dtype=torch.float32, device=self.device)
        if not self.beams:
            return out.zero_()
        for k in range(len(self.beams)):
            self.beam_dose(k, None if mus is None else mus[k], None if xys is None else xys[k], out=out,
                           accumulate=k > 0)
        if self.n_fractions != 1:
            out.mul_(float(self.n_fractions))
        return out

    def adjoint(self, grad_dose, mus=None, xys=None):
        """dL/dmu and dL/d(x, y) per beam, in each beam's spot_list row order,
        for dL/dD = grad_dose of the plan dose."""
        self.compute_wet()
        g = grad_dose.contiguous().to(torch.float32)
        if self.n_fractions != 1:
            g = g * float(self.n_fractions)
        res = []
        for k, b in enumerate(self.beams):
            self._set_spots(b, None if mus is None else mus[k], None if xys is None else xys[k])
            gmu, gxy = D.pb_adjoint(b.wet, g, self.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                                    cutoff_sigma=self.cutoff_sigma)
            res.append((b.spots.to_input_order(gmu), b.spots.to_input_order(gxy)))
        return res

    # -- autograd --------------------------------------------------------------
    def parameters(self, requires_grad=True):
        """Fresh leaf tensors (mu_k, xy_k) initialised from the plan's spot lists."""
        out = []
        for b in self.beams:
            out.append((b.mu0.clone().requires_grad_(requires_grad), b.xy0.clone().requires_grad_(requires_grad)))
        return out

    def differentiable_dose(self, params):
        flat = []
        for mu, xy in params:
            flat += [mu, xy]
        return ProtonDose.apply(self, *flat)


class ProtonDose(torch.autograd.Function):
    """Plan dose as a function of every beam's spot weights and positions."""

    @staticmethod
    def forward(ctx, engine, *flat):
        mus = list(flat[0::2])
        xys = list(flat[1::2])
        ctx.engine = engine
        ctx.save_for_backward(*flat)
        with torch.no_grad():
            return engine.dose(mus, xys)

    @staticmethod
    def backward(ctx, grad_dose):
        flat = ctx.saved_tensors
        mus = list(flat[0::2])
        xys = list(flat[1::2])
        res = ctx.engine.adjoint(grad_dose, mus, xys)
        grads = []
        for k, (gmu, gxy) in enumerate(res):
            grads.append(gmu.to(mus[k].dtype) if ctx.needs_input_grad[1 + 2 * k] else None)
            grads.append(gxy.to(xys[k].dtype) if ctx.needs_input_grad[2 + 2 * k] else None)
        return (None, *grads)
