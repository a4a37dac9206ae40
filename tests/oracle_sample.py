"""The FP64 oracle evaluated at sampled voxels (test helper).

Plans of the head, uneven-batch and body sizes (6.4e6 .. 7.9e7 voxels, up to
63k spots) are too large for a full float64 oracle pass per test, but the dose
at a voxel needs only that voxel's smoothed WET, and its smoothed WET only the
raw WET of the voxel and of its <= 54 smoothing neighbours, each one ray
(impt_jax.py:235-418).  ``OraclePlan`` therefore marches exactly those rays
(``O.wet_at``) and evaluates the dose (``O.pb_dose`` with ``idx``) at the
sample, and the gradient (``O.pb_grad``) for a dL/dD that is zero off the
sample.  ``gamma_at`` is the oracle's global gamma index evaluated at sampled
reference voxels against a full test volume.  CPU tests check each piece
against the full-volume oracle on small plans (tests/test_oracle_sample.py).
"""

import math

import numpy as np

from oracle import impt_oracle as O


def sample_voxels(dose, n_hi, n_rand, seed=0, frac=0.1, extra=()):
    """Sorted unique flat indices: n_hi voxels where dose > frac * max, n_rand
    uniformly over the volume, the argmax of dose, and any `extra`."""
    flat = np.asarray(dose).reshape(-1)
    rng = np.random.default_rng(seed)
    parts = [np.array([int(np.argmax(flat))], np.int64), np.asarray(extra, np.int64).reshape(-1)]
    hi = np.flatnonzero(flat > frac * flat.max()) if flat.max() > 0 else np.zeros(0, np.int64)
    if hi.size:
        parts.append(rng.choice(hi, size=min(n_hi, hi.size), replace=False))
    parts.append(rng.choice(flat.size, size=min(n_rand, flat.size), replace=False))
    return np.unique(np.concatenate(parts).astype(np.int64))


class OracleBeam:
    """Oracle inputs of one beam: geometry, sorted spots, layers, model tables."""

    def __init__(self, g, plan, beam):
        self.model = O.beam_model_for(plan, beam)
        self.geom = O.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                    beam.gantry_angle, beam.couch_angle, self.model.VSADX, self.model.VSADY)
        self.spot_list = np.array(beam.spot_list, dtype=np.float32).reshape(-1, 4)
        self.marr = O.ModelArrays(self.model)

    def spots(self, mu=None, xy=None):
        sl = self.spot_list.copy()
        if mu is not None:
            sl[:, 2] = np.asarray(mu, np.float32).reshape(-1)
        if xy is not None:
            sl[:, :2] = np.asarray(xy, np.float32).reshape(-1, 2)
        return O.spots_and_layers(sl, self.model.divergence_params)


class OraclePlan:
    """Oracle dose and gradients of a plan at the voxels `idx` (flat indices of
    the (z, y, x) grid), as computeIMPTPlanJax (impt_jax.py:866-929) gives
    them there: beam doses summed and times n_fractions."""

    def __init__(self, g, plan, machine_dir, idx, rlsp=None):
        self.shape = tuple(int(v) for v in np.asarray(g.HU).shape)
        self.idx = np.ascontiguousarray(idx, dtype=np.int64)
        self.s = float(np.float32(g.spacing[0]))
        self.n_fx = plan.n_fractions
        rho = O.rlsp_from_hu(g.HU, machine_dir) if rlsp is None else rlsp
        self.beams = [OracleBeam(g, plan, b) for b in plan.beam_list]
        self.raw = [O.raymarch(rho, self.s, ob.geom, idx=self.idx) for ob in self.beams]
        self.wet = [O.wet_at(rho, self.s, ob.geom, self.idx) for ob in self.beams]

    def _wet_volume(self, k):
        w = np.zeros(self.shape)
        w.reshape(-1)[self.idx] = self.wet[k]
        return w

    def beam_dose(self, k, mu=None, xy=None):
        """Dose of beam k at the sample (not times n_fractions)."""
        ob = self.beams[k]
        _, spots, layers, r80 = ob.spots(mu, xy)
        return O.pb_dose(self._wet_volume(k), self.s, ob.geom, spots, layers, r80, ob.marr, idx=self.idx)

    def dose(self, mus=None, xys=None):
        d = np.zeros(self.idx.size)
        for k in range(len(self.beams)):
            d += self.beam_dose(k, None if mus is None else mus[k], None if xys is None else xys[k])
        return d * self.n_fx

    def grad(self, gvals, mus=None, xys=None):
        """[(dL/dmu [S], dL/d(x, y) [S, 2])] per beam in spot_list row order, for
        dL/dD(plan) = gvals at the sample and 0 elsewhere."""
        gvol = np.zeros(self.shape)
        gvol.reshape(-1)[self.idx] = np.asarray(gvals, np.float64) * self.n_fx
        out = []
        for k, ob in enumerate(self.beams):
            order, spots, layers, r80 = ob.spots(None if mus is None else mus[k], None if xys is None else xys[k])
            gmu_s, gxy_s = O.pb_grad(self._wet_volume(k), gvol, self.s, ob.geom, spots, layers, r80, ob.marr)
            gmu = np.zeros(order.size)
            gxy = np.zeros((order.size, 2))
            gmu[order] = gmu_s
            gxy[order] = gxy_s
            out.append((gmu, gxy))
        return out


def gamma_at(ref_vals, idx, test, spacing, dose_pct=2.0, dta_mm=2.0, threshold_pct=10.0, dmax=None):
    """Global gamma pass rate over the sampled reference voxels whose reference
    dose is >= threshold_pct of D_max (D_max = max of ref_vals unless given),
    searching the full test volume on the voxel grid inside the DTA ball with
    edge clamping -- O.gamma_pass_rate restricted to the sample."""
    test = np.asarray(test, np.float64)
    ref = np.asarray(ref_vals, np.float64)
    dmax = float(ref.max()) if dmax is None else float(dmax)
    if dmax <= 0:
        return 1.0
    dd = dose_pct / 100.0 * dmax
    keep = ref >= threshold_pct / 100.0 * dmax
    if not keep.any():
        return 1.0
    k, j, i = np.unravel_index(np.asarray(idx)[keep], test.shape)
    r0 = ref[keep]
    r = int(math.ceil(dta_mm / spacing))
    nz, ny, nx = test.shape
    best = np.full(r0.shape, np.inf)
    for dz in range(-r, r + 1):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                dist2 = (dx * dx + dy * dy + dz * dz) * spacing * spacing
                if dist2 > dta_mm * dta_mm * (1 + 1e-9) and (dx or dy or dz):
                    continue
                t = test[np.clip(k + dz, 0, nz - 1), np.clip(j + dy, 0, ny - 1), np.clip(i + dx, 0, nx - 1)]
                np.minimum(best, dist2 / (dta_mm * dta_mm) + ((t - r0) / dd) ** 2, out=best)
    return float(np.mean(best <= 1.0))


def rel_err(test, ref):
    """max |test - ref| / max |ref| (0 when both vanish)."""
    test = np.asarray(test, np.float64)
    ref = np.asarray(ref, np.float64)
    den = float(np.abs(ref).max()) if ref.size else 0.0
    num = float(np.abs(test - ref).max()) if ref.size else 0.0
    return num / den if den > 0 else num
