"""Device-resident building blocks over the C ABI of libdosehip.so.

Everything here keeps data in HIP device memory (torch tensors used as
allocations) and enqueues work on torch's current stream of the device; no
call synchronises.  Higher layers (``dose_kernels`` drop-in, ``engine``) are
built from these pieces.
"""

import ctypes

import numpy as np
import torch

from . import _native as N
from .geometry import beam_geometry, sort_spots, layer_table

# Default beam's-eye-view cutoff (in halo sigmas) for skipping spot/voxel pairs:
# every skipped term is below exp(-6^2/2) = 1.5e-8 of its spot's peak dose.
DEFAULT_CUTOFF_SIGMA = 6.0


def make_grid(shape_zyx, spacing):
    nz, ny, nx = (int(v) for v in shape_zyx)
    return N.dh_grid(nx, ny, nz, float(spacing))


def make_geom(g):
    """dh_beam_geom (float32) from a beam_geometry() dict."""
    geo = N.dh_beam_geom()
    for a in range(3):
        geo.iso[a] = float(g["iso"][a])
        geo.src[a] = float(g["src"][a])
    geo.sin_g, geo.cos_g = g["sin_g"], g["cos_g"]
    geo.sin_c, geo.cos_c = g["sin_c"], g["cos_c"]
    geo.vsadx, geo.vsady = g["vsadx"], g["vsady"]
    return geo


class DeviceBeamModel:
    """Beam-model lookup tables uploaded once: (IDD, sigma) interleaved float2
    rows, depth rows and divergence parameters (IMPTBeamModel attributes)."""

    def __init__(self, model, device):
        dvp = np.ascontiguousarray(model.divergence_params, dtype=np.float32)
        dep = np.ascontiguousarray(model.lut_depths, dtype=np.float32)
        sig = np.ascontiguousarray(model.lut_sigmas, dtype=np.float32)
        idd = np.ascontiguousarray(model.lut_idds, dtype=np.float32)
        if dep.shape != sig.shape or dep.shape != idd.shape or dep.ndim != 2:
            raise ValueError("lut_depths, lut_sigmas and lut_idds must have the same 2-D shape")
        self.n_energies, self.lut_len = dep.shape
        self.dvp_len = dvp.shape[1]
        if dvp.shape[0] < self.n_energies or self.dvp_len < 5:
            raise ValueError("divergence_params must be [n_energies, >=5]")
        inter = np.stack([idd, sig], axis=-1)  # [n_E, L, 2]
        self.device = device
        self.idd_sigma = torch.from_numpy(np.ascontiguousarray(inter)).to(device)
        self.depths = torch.from_numpy(dep).to(device)
        self.dvp = torch.from_numpy(dvp).to(device)
        d0 = float(dep[0, 0])
        dd = float(dep[0, 1] - dep[0, 0]) if self.lut_len > 1 else 1.0
        k = np.arange(self.lut_len, dtype=np.float32)
        uniform = dd > 0 and bool(np.all(dep == (np.float32(d0) + k * np.float32(dd))[None, :]))
        self.c = N.dh_beam_model(N.ptr(self.idd_sigma).value, N.ptr(self.depths).value, N.ptr(self.dvp).value,
                                 self.n_energies, self.lut_len, self.dvp_len, 1 if uniform else 0,
                                 d0, dd if uniform else 1.0)
        self.uniform = uniform
        self.divergence_params = dvp


class DeviceSpots:
    """One beam's spots sorted by energy id (spatially inside a layer) plus its
    layer table, on the device.  ``order`` maps sorted position -> input row."""

    def __init__(self, spot_list, divergence_params, device, spatial=True):
        sl = np.asarray(spot_list, dtype=np.float32).reshape(-1, 4)
        order, sorted_ids = sort_spots(sl, spatial=spatial)
        tab, r80, energy = layer_table(sorted_ids, divergence_params)
        self.order = order
        self.n_input = sl.shape[0]
        self.n_spots = int(sl.shape[0])
        self.n_layers = int(tab.shape[0])
        rec = np.zeros((self.n_spots, 4), dtype=np.float32)
        rec[:, :3] = sl[order, :3]
        rec.view(np.int32)[:, 3] = sorted_ids
        lay = np.zeros((self.n_layers, 5), dtype=np.float32)
        lay.view(np.int32)[:, :3] = tab
        lay[:, 3] = r80
        lay[:, 4] = energy
        self.layers_np = tab
        self.device = device
        self.spots = torch.from_numpy(rec).to(device)
        self.layers = torch.from_numpy(lay).to(device)
        self.order_t = torch.from_numpy(order.astype(np.int64)).to(device)
        self.c = N.dh_spot_set(N.ptr(self.spots).value if self.n_spots else None,
                               N.ptr(self.layers).value if self.n_layers else None,
                               self.n_spots, self.n_layers)

    def set_weights_positions(self, mu=None, xy=None):
        """Overwrite weights / positions from tensors in INPUT spot order
        (device): mu [S], xy [S, 2]; one library launch (dh_spot_params)."""
        if mu is None and xy is None:
            return
        if mu is not None:
            mu = mu.detach().to(device=self.device, dtype=torch.float32).contiguous()
            if mu.numel() != self.n_spots:
                raise ValueError("spot weights: %d values for %d spots" % (mu.numel(), self.n_spots))
        if xy is not None:
            xy = xy.detach().to(device=self.device, dtype=torch.float32).contiguous()
            if xy.numel() != 2 * self.n_spots:
                raise ValueError("spot positions: %d values for %d spots" % (xy.numel(), self.n_spots))
        if self.n_spots == 0:
            return
        lib = N.load()
        N.check(lib.dh_spot_params(N.ptr(self.spots), N.ptr(self.order_t), None if mu is None else N.ptr(mu),
                                   None if xy is None else N.ptr(xy), self.n_spots, N.stream_handle(self.device)),
                "dh_spot_params")

    def to_input_order(self, sorted_vals):
        out = torch.empty_like(sorted_vals)
        out[self.order_t] = sorted_vals
        return out


def alloc_workspace(n_spots, n_layers, grid, device):
    nbytes = int(N.load().dh_pb_workspace_size(n_spots, n_layers, grid))
    # torch's caching allocator returns 512-byte aligned blocks
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
    return ws, nbytes


def rlsp_pad(rlsp, grid, device):
    lib = N.load()
    out = torch.empty(int(lib.dh_padded_len(grid)), dtype=torch.float32, device=device)
    N.check(lib.dh_rlsp_pad(N.ptr(rlsp), N.ptr(out), grid, N.stream_handle(device)), "dh_rlsp_pad")
    return out


def rlsp_from_hu(hu, grid, hu_knots, rlsp_knots, device):
    lib = N.load()
    out = torch.empty(int(lib.dh_padded_len(grid)), dtype=torch.float32, device=device)
    hk = np.ascontiguousarray(hu_knots, dtype=np.float64)
    rk = np.ascontiguousarray(rlsp_knots, dtype=np.float64)
    rc = lib.dh_rlsp_from_hu(N.ptr(hu), N.ptr(out), grid, hk.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             rk.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(hk), N.stream_handle(device))
    N.check(rc, "dh_rlsp_from_hu")
    return out


def wet_from_padded(rlsp_padded, grid, geom, device, raw_out=None, out=None):
    """Ray-march + smoothing: returns (raw WET, smoothed WET) as (z, y, x)."""
    lib = N.load()
    shape = (grid.nz, grid.ny, grid.nx)
    raw = raw_out if raw_out is not None else torch.empty(shape, dtype=torch.float32, device=device)
    wet = out if out is not None else torch.empty(shape, dtype=torch.float32, device=device)
    st = N.stream_handle(device)
    N.check(lib.dh_wet_raymarch(N.ptr(rlsp_padded), N.ptr(raw), grid, ctypes.byref(geom), st), "dh_wet_raymarch")
    N.check(lib.dh_wet_smooth(N.ptr(raw), N.ptr(wet), grid, ctypes.byref(geom), st), "dh_wet_smooth")
    return raw, wet


def pb_forward(wet, dose, grid, geom, dmodel, dspots, workspace, ws_bytes, cutoff_sigma=DEFAULT_CUTOFF_SIGMA,
               accumulate=False):
    lib = N.load()
    rc = lib.dh_pb_forward(N.ptr(wet), N.ptr(dose), grid, ctypes.byref(geom), ctypes.byref(dmodel.c),
                           ctypes.byref(dspots.c), float(cutoff_sigma), 1 if accumulate else 0,
                           N.ptr(workspace), int(ws_bytes), N.stream_handle(dose.device))
    N.check(rc, "dh_pb_forward")
    return dose


def pb_adjoint(wet, grad_dose, grid, geom, dmodel, dspots, workspace, ws_bytes,
               cutoff_sigma=DEFAULT_CUTOFF_SIGMA, grad_mu=None, grad_xy=None):
    """dL/dmu [S] and dL/d(x, y) [S, 2] in SORTED spot order."""
    lib = N.load()
    dev = grad_dose.device
    S = dspots.n_spots
    gmu = grad_mu if grad_mu is not None else torch.empty(S, dtype=torch.float32, device=dev)
    gxy = grad_xy if grad_xy is not None else torch.empty((S, 2), dtype=torch.float32, device=dev)
    if S == 0:
        return gmu, gxy
    rc = lib.dh_pb_adjoint(N.ptr(wet), N.ptr(grad_dose), grid, ctypes.byref(geom), ctypes.byref(dmodel.c),
                           ctypes.byref(dspots.c), float(cutoff_sigma), N.ptr(gmu), N.ptr(gxy),
                           N.ptr(workspace), int(ws_bytes), N.stream_handle(dev))
    N.check(rc, "dh_pb_adjoint")
    return gmu, gxy


def pb_count_pairs(wet, grad_dose, grid, geom, dmodel, dspots, workspace, ws_bytes,
                   cutoff_sigma=DEFAULT_CUTOFF_SIGMA, counts=None):
    """Accumulate the work counts of dh_pb_count_pairs into a device int64 [4]
    tensor (forward dense, forward evaluated, adjoint dense, adjoint evaluated)."""
    lib = N.load()
    dev = wet.device
    if counts is None:
        counts = torch.zeros(4, dtype=torch.int64, device=dev)
    rc = lib.dh_pb_count_pairs(N.ptr(wet), N.ptr(grad_dose) if grad_dose is not None else None, grid,
                               ctypes.byref(geom), ctypes.byref(dmodel.c), ctypes.byref(dspots.c),
                               float(cutoff_sigma), N.ptr(counts), N.ptr(workspace), int(ws_bytes),
                               N.stream_handle(dev))
    N.check(rc, "dh_pb_count_pairs")
    return counts


def geometry_for(beam, origin, model):
    return beam_geometry(beam.iso, origin, beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY)
