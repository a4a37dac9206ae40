"""dh_pb_adjoint_noalloc (library-held scratch, no workspace argument) vs
dh_pb_adjoint (caller's workspace), through the C ABI (GPU).

Two calls in a row on one device, the second with a smaller input:
  1. 65610 spots in 10 layers on an 80^3 grid with a 6-sigma cutoff: the
     workspace layout of dh_pb_adjoint (64 slabs of partial sums) exceeds
     the 64 MiB array, so the no-alloc call runs fewer slabs; gradients
     within 1e-5 of the workspace call (max norm, relative);
  2. 320 spots on a 40^3 water-slab grid, dense (cutoff 0): same layout as
     dh_pb_adjoint, gradients bit for bit -- and nothing left in the array by
     call 1 leaks into them.
A beam whose layout exceeds the array even with one slab is rejected with
DH_ERR_ARG and a message naming the scratch limit, before any kernel runs.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[no-alloc] " + msg, flush=True)


def _beam_inputs(dev, g, plan, k=0):
    from DoseCUDA import device as D
    from DoseCUDA.plan_impt import _model_for_beam
    beam = plan.beam_list[k]
    model = _model_for_beam(plan, beam)
    grid = D.make_grid(g.HU.shape, float(np.float32(g.spacing[0])))
    geo = D.make_geom(D.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                      beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY))
    dmodel = D.DeviceBeamModel(model, dev)
    spots = D.DeviceSpots(beam.spot_list, model.divergence_params, dev)
    rlsp = torch.from_numpy(np.asarray(g.RLSPFromHU(plan.machine_name), np.float32)).to(dev)
    _, wet = D.wet_from_padded(D.rlsp_pad(rlsp, grid, dev), grid, geo, dev)
    return grid, geo, dmodel, spots, wet


def _both(dev, grid, geo, dmodel, spots, wet, gd, cutoff):
    from DoseCUDA import _native as N
    from DoseCUDA import device as D
    lib = N.load()
    st = N.stream_handle(dev)
    S = spots.n_spots
    ws, nb = D.alloc_workspace(S, spots.n_layers, grid, dev)
    ref_mu = torch.full((S,), float("nan"), device=dev)
    ref_xy = torch.full((S, 2), float("nan"), device=dev)
    N.check(lib.dh_pb_adjoint(N.ptr(wet), N.ptr(gd), grid, ctypes.byref(geo), ctypes.byref(dmodel.c),
                              ctypes.byref(spots.c), float(cutoff), N.ptr(ref_mu), N.ptr(ref_xy), N.ptr(ws), nb, st),
            "dh_pb_adjoint")
    mu = torch.full((S,), float("nan"), device=dev)
    xy = torch.full((S, 2), float("nan"), device=dev)
    N.check(lib.dh_pb_adjoint_noalloc(N.ptr(wet), N.ptr(gd), grid, ctypes.byref(geo), ctypes.byref(dmodel.c),
                                      ctypes.byref(spots.c), float(cutoff), N.ptr(mu), N.ptr(xy), st),
            "dh_pb_adjoint_noalloc")
    torch.cuda.synchronize()
    return ref_mu.cpu().numpy(), ref_xy.cpu().numpy(), mu.cpu().numpy(), xy.cpu().numpy(), nb


def test_noalloc_adjoint_matches_workspace_adjoint():
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, _grid, grid_spots, water_slab_plan
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dev = torch.device("cuda", 0)
    # call 1: 65610 spots in 10 layers (81 x 81 on a 1 mm lattice each), 80^3 @ 2.5 mm, cutoff 6
    g = _grid((80, 80, 80), 2.5)
    rng = np.random.default_rng(3)
    g.HU = rng.normal(20.0, 30.0, size=(80, 80, 80)).astype(np.float32)
    c = np.arange(-40.0, 40.5, 1.0)
    sl = grid_spots(list(range(30, 40)), c, c)
    sl[:, 2] = rng.uniform(0.1, 1.0, size=sl.shape[0]).astype(np.float32)
    plan = IMPTPlan()
    plan.addBeam(_beam(30.0, sl, couch=10.0))
    grid, geo, dmodel, spots, wet = _beam_inputs(dev, g, plan)
    gd = torch.from_numpy(rng.normal(size=g.HU.shape).astype(np.float32)).to(dev)
    _say("call 1: %d spots, %d layers, grid 80^3, cutoff 6" % (spots.n_spots, spots.n_layers))
    rm, rx, m, x, nb = _both(dev, grid, geo, dmodel, spots, wet, gd, 6.0)
    assert nb > (64 << 20), nb   # dh_pb_adjoint's layout exceeds the array: the no-alloc call cuts slabs
    em = np.abs(m - rm).max() / np.abs(rm).max()
    ex = np.abs(x - rx).max() / np.abs(rx).max()
    _say("call 1: workspace layout %.1f MiB; rel diff mu %.2e xy %.2e" % (nb / 2 ** 20, em, ex))
    assert np.all(np.isfinite(m)) and np.all(np.isfinite(x))
    assert em <= 1e-5 and ex <= 1e-5
    del wet, gd, spots
    # call 2: smaller, dense: bit for bit
    g2, plan2 = water_slab_plan(n=40, spacing=2.5)
    grid, geo, dmodel, spots, wet = _beam_inputs(dev, g2, plan2)
    gd = torch.from_numpy(np.random.default_rng(4).normal(size=g2.HU.shape).astype(np.float32)).to(dev)
    _say("call 2: %d spots, %d layers, grid 40^3, cutoff 0" % (spots.n_spots, spots.n_layers))
    rm, rx, m, x, nb = _both(dev, grid, geo, dmodel, spots, wet, gd, 0.0)
    _say("call 2: identical mu %s xy %s" % (np.array_equal(m, rm), np.array_equal(x, rx)))
    assert np.abs(rm).max() > 0
    assert np.array_equal(m, rm) and np.array_equal(x, rx)


def test_noalloc_adjoint_rejects_beams_beyond_the_array():
    from DoseCUDA import _native as N
    from DoseCUDA import device as D
    from DoseCUDA import IMPTPlan
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dev = torch.device("cuda", 0)
    lib = N.load()
    S = 2_000_000   # 48 bytes per spot alone exceed 64 MiB
    spots = torch.zeros((S, 4), dtype=torch.float32, device=dev)
    layers = torch.zeros((1, 5), dtype=torch.float32, device=dev)
    ss = N.dh_spot_set(N.ptr(spots).value, N.ptr(layers).value, S, 1)
    grid = D.make_grid((8, 8, 8), 2.0)
    model = IMPTPlan().beam_models[0]
    dmodel = D.DeviceBeamModel(model, dev)
    geo = D.make_geom(D.beam_geometry(np.zeros(3, np.float32), np.zeros(3, np.float32), 0.0, 0.0,
                                      model.VSADX, model.VSADY))
    vol = torch.zeros((8, 8, 8), dtype=torch.float32, device=dev)
    out_mu = torch.zeros(S, device=dev)
    out_xy = torch.zeros((S, 2), device=dev)
    rc = lib.dh_pb_adjoint_noalloc(N.ptr(vol), N.ptr(vol), grid, ctypes.byref(geo), ctypes.byref(dmodel.c),
                                   ctypes.byref(ss), 6.0, N.ptr(out_mu), N.ptr(out_xy), N.stream_handle(dev))
    msg = N.last_error()
    _say("rejected: rc %d, %s" % (rc, msg))
    assert rc == N.DH_ERR_ARG and "scratch" in msg and str(64 << 20) in msg
    torch.cuda.synchronize()
    assert float(out_mu.abs().max()) == 0.0
