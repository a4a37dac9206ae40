"""The C ABI library: loads, exports every declared symbol, struct layouts agree
with the header, argument errors come back without touching a GPU."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dosehip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(dh_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from DoseCUDA import _native as N
    lib = N.load()
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, "ctypes binding missing for %s" % name
    assert lib.dh_version() >= 100


def test_struct_layout_matches_c(tmp_path):
    from DoseCUDA import _native as N
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "dosehip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dh_grid), sizeof(dh_beam_geom), sizeof(dh_spot),
         sizeof(dh_layer), sizeof(dh_beam_model), sizeof(dh_spot_set));
  printf("%zu %zu %zu %zu\n", offsetof(dh_beam_geom, sin_g), offsetof(dh_beam_model, n_energies),
         offsetof(dh_beam_model, depth0), offsetof(dh_spot_set, n_spots));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = [int(v) for v in out[0].split()]
    offs = [int(v) for v in out[1].split()]
    assert sizes == [ctypes.sizeof(c) for c in (N.dh_grid, N.dh_beam_geom, N.dh_spot, N.dh_layer,
                                                N.dh_beam_model, N.dh_spot_set)]
    assert offs == [N.dh_beam_geom.sin_g.offset, N.dh_beam_model.n_energies.offset,
                    N.dh_beam_model.depth0.offset, N.dh_spot_set.n_spots.offset]


def test_argument_errors_without_gpu():
    from DoseCUDA import _native as N
    lib = N.load()
    bad = N.dh_grid(0, 4, 4, 2.0)
    geo = N.dh_beam_geom()
    rc = lib.dh_wet_raymarch(None, None, bad, ctypes.byref(geo), None)
    assert rc == N.DH_ERR_ARG
    assert "invalid grid" in N.last_error()
    with pytest.raises(ValueError):
        N.check(rc, "dh_wet_raymarch")
    good = N.dh_grid(8, 8, 8, 2.0)
    # quads, then per 4x4x4 cell block a (value, run radius) pair and two ints of scratch
    assert lib.dh_padded_len(good) == 4 * 8 * 8 * 9 + 4 * 2 * 2 * 2
    assert lib.dh_padded_len(N.dh_grid(9, 5, 4, 2.0)) == 4 * 9 * 5 * 5 + 4 * 3 * 2 * 1
    huge = N.dh_grid(2000, 2000, 2000, 1.0)
    assert lib.dh_rlsp_pad(ctypes.c_void_p(16), ctypes.c_void_p(16), huge, None) == N.DH_ERR_ARG
    knots = (ctypes.c_double * 2)(0.0, 0.0)
    assert lib.dh_rlsp_from_hu(ctypes.c_void_p(16), ctypes.c_void_p(16), good, knots, knots, 2, None) \
        == N.DH_ERR_ARG  # knots not ascending
    assert lib.dh_pb_workspace_size(0, 0, good) == 0
    ws = lib.dh_pb_workspace_size(1000, 10, good)
    assert ws >= 1000 * 32 and ws % 256 == 0
    ss = N.dh_spot_set(None, None, 5, 1)
    model = N.dh_beam_model()
    rc = lib.dh_pb_forward(ctypes.c_void_p(16), ctypes.c_void_p(16), good, ctypes.byref(geo),
                           ctypes.byref(model), ctypes.byref(ss), 0.0, 0, None, 0, None)
    assert rc == N.DH_ERR_ARG


def test_batch_struct_layout_matches_c(tmp_path):
    from DoseCUDA import _native as N
    src = tmp_path / "blayout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "dosehip.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(dh_batch_plan), sizeof(dh_batch_beam), sizeof(dh_batch_tune),
         sizeof(dh_batch));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(dh_batch_plan, beam_start), offsetof(dh_batch_beam, spots),
         offsetof(dh_batch_beam, wet), offsetof(dh_batch_beam, grad_xy), offsetof(dh_batch, n_bricks),
         offsetof(dh_batch, events));
  return 0;
}
''')
    exe = tmp_path / "blayout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = [int(v) for v in out[0].split()]
    offs = [int(v) for v in out[1].split()]
    assert sizes == [ctypes.sizeof(c) for c in (N.dh_batch_plan, N.dh_batch_beam, N.dh_batch_tune, N.dh_batch)]
    assert offs == [N.dh_batch_plan.beam_start.offset, N.dh_batch_beam.spots.offset, N.dh_batch_beam.wet.offset,
                    N.dh_batch_beam.grad_xy.offset, N.dh_batch.n_bricks.offset, N.dh_batch.events.offset]


def _fake_batch(N, n_spots=(100, 0, 7), grid=(16, 12, 10)):
    """Descriptors with non-null fake device pointers (nothing is dereferenced
    by the sizing / validation code)."""
    fake = 4096
    nb = len(n_spots)
    plans = (N.dh_batch_plan * 2)()
    plans[0].grid = N.dh_grid(grid[0], grid[1], grid[2], 2.0)
    plans[0].rlsp_padded = fake
    plans[0].dose = fake
    plans[0].grad_dose = fake
    plans[0].beam_start = 0
    plans[0].n_beams = nb - 1
    plans[1].grid = N.dh_grid(8, 8, 8, 2.0)
    plans[1].rlsp_padded = fake
    plans[1].dose = fake
    plans[1].grad_dose = fake
    plans[1].beam_start = nb - 1
    plans[1].n_beams = 1
    beams = (N.dh_batch_beam * nb)()
    for k, S in enumerate(n_spots):
        B = beams[k]
        B.model = N.dh_beam_model(fake, fake, fake, 98, 400, 5, 1, 0.0, 1.0)
        B.spots = N.dh_spot_set(fake if S else None, fake if S else None, S, 3 if S else 0)
        B.wet = fake
        B.grad_mu = fake
        B.grad_xy = fake
    return plans, beams


def test_batch_sizes_and_errors_without_gpu():
    from DoseCUDA import _native as N
    lib = N.load()
    plans, beams = _fake_batch(N)
    tune = N.dh_batch_tune(0, 0, 0, 0, 0, 0)
    n = lib.dh_batch_workspace_size(plans, 2, beams, 3, ctypes.byref(tune))
    assert n > 0 and n % 256 == 0
    # more slabs -> more adjoint partials -> larger workspace
    big = N.dh_batch_tune(0, 1 << 20, 1, 70000, 0, 0)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, ctypes.byref(big)) == N.DH_ERR_ARG  # over 65535
    big = N.dh_batch_tune(0, 1 << 20, 1, 60000, 0, 0)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, ctypes.byref(big)) > n
    bad = N.dh_batch_tune(48, 0, 0, 0, 0, 0)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, ctypes.byref(bad)) == N.DH_ERR_ARG
    assert "schedule" in N.last_error()
    # beams must be listed in order without gaps
    plans[1].beam_start = 1
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, None) == N.DH_ERR_ARG
    assert "without gaps" in N.last_error()
    plans[1].beam_start = 2
    beams[0].wet = None
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, None) == N.DH_ERR_ARG
    beams[0].wet = 4096
    # spots without layers are a valid (empty) delivery; spots without a spot
    # array, or layers without a layer table, are not
    beams[0].spots = N.dh_spot_set(4096, None, 100, 0)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, None) > 0
    beams[0].spots = N.dh_spot_set(None, 4096, 100, 3)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, None) == N.DH_ERR_ARG
    assert "invalid spot set" in N.last_error()
    beams[0].spots = N.dh_spot_set(4096, None, 100, 3)
    assert lib.dh_batch_workspace_size(plans, 2, beams, 3, None) == N.DH_ERR_ARG
    beams[0].spots = N.dh_spot_set(4096, 4096, 100, 3)
    # setup rejects a small workspace before touching the device
    b = N.dh_batch()
    rc = lib.dh_batch_setup(ctypes.byref(b), plans, 2, beams, 3, None, 6.0, ctypes.c_void_p(4096), 256, None)
    assert rc == N.DH_ERR_ARG and "workspace too small" in N.last_error()
    neg = N.dh_batch_tune(0, 0, 0, 0, -1, 0)
    rc = lib.dh_batch_setup(ctypes.byref(b), plans, 2, beams, 3, ctypes.byref(neg), 6.0, ctypes.c_void_p(4096),
                            1 << 40, None)
    assert rc == N.DH_ERR_ARG and "brick schedule" in N.last_error()
    for shape in (-1, 6):
        bad_shape = N.dh_batch_tune(0, 0, 0, 0, 0, 0, shape)
        rc = lib.dh_batch_setup(ctypes.byref(b), plans, 2, beams, 3, ctypes.byref(bad_shape), 6.0,
                                ctypes.c_void_p(4096), 1 << 40, None)
        assert rc == N.DH_ERR_ARG and "ray_shape" in N.last_error()
    # phases refuse a batch that was never set up
    assert lib.dh_batch_forward(ctypes.byref(b), 0, None) == N.DH_ERR_ARG
    assert "not set up" in N.last_error()
    assert lib.dh_batch_timing(ctypes.byref(b), None) == N.DH_ERR_ARG
    assert b.prepared == 0


def test_geometry_matches_oracle_restatement():
    from DoseCUDA.geometry import beam_geometry
    from oracle import impt_oracle as O
    for ga, ca in [(0, 0), (90, 0), (120, 30), (240, -15), (359.5, 270)]:
        a = beam_geometry(np.array([1.5, -2.0, 3.0], np.float32), np.array([-100, -90, -80], np.float32), ga, ca,
                          1340.0, 1930.0)
        b = O.beam_geometry(np.array([1.5, -2.0, 3.0], np.float32), np.array([-100, -90, -80], np.float32), ga,
                            ca, 1340.0, 1930.0)
        flat = np.concatenate([a["iso"], a["src"], [a["sin_g"], a["cos_g"], a["sin_c"], a["cos_c"],
                                                    a["vsadx"], a["vsady"]]])
        np.testing.assert_allclose(flat, b, rtol=0, atol=1e-12)


def test_layer_table_matches_oracle(plan_models):
    from DoseCUDA.geometry import sort_spots, layer_table
    from oracle import impt_oracle as O
    dvp = plan_models.beam_models[0].divergence_params
    rng = np.random.default_rng(0)
    sl = np.column_stack([rng.uniform(-20, 20, 300), rng.uniform(-20, 20, 300), rng.uniform(0, 1, 300),
                          rng.integers(0, 110, 300)]).astype(np.float32)
    order, ids = sort_spots(sl)
    tab, r80, energy = layer_table(ids, dvp)
    _, _, lay_o, r80_o = O.spots_and_layers(sl, dvp)
    np.testing.assert_array_equal(tab, lay_o)
    np.testing.assert_array_equal(r80.astype(np.float64), r80_o)
    # spatial order inside a layer only permutes spots of equal energy
    assert np.all(np.diff(ids) >= 0)
    sl[5, 3] = -2
    order, ids = sort_spots(sl)
    assert layer_table(ids, dvp)[0].shape == (0, 3)


def test_photon_struct_layout_and_errors_without_gpu(tmp_path):
    """dh_mlc_pair / dh_photon_model / dh_photon_beam as C lays them out, and the
    photon entry points' argument checks (no kernel launched)."""
    from DoseCUDA import _native as N
    src = tmp_path / "layout_ph.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "dosehip.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(dh_mlc_pair), sizeof(dh_photon_model), sizeof(dh_photon_beam));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(dh_photon_model, kernel), offsetof(dh_photon_model, mu_cal),
         offsetof(dh_photon_model, electron_fitted_dmax), offsetof(dh_photon_beam, sin_col),
         offsetof(dh_photon_beam, mlc), offsetof(dh_photon_beam, mu));
  return 0;
}
''')
    exe = tmp_path / "layout_ph"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert [int(v) for v in out[0].split()] == [ctypes.sizeof(c) for c in (N.dh_mlc_pair, N.dh_photon_model,
                                                                            N.dh_photon_beam)]
    assert [int(v) for v in out[1].split()] == [N.dh_photon_model.kernel.offset, N.dh_photon_model.mu_cal.offset,
                                                N.dh_photon_model.electron_fitted_dmax.offset,
                                                N.dh_photon_beam.sin_col.offset, N.dh_photon_beam.mlc.offset,
                                                N.dh_photon_beam.mu.offset]
    lib = N.load()
    g = N.dh_grid(10, 12, 14, 2.0)
    # 12 floats per padded cell, then 4 per block of 4x4x4 sample cells (11 x 13 x 15 cells)
    assert lib.dh_photon_record_len(g) == 12 * 12 * 14 * 16 + 4 * 3 * 4 * 4
    assert lib.dh_photon_record_len(N.dh_grid(0, 12, 14, 2.0)) < 0
    assert lib.dh_photon_record_len(N.dh_grid(2000, 2000, 2000, 1.0)) < 0   # 32-bit record offsets
    assert "too large" in N.last_error()
    beam = N.dh_photon_beam()
    model = N.dh_photon_model()
    assert lib.dh_photon_density(None, None, g, None) == N.DH_ERR_ARG
    assert lib.dh_photon_raytrace(None, None, g, ctypes.byref(beam), None) == N.DH_ERR_ARG
    assert lib.dh_photon_terma(None, None, None, g, ctypes.byref(model), ctypes.byref(beam), None) == N.DH_ERR_ARG
    assert lib.dh_photon_terma_records(None, None, g, None) == N.DH_ERR_ARG
    assert lib.dh_photon_ccc(None, None, None, None, g, ctypes.byref(model), ctypes.byref(beam), 1.0, 0,
                             None) == N.DH_ERR_ARG
    # a misaligned record volume is refused before any launch
    assert lib.dh_photon_raytrace(ctypes.c_void_p(16), ctypes.c_void_p(64), g, ctypes.byref(beam), None) == \
        N.DH_ERR_ARG
    assert "32-byte aligned" in N.last_error()
    # model checks, all before any launch (the pointers are never dereferenced on the host)
    fake = ctypes.c_void_p(4096)
    model.profile = model.spectrum = 4096
    model.n_profile, model.n_spectrum = 3, 2
    model.primary_src_dist, model.scatter_src_dist = 1000.0, 1000.0
    model.primary_src_size, model.scatter_src_size = 0.0, 30.0
    assert lib.dh_photon_terma(fake, fake, fake, g, ctypes.byref(model), ctypes.byref(beam), None) == N.DH_ERR_ARG
    assert "must be positive" in N.last_error()
    model.primary_src_size = 1.0
    model.kernel[2][5] = float("nan")
    assert lib.dh_photon_ccc(fake, fake, fake, fake, g, ctypes.byref(model), ctypes.byref(beam), 1.0, 0,
                             None) == N.DH_ERR_ARG
    assert "cone 2" in N.last_error() and "not a number" in N.last_error()
    model.kernel[2][5] = 1.0e6   # 5e6 steps of 0.2 cm
    assert lib.dh_photon_ccc(fake, fake, fake, fake, g, ctypes.byref(model), ctypes.byref(beam), 1.0, 0,
                             None) == N.DH_ERR_ARG
    assert "beyond 4096 steps" in N.last_error()
