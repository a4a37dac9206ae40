"""Plan API and drop-in argument protocol (CPU; no kernel launches)."""

import numpy as np
import pytest

from DoseCUDA import IMPTDoseGrid, IMPTPlan, IMPTBeam, VolumeObject
from DoseCUDA import dose_kernels


def test_cube_phantom_matches_reference_layout():
    g = IMPTDoseGrid()
    g.createCubePhantom()
    assert g.HU.shape == (138, 138, 138) and g.HU.dtype == np.float32
    np.testing.assert_array_equal(g.origin, [-207.0, -207.0, -207.0])
    np.testing.assert_array_equal(g.spacing, [3.0, 3.0, 3.0])
    m = round(10.0 / 3.0)
    assert np.all(g.HU[:m] == -1000) and np.all(g.HU[-m:] == -1000)
    assert np.all(g.HU[m:-m, m:-m, m:-m] == 0)
    assert g.HU[m - 1, 50, 50] == -1000 and g.HU[m, m, m] == 0
    g2 = IMPTDoseGrid()
    g2.createCubePhantom(size=[4, 4, 4], spacing=25.0)  # margin rounds to 0: all air, as the reference slices
    assert np.all(g2.HU == -1000)


def test_beam_spot_helpers():
    b = IMPTBeam()
    b.addSingleSpot(1.0, 2.0, 0.5, 3)
    assert b.spot_list.ndim == 1 and b.n_spots == 1        # reference quirk kept
    b.addSingleSpot(-1.0, 0.0, 0.25, 4)
    assert b.spot_list.shape == (2, 4) and b.spot_list.dtype == np.float32
    b.changeSpotEnergy(7)
    assert np.all(b.spot_list[:, 3] == 7)

    class CP:
        ScanSpotPositionMap = [0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
        ScanSpotMetersetWeights = [0.1, 0.2, 0.3]

    b.addSpotData(CP(), 9)
    assert b.spot_list.shape == (5, 4)
    np.testing.assert_allclose(b.spot_list[2:], [[0, 1, 0.1, 9], [2, 3, 0.2, 9], [4, 5, 0.3, 9]], rtol=1e-7)
    c = IMPTBeam(b)
    assert c.n_spots == 5 and c.spot_list is b.spot_list
    b.resetSpots()
    assert b.spot_list.shape == (0, 4) and b.n_spots == 0


def test_plan_models_and_energy_lookup():
    p = IMPTPlan()
    assert len(p.beam_models) == 2 and p.machine_name == "HitachiProbeatJHU"
    m = p.beam_models[0]
    assert m.energyIDFromLabel(70.2) == 0
    with pytest.raises(Exception):
        m.energyIDFromLabel(12.345)


def test_isotropy_check():
    g = IMPTDoseGrid()
    g.createCubePhantom(size=[8, 8, 8], spacing=3.0)
    g.spacing = np.array([3.0, 3.0, 2.0])
    with pytest.raises(Exception, match="isotropic"):
        g.computeIMPTPlan(IMPTPlan())


class _Obj:
    pass


def _args():
    p = IMPTPlan()
    model = p.beam_models[0]
    vol = VolumeObject()
    vol.voxel_data = np.zeros((4, 5, 6), np.float32)
    vol.spacing = np.array([2.0, 2.0, 2.0], np.float32)
    vol.origin = np.array([-6.0, -5.0, -4.0], np.float32)
    beam = IMPTBeam()
    beam.spot_list = np.array([[0, 0, 1, 3], [1, 1, 1, 3]], np.float32)
    beam.n_spots = 2
    wet = VolumeObject()
    wet.voxel_data = np.zeros((4, 5, 6), np.float32)
    return model, vol, wet, beam


def test_dropin_attribute_errors():
    model, vol, wet, beam = _args()
    vol.voxel_data = np.zeros((4, 5, 6), np.float64)
    with pytest.raises(ValueError, match="'voxel_data' must be 3-dimensional and of type float."):
        dose_kernels.proton_raytrace_cuda(model, vol, beam, 0)
    model, vol, wet, beam = _args()
    beam.iso = np.zeros((1, 3), np.float32)
    with pytest.raises(ValueError, match="'iso' must be 1-dimensional and of type float."):
        dose_kernels.proton_raytrace_cuda(model, vol, beam, 0)
    model, vol, wet, beam = _args()
    bad = _Obj()
    with pytest.raises(AttributeError):
        dose_kernels.proton_raytrace_cuda(bad, vol, beam, 0)
    model, vol, wet, beam = _args()
    beam.gantry_angle = "90"
    with pytest.raises(TypeError):
        dose_kernels.proton_raytrace_cuda(model, vol, beam, 0)
    model, vol, wet, beam = _args()
    beam.spot_list = np.array([0, 0, 1, 3], np.float32)   # a lone addSingleSpot row
    with pytest.raises(ValueError, match="'spot_list' must be 2-dimensional and of type float."):
        dose_kernels.proton_spot_cuda(model, vol, wet, beam, 0)
    model, vol, wet, beam = _args()
    wet.voxel_data = np.zeros((4, 5), np.float32)
    with pytest.raises(ValueError, match="'voxel_data' must be 3-dimensional and of type float."):
        dose_kernels.proton_spot_cuda(model, vol, wet, beam, 0)


def test_no_cpu_fallback():
    """Without a HIP device the drop-in refuses to run (it never computes on the CPU)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    model, vol, wet, beam = _args()
    with pytest.raises(RuntimeError, match="CUDA error"):
        dose_kernels.proton_raytrace_cuda(model, vol, beam, 0)
