"""The reference's other call sites of the dose path (GPU): the WET cache
(computeIMPTWET, computeIMPTPlanFromWET; plan_impt.py:208-257) and
writeWETNRRD (plan_impt.py:259-287) give exactly what computeIMPTPlan
computes beam by beam."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_wet_cache_and_wet_nrrd(tmp_path):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from DoseCUDA import IMPTPlan
    from DoseCUDA.imageio import read_nrrd
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    print("[gpu-dropin] WET cache, plan from WET, WET NRRD", flush=True)
    g = _grid((30, 36, 32), 2.5)
    rng = np.random.default_rng(4)
    g.HU = rng.uniform(-200.0, 400.0, size=(30, 36, 32)).astype(np.float32)
    plan = IMPTPlan()
    plan.n_fractions = 2
    c = np.arange(-9.0, 9.0 + 1e-6, 3.0)
    for ga, ca in ((0.0, 0.0), (110.0, 20.0)):
        sl = grid_spots([18, 26], c, c, disk_radius=9.0)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=ca))

    g.computeIMPTPlan(plan)
    dose, beam_doses = g.dose.copy(), [d.copy() for d in g.beam_doses]
    assert float(dose.max()) > 0.0
    wets = g.computeIMPTWET(plan)
    assert len(wets) == 2
    total = np.zeros(g.size, dtype=np.single)
    for k, (w, beam) in enumerate(zip(wets, plan.beam_list)):
        d = g.computeIMPTPlanFromWET(plan, w, beam)
        assert np.array_equal(d * plan.n_fractions, beam_doses[k]), "beam %d" % k
        total += d
    total *= plan.n_fractions
    assert np.array_equal(total, dose)

    g.writeWETNRRD(plan, str(tmp_path / "wet.nrrd"))
    for k, w in enumerate(wets):
        vol, origin, spacing = read_nrrd(str(tmp_path / ("wet_beam%02d.nrrd" % (k + 1))))
        assert np.array_equal(vol, w.voxel_data)
        assert np.allclose(origin, g.origin) and np.allclose(spacing, g.spacing)
