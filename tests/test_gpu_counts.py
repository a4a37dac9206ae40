"""Work counts of dh_pb_count_pairs (GPU), which the throughput figure of
bench.py is built on: the reference's pairs per (voxel, layer) and the pairs
the culled kernels evaluate.

Dense counts have a closed form from the WET volume and the layer table
(importLayers, IMPTClasses.cu:29-67; the 1.1 R80 skip, IMPTClasses.cu:224-226):
forward = sum over layers of n_spots x #(voxels with 10 WET <= 1.1 R80),
adjoint = the same over voxels with dL/dD != 0, in the kernels' float32
arithmetic.  Without a cutoff every such pair is evaluated; with one the
evaluated counts lie in (0, dense].
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cutoff", [6.0, 0.0])
def test_count_pairs_closed_form(cutoff):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from DoseCUDA import device as D
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    from DoseCUDA import IMPTPlan
    print("[gpu-counts] cutoff %g" % cutoff, flush=True)
    g = _grid((36, 44, 40), 2.5)
    rng = np.random.default_rng(21)
    g.HU = rng.uniform(-150.0, 350.0, size=(36, 44, 40)).astype(np.float32)
    plan = IMPTPlan()
    c = np.arange(-12.0, 12.0 + 1e-6, 3.0)
    for ga, ca in ((0.0, 0.0), (135.0, 20.0)):
        sl = grid_spots([19, 25, 30, 34], c, c, disk_radius=12.0)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=ca))
    eng = IMPTDoseEngine(g, plan, cutoff_sigma=cutoff)
    eng.compute_wet(force=True)
    gd = rng.normal(size=eng.shape).astype(np.float32)
    gd[rng.uniform(size=eng.shape) < 0.4] = 0.0
    gd_t = torch.from_numpy(gd).to(eng.device)
    for b in eng.beams:
        counts = D.pb_count_pairs(b.wet, gd_t, eng.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                                  cutoff_sigma=cutoff)
        got = [int(v) for v in counts.cpu().tolist()]
        wet_mm = b.wet.cpu().numpy().astype(np.float32) * np.float32(10.0)
        tab = b.spots.layers_np
        r80 = b.spots.layers.cpu().numpy()[:, 3].astype(np.float32)
        lim = np.float32(1.1) * r80
        fwd = sum(int(tab[l, 1]) * int(np.count_nonzero(wet_mm <= lim[l])) for l in range(len(tab)))
        adj = sum(int(tab[l, 1]) * int(np.count_nonzero((wet_mm <= lim[l]) & (gd != 0.0)))
                  for l in range(len(tab)))
        print("[gpu-counts] counts %s, closed form %d / %d" % (got, fwd, adj), flush=True)
        assert got[0] == fwd and got[2] == adj and fwd > 0 and adj > 0
        if cutoff > 0.0:
            assert 0 < got[1] <= got[0] and 0 < got[3] <= got[2]
        else:
            assert got[1] == got[0] and got[3] == got[2]
