"""HIP path vs the committed golden fixtures (GPU).  Reads only
tests/golden/*.npz: no oracle call.

Per fixture plan (one beam, 27 spots): the engine's raw and smoothed WET
(max |dWET| <= 2e-4 cm), the dose with the dense kernel (cutoff 0, max |dD| /
D_max <= 1e-4) and with the default 6-sigma cutoff (<= 1e-3), and the
gradients of L = sum(grad_field * dose) through torch.autograd (max-norm
relative error <= 1e-4 for weights and positions).
"""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["water_slab_g0", "hetero_g120_c30"]


def _say(msg):
    print("[gpu-golden] " + msg, flush=True)


def _plan(f):
    from DoseCUDA import IMPTDoseGrid, IMPTPlan
    from DoseCUDA.plan_impt import IMPTBeam
    g = IMPTDoseGrid()
    s = float(f["spacing"])
    g.spacing = np.array([s] * 3, dtype=np.float64)
    g.size = np.array(f["hu"].shape)
    g.origin = np.asarray(f["origin"], np.float64)
    g.HU = f["hu"]
    b = IMPTBeam()
    b.dicom_rangeshifter_label = "0"
    b.gantry_angle = float(f["gantry"])
    b.couch_angle = float(f["couch"])
    b.iso = np.asarray(f["iso"], np.float32)
    b.spot_list = f["spot_list"]
    b.n_spots = b.spot_list.shape[0]
    plan = IMPTPlan()
    plan.n_fractions = int(f["n_fractions"])
    plan.addBeam(b)
    return g, plan


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_engine_matches_golden(name, cutoff):
    from DoseCUDA.engine import IMPTDoseEngine
    f = dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))
    _say("%s cutoff %g" % (name, cutoff))
    g, plan = _plan(f)
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=cutoff)
    eng.compute_wet()
    b = eng.beams[0]
    d_raw = np.abs(b.wet_raw.cpu().numpy() - f["wet_raw"]).max()
    d_wet = np.abs(b.wet.cpu().numpy() - f["wet"]).max()
    params = eng.parameters()
    dose = eng.differentiable_dose(params)
    gf = torch.from_numpy(f["grad_field"].astype(np.float32)).to(dose.device)
    (dose * gf).sum().backward()
    D = dose.detach().cpu().numpy().astype(np.float64)
    dmax = f["dose"].max()
    e_dose = np.abs(D - f["dose"]).max() / dmax
    gmu = params[0][0].grad.cpu().numpy().astype(np.float64)
    gxy = params[0][1].grad.cpu().numpy().astype(np.float64)
    e_mu = np.abs(gmu - f["grad_mu"]).max() / np.abs(f["grad_mu"]).max()
    e_xy = np.abs(gxy - f["grad_xy"]).max() / np.abs(f["grad_xy"]).max()
    _say("raw %.2e wet %.2e cm, dose %.2e of D_max, grad mu %.2e xy %.2e" % (d_raw, d_wet, e_dose, e_mu, e_xy))
    assert d_raw <= 2e-4 and d_wet <= 2e-4
    assert e_dose <= (1e-4 if cutoff == 0.0 else 1e-3)
    assert e_mu <= 1e-4 and e_xy <= 1e-4
