"""Gradients of the 10k-spot target plan (the differentiation workload) vs central
finite differences (GPU).

Plan: 100^3 water phantom at 2.5 mm with a 20 mm HU = 1000 slab, gantry 0,
25 layers x 400 spots (10000 spots), 50 mm spherical target at the isocentre,
loss L = mean_target (D - 2)^2 + 0.1 mean_shell D^2 (10 mm shell).

dL/dmu and dL/d(x, y) come from the hand-written adjoint through
torch.autograd (ProtonDose).  The finite differences perturb one spot at a
time: the float64 oracle computes that spot's dose at the perturbed weight or
position on the oracle's float64 WET, and L is re-evaluated in float64 on the
GPU plan dose with the spot's term swapped (central differences, h = 1e-3 in
weight, 0.05 mm in position).  Bar: max |grad - FD| <= 1e-3 x max |FD| over
the checked spots, for weights and for positions.
"""

import numpy as np
import pytest
import torch

from oracle import impt_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-grad] " + msg, flush=True)


def test_target_plan_gradients_vs_central_fd():
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import target_plan
    dev = torch.device("cuda", 0)
    g, plan, target, shell = target_plan()
    beam = plan.beam_list[0]
    model = plan.beam_models[0]
    assert beam.spot_list.shape[0] == 10000
    eng = IMPTDoseEngine(g, plan, cutoff_sigma=6.0)
    params = eng.parameters()
    dose = eng.differentiable_dose(params)
    tm = torch.from_numpy(target).to(dev)
    sm = torch.from_numpy(shell).to(dev)
    loss = ((dose[tm] - 2.0) ** 2).mean() + 0.1 * (dose[sm] ** 2).mean()
    loss.backward()
    gmu = params[0][0].grad.cpu().numpy().astype(np.float64)
    gxy = params[0][1].grad.cpu().numpy().astype(np.float64)
    D0 = dose.detach().cpu().numpy().astype(np.float64)
    n_t, n_s = float(target.sum()), float(shell.sum())

    def L(D):
        return float(np.sum((D[target] - 2.0) ** 2) / n_t + 0.1 * np.sum(D[shell] ** 2) / n_s)

    rlsp = g.RLSPFromHU(plan.machine_name)
    s = float(np.float32(g.spacing[0]))
    geom = O.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32), beam.gantry_angle,
                           beam.couch_angle, model.VSADX, model.VSADY)
    wet_o = O.wet(rlsp, s, geom)
    marr = O.ModelArrays(model)
    sl = np.asarray(beam.spot_list, dtype=np.float32)

    def spot_dose(x, y, e):
        one = np.array([[x, y, 1.0, e]], dtype=np.float32)
        _, spots, layers, r80 = O.spots_and_layers(one, model.divergence_params)
        return O.pb_dose(wet_o, s, geom, spots, layers, r80, marr)

    # spots with the largest gradients plus a spread of others
    rng = np.random.default_rng(0)
    order = np.argsort(-np.abs(gmu))
    pick = np.unique(np.concatenate([order[:10], rng.choice(10000, size=6, replace=False)]))
    hw, hx = 1e-3, 0.05
    fd_mu, fd_xy, got_mu, got_xy = [], [], [], []
    for k in pick:
        x, y, mu, e = float(sl[k, 0]), float(sl[k, 1]), float(sl[k, 2]), int(sl[k, 3])
        K = spot_dose(x, y, e)
        fd_mu.append((L(D0 + hw * K) - L(D0 - hw * K)) / (2 * hw))
        got_mu.append(gmu[k])
        base = D0 - mu * K
        for a in range(2):
            dp = [x, y]
            dm = [x, y]
            dp[a] += hx
            dm[a] -= hx
            fd_xy.append((L(base + mu * spot_dose(dp[0], dp[1], e)) - L(base + mu * spot_dose(dm[0], dm[1], e)))
                         / (2 * hx))
            got_xy.append(gxy[k, a])
    fd_mu, fd_xy = np.array(fd_mu), np.array(fd_xy)
    e_mu = np.abs(np.array(got_mu) - fd_mu).max() / np.abs(fd_mu).max()
    e_xy = np.abs(np.array(got_xy) - fd_xy).max() / np.abs(fd_xy).max()
    _say("spots %d, loss %.5f, rel err dL/dmu %.2e, dL/dxy %.2e" % (len(pick), float(loss.detach()), e_mu, e_xy))
    assert e_mu <= 1e-3 and e_xy <= 1e-3
