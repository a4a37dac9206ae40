"""CPU checks of the sampled-voxel oracle helper against the full-volume oracle.

The big-plan GPU parity tests compare with ``oracle_sample.OraclePlan``, which
marches only the rays a sampled voxel's smoothed WET needs.  Here, on small
plans, it must equal the full-volume oracle (``O.compute_plan``, ``O.wet``,
``O.pb_grad``) at the sample, its gradient must match central finite
differences of its own dose, and ``gamma_at`` over every voxel must equal
``O.gamma_pass_rate``.
"""

import numpy as np
import pytest

from oracle import impt_oracle as O
from oracle_sample import OraclePlan, gamma_at, rel_err, sample_voxels


def _plan(n_fx=2):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    shape = (18, 22, 20)
    g = _grid(shape, 2.0)
    rng = np.random.default_rng(21)
    hu = rng.normal(40.0, 20.0, size=shape).astype(np.float32)
    hu[:, 4:7, :] = 1100.0
    hu[5:9, 10:14, 3:8] = -1000.0
    g.HU = hu
    plan = IMPTPlan()
    plan.n_fractions = n_fx
    c = np.arange(-6.0, 6.1, 3.0)
    for ga, couch in ((120.0, 0.0), (240.0, 30.0)):
        sl = grid_spots([18, 24], c, c)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=couch))
    return g, plan


def test_sampled_oracle_equals_full_oracle(machine_dir):
    g, plan = _plan()
    ref, beams, wets = O.compute_plan(g, plan, machine_dir, return_wet=True)
    idx = sample_voxels(ref, 60, 60, seed=3)
    op = OraclePlan(g, plan, machine_dir, idx)
    np.testing.assert_allclose(op.dose(), ref.reshape(-1)[idx], rtol=1e-12, atol=1e-14 * ref.max())
    for k, w in enumerate(wets):
        np.testing.assert_allclose(op.wet[k], w.reshape(-1)[idx], rtol=0, atol=1e-12)
        np.testing.assert_allclose(op.beam_dose(k) * plan.n_fractions, beams[k].reshape(-1)[idx], rtol=1e-12,
                                   atol=1e-14 * ref.max())
    assert int(np.argmax(ref)) in set(idx.tolist())


def test_sampled_oracle_gradient(machine_dir):
    g, plan = _plan()
    ref, _, wets = O.compute_plan(g, plan, machine_dir, return_wet=True)
    idx = sample_voxels(ref, 40, 40, seed=5)
    op = OraclePlan(g, plan, machine_dir, idx)
    gv = np.random.default_rng(8).normal(size=idx.size)
    got = op.grad(gv)
    # the full-WET oracle gradient with the same sparse dL/dD
    s = float(np.float32(g.spacing[0]))
    gvol = np.zeros(ref.shape)
    gvol.reshape(-1)[idx] = gv * plan.n_fractions
    for k, ob in enumerate(op.beams):
        order, spots, layers, r80 = ob.spots()
        gm, gxy = O.pb_grad(wets[k], gvol, s, ob.geom, spots, layers, r80, ob.marr)
        np.testing.assert_allclose(got[k][0][order], gm, rtol=1e-12, atol=1e-14 * np.abs(gm).max())
        np.testing.assert_allclose(got[k][1][order], gxy, rtol=1e-12, atol=1e-14 * np.abs(gxy).max())
    # central differences of L = sum gv * dose at the sample, beam 0, a few spots
    mu0 = op.beams[0].spot_list[:, 2].astype(np.float64)
    xy0 = op.beams[0].spot_list[:, :2].astype(np.float64)

    def loss(mu, xy):
        return float(np.dot(gv, op.dose([mu, None], [xy, None])))

    # the dose is linear in the weights and spot_list is float32: a step of 0.125 keeps mu +- h exact
    for s_ in (0, 7, 19):
        e = np.zeros_like(mu0)
        e[s_] = 0.125
        fd = (loss(mu0 + e, xy0) - loss(mu0 - e, xy0)) / 0.25
        assert abs(fd - got[0][0][s_]) <= 1e-6 * max(1.0, abs(fd))
        for a in (0, 1):
            d = np.zeros_like(xy0)
            d[s_, a] = 0.05
            fdx = (loss(mu0, xy0 + d) - loss(mu0, xy0 - d)) / 0.1
            assert abs(fdx - got[0][1][s_, a]) <= 2e-3 * np.abs(got[0][1]).max()


@pytest.mark.parametrize("spacing", [2.0, 1.0, 2.5])
def test_gamma_at_every_voxel_equals_full_gamma(spacing):
    rng = np.random.default_rng(int(spacing * 10))
    shape = (12, 14, 13)
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    ref = np.exp(-((x - 6.0) ** 2 + (y - 7.0) ** 2 + (z - 5.0) ** 2) / 18.0)
    test = ref * (1.0 + rng.normal(0.0, 0.03, size=shape)) + np.roll(ref, 1, axis=2) * 0.02
    idx = np.arange(ref.size)
    want = O.gamma_pass_rate(test, ref, spacing)
    got = gamma_at(ref.reshape(-1), idx, test, spacing)
    assert got == want
    assert 0.0 < want < 1.0


def test_sample_voxels_and_rel_err():
    d = np.zeros((5, 6, 7))
    d[2, 3, 4] = 10.0
    d[1, 1, 1] = 2.0
    idx = sample_voxels(d, 5, 20, seed=1, extra=[3])
    assert np.all(np.diff(idx) > 0) and idx.min() >= 0 and idx.max() < d.size
    assert np.ravel_multi_index((2, 3, 4), d.shape) in idx and 3 in idx
    assert rel_err([1.0, 2.1], [1.0, 2.0]) == pytest.approx(0.05)
    assert rel_err([0.0], [0.0]) == 0.0
