"""Pins the oracle's restatement of ReconDBT's SART and ML loops
(oracle/recondbt.py; ReconDBT/SART_dbt.m, ML_dbt.m) by what the algorithms fix:

* consistent data make the true image a fixed point of both (exactly);
* SART from a zero image lowers the data residual ||p - Gx|| every iteration;
* ML lowers its cost sum(y + proj * drr) (ML_dbt.m:98) every iteration.
"""
import numpy as np

from newrecon_amd.geometry import Geometry, arc_angles
from oracle import recondbt as R


def geom():
    return Geometry.half_cone(8, 6, 8, 2.0, 2.0, 1.0, 24, 14, 1.5, 1.5,
                              source_to_detector=60.0)


def phantom(g, seed):
    rng = np.random.default_rng(seed)
    return rng.random(g.vol_shape) * 0.02 + 0.01


def test_sart_fixed_point(oracle_lib):
    g, ang = geom(), arc_angles(4, 7.5)
    x = phantom(g, 1)
    p = np.concatenate([R._G(g, ang, x, v) for v in range(4)])
    got = R.sart_dbt(g, ang, p, x, 2, 0.7)
    np.testing.assert_array_equal(got, x)


def test_ml_fixed_point(oracle_lib):
    g, ang = geom(), arc_angles(4, 7.5)
    x = phantom(g, 2)
    counts = np.concatenate([1e4 * np.exp(-R._G(g, ang, x, v)) for v in range(4)])
    got = R.ml_dbt(g, ang, counts, x, 1e4, 2, 1.0)
    np.testing.assert_array_equal(got, x)


def test_sart_residual_falls(oracle_lib):
    g, ang = geom(), arc_angles(4, 7.5)
    x = phantom(g, 3)
    p = np.concatenate([R._G(g, ang, x, v) for v in range(4)])
    rows = np.concatenate([oracle_lib.gamma(g, ang, v, v + 1, 1, dtype=np.float64)
                           for v in range(4)])
    assert (rows > 5 * g.dz).any() and (rows[rows > 0] <= 5 * g.dz).any()  # both sides of the mask
    snaps = R.sart_dbt(g, ang, p, np.zeros_like(x), 4, 0.5, saveiter=True)
    res = [np.linalg.norm(np.concatenate([R._G(g, ang, s, v) for v in range(4)]) - p)
           for s in snaps]
    assert all(b < a for a, b in zip(res, res[1:])), res
    assert res[-1] < 0.5 * np.linalg.norm(p)


def test_ml_cost_falls(oracle_lib):
    g, ang = geom(), arc_angles(4, 7.5)
    x = phantom(g, 4)
    rng = np.random.default_rng(5)
    mean = np.concatenate([2e3 * np.exp(-R._G(g, ang, x, v)) for v in range(4)])
    counts = rng.poisson(mean).astype(np.float64)
    x0 = np.full_like(x, 0.01)
    snaps = R.ml_dbt(g, ang, counts, x0, 2e3, 4, 1.0, saveiter=True)
    cost = [R.ml_cost(g, ang, counts, s, 2e3) for s in [x0, *snaps]]
    assert all(b < a for a, b in zip(cost, cost[1:])), cost
