"""Pins the oracle's impact driver (oracle.impact_run) with closed forms:
free flight under DynamicTimeStep, the FirstRun zero-force first step, the
order of MeasureKinetic and RangeLimiter, the CoordClamp axis mask, the
ForceDamper decay and the SW / pair-table split of impact_forces."""
import math
import os

import numpy as np
import pytest

from objectmd_amd import impact, tables
from oracle import oracle as O
from tests.common import GOLDEN

M_SI = tables.mass_sim(tables.MASS_SI)
M_AR = tables.mass_sim(tables.MASS_AR)
BIG = np.array([-50.0, -50.0, -50.0]), np.array([50.0, 50.0, 50.0])


@pytest.fixture(scope="module")
def pot():
    y, dr, r0, rm = O.table_load(os.path.join(GOLDEN, "ref", "pair.ar-si.zbl"), "")
    p = impact.sw_params(os.path.join(GOLDEN, "ref", "sw.silicon"))
    return O.SwParam(*p[:10], int(p[10]), *p[11:]), O.Spline(y, dr, r0, rm)


def run(pot, x, v, tid, flag, nsteps, dt_max=1e-3, maxpath=0.0, update=5, clamp=7,
        damp=0.0, border=-1.0, box=BIG):
    swp, pair = pot
    return O.impact_run(np.array(x, float), np.array(v, float), np.array(tid), np.array(flag),
                        [M_SI, M_AR], box[0], box[1], 0, swp, pair, nsteps, dt_max, maxpath,
                        update, clamp, damp, border)


def test_free_flight_dynamic_dt(pot):
    """A lone atom keeps its speed: dt = maxpath / (1.1 |v|), x = x0 + v t."""
    v0 = np.array([300.0, -200.0, 100.0])
    r = run(pot, [[0.0, 0.0, 0.0]], [v0], [1], [1], 20, maxpath=0.13, update=5)
    dt = 0.13 / (np.linalg.norm(v0) * 1.10)
    assert np.allclose(r["dts"], dt, rtol=1e-14, atol=0)
    assert r["time"] == pytest.approx(20 * dt, rel=1e-13)
    assert np.allclose(r["x"][0], v0 * r["time"], rtol=1e-12, atol=1e-12)
    # a slow atom takes dt_max
    r = run(pot, [[0.0, 0.0, 0.0]], [[1.0, 0.0, 0.0]], [1], [1], 7, maxpath=0.13)
    assert np.all(r["dts"] == 1e-3)


def test_first_step_has_zero_force(pot):
    """FirstRun zeroes the forces, so the first drift moves nothing from rest
    and the first step ends with v = (dt/2) f(x0) / m."""
    x0 = np.array([[0.0, 0.0, 0.0], [2.1, 0.0, 0.0]])
    dt = 5e-4
    r = run(pot, x0, np.zeros((2, 3)), [0, 0], [1, 1], 1, dt_max=dt)
    assert np.array_equal(r["x"], x0)
    f0, _ = O.impact_forces(x0, [0, 0], [1, 1], BIG[0], BIG[1] - BIG[0], 0, *pot)
    assert abs(f0[0, 0]) > 1.0
    assert np.allclose(r["v"], 0.5 * dt * f0 / M_SI, rtol=1e-14, atol=0)


def test_range_limiter_after_measure_kinetic(pot):
    """The step that removes an atom still counts its speed for the next dt;
    once inactive it stops moving."""
    lo, hi = np.array([-5.0, -5.0, -5.0]), np.array([5.0, 5.0, 5.0])
    border = 2.0
    vf = 400.0
    x = [[6.0, 0.0, 0.0], [-3.0, -3.0, -3.0]]
    v = [[vf, 0.0, 0.0], [5.0, 0.0, 0.0]]
    r = run(pot, x, v, [1, 1], [1 | 32, 1], 12, maxpath=0.13, update=1, border=border,
            box=(lo, hi))
    d_fast = 0.13 / (vf * 1.10)
    # first step whose end position lies beyond hi + border
    k = next(s for s in range(1, 100) if 6.0 + s * d_fast * vf > hi[0] + border)
    expect = [d_fast] * (k + 1) + [1e-3] * (12 - k - 1)
    assert np.allclose(r["dts"], expect, rtol=1e-14)
    assert r["flag"][0] == 32
    assert r["x"][0, 0] == pytest.approx(6.0 + k * d_fast * vf, rel=1e-12)


def test_clamp_axis_mask(pot):
    """CoordClamp zeroes f and v on the masked axes only; FirstRun already
    stops the clamped motion.  Mask 6 is the reference's parse of "xyz"."""
    v0 = [[2.0, 3.0, 4.0]]
    r = run(pot, [[0.0, 0.0, 0.0]], v0, [1], [1 | 8], 10, dt_max=1e-3, clamp=6)
    assert r["x"][0, 0] == pytest.approx(10 * 1e-3 * 2.0, rel=1e-12)
    assert r["x"][0, 1] == 0.0 and r["x"][0, 2] == 0.0
    assert np.array_equal(r["v"][0], [2.0, 0.0, 0.0])


def test_damper_decay(pot):
    """f = -gamma v on a free atom: |v| decays as exp(-gamma t / m)."""
    gamma, dt, n = 0.34328, 1e-4, 200
    r = run(pot, [[0.0, 0.0, 0.0]], [[10.0, 0.0, 0.0]], [1], [1 | 16], n, dt_max=dt,
            damp=gamma)
    ratio = r["v"][0, 0] / 10.0
    expect = math.exp(-gamma / M_AR * n * dt)
    assert 0.1 < expect < 0.3
    assert ratio == pytest.approx(expect, rel=0.015)


def test_force_split_sw_and_table(pot):
    """Ar-Si pairs use the table, Si-Si pairs Stillinger-Weber, Ar-Ar none."""
    swp, pair = pot
    L = BIG[1] - BIG[0]
    for r_ in (2.2, 2.9, 3.4):
        f, e = O.impact_forces([[0, 0, 0], [r_, 0, 0]], [1, 0], [1, 1], BIG[0], L, 0, swp,
                               pair)
        rc, phi, dphi = pair.read(r_)
        assert e.sum() == pytest.approx(phi, rel=1e-12)
        assert f[0, 0] == pytest.approx(dphi, rel=1e-10)
        f, e = O.impact_forces([[0, 0, 0], [r_, 0, 0]], [1, 1], [1, 1], BIG[0], L, 0, swp,
                               pair)
        assert e.sum() == 0.0 and np.all(f == 0.0)
    # Si-Si dimer outside the ZBL splice: the SW two-body term
    s = swp
    for r_ in (2.2, 2.6, 3.2):
        _, e = O.impact_forces([[0, 0, 0], [r_, 0, 0]], [0, 0], [1, 1], BIG[0], L, 0, swp,
                               pair)
        v2 = s.A * s.eps * (s.B * (s.sigma / r_) ** s.p - (s.sigma / r_) ** s.q) * \
            math.exp(s.sigma / (r_ - s.a * s.sigma))
        assert e.sum() == pytest.approx(v2, rel=1e-12)


def test_impact_energy_conservation(pot):
    """Without modifies a slow projectile run conserves Ek + Ep."""
    b = impact.replica(300, 5.0, 10.0, 20.0, 3)
    swp, pair = pot
    flag = np.ones(len(b["x"]), np.int32)
    e = []
    for n in (0, 150, 300):
        r = O.impact_run(b["x"], b["v"], b["tid"], flag, [M_SI, M_AR], b["lo"], b["hi"], 3,
                         swp, pair, n, 2e-4, 0.0, 5, 7, 0.0, -1.0)
        if n:
            e.append(r["ek"] + r["ep"])
    ek0 = 0.5 * (np.array([M_SI, M_AR])[b["tid"]][:, None] * b["v"] ** 2).sum()
    assert abs(e[1] - e[0]) < 2e-4 * ek0
