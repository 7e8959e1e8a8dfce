"""Pins the oracle's velocity Verlet (integrator.cpp:225-259) and run driver."""
import numpy as np
import pytest

from objectmd_amd import tables
from oracle import oracle as O
from tests.common import lj_system


def _harmonic_table(k, r_eq, r0=0.5, r1=4.0, n=4000):
    dr = (r1 - r0) / (n - 1)
    r = r0 + dr * np.arange(n)
    return O.Spline(0.5 * k * (r - r_eq) ** 2, dr, r0)


def test_harmonic_dimer_closed_form():
    k, req, m = 2.0, 2.0, 0.01
    tab = _harmonic_table(k, req)
    x = np.array([[0.0, 0.0, 0.0], [2.1, 0.0, 0.0]])
    v = np.zeros((2, 3))
    dt, nsteps = 1e-3, 2000
    xe, ve, en = O.run(0, x, v, m, np.zeros(3), np.ones(3) * 100, 0, [tab], 4.0, 0.0, 1,
                       nsteps, dt)
    # FirstRun zeroes the forces (system.cpp:872): the first half kick is lost,
    # so the motion starts one half step late; compare with that closed form
    w = np.sqrt(2 * k / m)
    t = nsteps * dt
    sep = req + 0.1 * np.cos(w * (t - 0.5 * dt))
    assert xe[1, 0] - xe[0, 0] == pytest.approx(sep, abs=2e-3)
    # energy bounded (symplectic)
    e = en[:, 2]
    assert np.ptp(e) < 1e-3


def test_time_reversibility():
    # velocity Verlet with forces at the start is exactly reversible up to
    # rounding: n steps forward, flip v, n steps back returns to the start
    s = lj_system(cells=3)
    phi = O.Spline.from_table(tables.lj_table())
    x = s["x"].copy()
    v = s["v"].copy()
    nb = O.neighbors(x, s["lo"], s["L"], 7, 8.5 + 1.0)

    def forces(xx):
        return O.pair_forces(xx, s["lo"], s["L"], 7, nb, phi)[0]

    f = forces(x)
    for direction in (1, -1):
        if direction < 0:
            v = -v
        for _ in range(100):
            O.vv_kick_drift(x, v, f, s["mass"], 1e-3)
            f = forces(x)
            O.vv_kick(v, f, s["mass"], 1e-3)
    assert np.abs(x - s["x"]).max() < 1e-9
    assert np.abs(v + s["v"]).max() < 1e-9


def test_lj_energy_steps_are_cutoff_crossings():
    # the truncated LJ of config 0 (rc = 2.5 sigma, not shifted) changes the
    # energy by V(rc) whenever a pair crosses rc; between crossings the
    # energy is conserved
    s = lj_system(cells=4)
    phi = O.Spline.from_table(tables.lj_table())
    _, _, en = O.run(0, s["x"], s["v"], s["mass"], s["lo"], s["L"], 7, [phi], 8.5, 0.3, 5,
                     300, 1e-3)
    jumps = np.diff(en[:, 2])
    vrc = abs(tables.lj_energy(8.5))
    small = np.abs(jumps) < 0.25 * vrc
    assert np.abs(jumps[small]).max() < 5e-6
    assert np.allclose(np.abs(jumps[~small]) / vrc, np.round(np.abs(jumps[~small]) / vrc),
                       atol=0.1)


def test_eam_energy_conservation():
    from tests.common import eam_system, eam_tables
    s = eam_system(cells=4)
    t = eam_tables()
    sp = [O.Spline.from_table(t[k]) for k in ("PAIR", "EDENS", "EMBED")]
    _, _, en = O.run(1, s["x"], s["v"], s["mass"], s["lo"], s["L"], 7, sp, 5.5, 0.3, 5,
                     400, 2e-3)
    e = en[:, 2] / len(s["x"])
    # bounded O(dt^2) oscillation, no secular drift
    assert np.ptp(e[5:]) < 5e-5
    assert abs(e[-100:].mean() - e[5:105].mean()) < 5e-6
