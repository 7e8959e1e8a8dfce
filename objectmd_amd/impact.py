"""Impact replicas: Si(001) slabs hit by one Ar atom each.

Per replica this is the set-up of test/si-impact/main.cpp:17-80 and its
omd-parameter: a diamond (001) slab whose top layer lies at z = 0 with the
box a/8 outside the outer layers (as test/si-impact/atom.dat:7), the
projectile of include/crystal/Projectile.h:22-49 aimed at a surface point
from z = 5 A, the bottom layers fixed (CoordClamp), a damped border
(ForceDamper), the projectile range-limited (RangeLimiter), Stillinger-Weber
among Si and the Ar-Si ZBL pair table between projectile and target, and
DynamicTimeStep with maxpath 0.13 A, dt <= 1 fs, update every 5 steps
(omd-parameter:27-29).

Replicas of uneven size are concatenated into one atom set for
``Engine.set_replicas``; each keeps its own box and periodicity.
"""
import math
import os

import numpy as np

from . import FLAG_ACTIVE, FLAG_DAMP, FLAG_FIXED, FLAG_LIMIT, lattice, tables
from .params import sw_params

A_SI = 5.431
MAXPATH = 0.13          # omd-parameter:27
DT_MAX = 1e-3           # omd-parameter:28 (ps)
UPDATE = 5              # omd-parameter:29
DAMP_FACTOR = 0.34328   # omd-parameter:24
RANGE_BORDER = 1000.0   # main.cpp:76
PROJ_HEIGHT = 5.0       # omd-parameter:3


def projectile(tx, ty, h, theta, phi, energy, mass):
    """Position and velocity of Projectile (include/crystal/Projectile.h:
    30-47): speed from the energy, polar angle theta from -z, azimuth phi
    (degrees); the start point is moved back along the path so that the
    line passes (tx, ty) at z = 0."""
    spi = math.pi / 180.0
    vv = math.sqrt(2.0 * energy / mass)
    vxy = vv * math.sin(theta * spi)
    v = np.array([vxy * math.cos(phi * spi), vxy * math.sin(phi * spi),
                  -vv * math.cos(theta * spi)])
    pxy = h * math.tan(theta * spi)
    x = np.array([tx - pxy * math.cos(phi * spi), ty - pxy * math.sin(phi * spi), h])
    return x, v


def slab_shape(n_target):
    """Unit cells (nx = ny, nz) of a roughly cubic slab of about n_target atoms."""
    cells = max(n_target / 8.0, 27.0)
    k = max(3, int(round(cells ** (1.0 / 3.0))))
    nz = max(3, int(round(cells / (k * k))))
    return k, k, nz


def replica(n_target, energy, theta, phi, seed, temperature=300.0, a=A_SI,
            fixed_layers=2, damp_width=3.0):
    """One impact system: dict with x, v, tid, flag, lo, hi, pbc (PBC x, y)."""
    nx, ny, nz = slab_shape(n_target)
    pos, (lo, hi) = lattice.diamond100(nx, ny, nz, a)
    rng = np.random.default_rng(seed)
    m_si = tables.mass_sim(tables.MASS_SI)
    m_ar = tables.mass_sim(tables.MASS_AR)
    v = lattice.maxwell_boltzmann(len(pos), m_si, temperature, int(rng.integers(1 << 31)))
    zmin = pos[:, 2].min()
    flag = np.full(len(pos), FLAG_ACTIVE, np.int32)
    # bottom monolayers (a/4 apart) fixed on every axis
    fixed = pos[:, 2] < zmin + (fixed_layers - 0.5) * 0.25 * a
    flag[fixed] |= FLAG_FIXED
    # damped border: within damp_width of the side faces or above the
    # fixed layers (the inverse box of main.cpp:57-62)
    damp = ((pos[:, 0] < lo[0] + damp_width) | (pos[:, 0] > hi[0] - damp_width) |
            (pos[:, 1] < lo[1] + damp_width) | (pos[:, 1] > hi[1] - damp_width) |
            (pos[:, 2] < lo[2] + damp_width + fixed_layers * 0.25 * a))
    flag[damp] |= FLAG_DAMP
    # aim at a random surface point away from the damped border
    tx = rng.uniform(lo[0] + damp_width, hi[0] - damp_width)
    ty = rng.uniform(lo[1] + damp_width, hi[1] - damp_width)
    xp, vp = projectile(tx, ty, PROJ_HEIGHT, theta, phi, energy, m_ar)
    x = np.vstack([pos, xp[None]])
    v = np.vstack([v, vp[None]])
    tid = np.zeros(len(x), np.int32)
    tid[-1] = 1
    flag = np.append(flag, FLAG_ACTIVE | FLAG_LIMIT).astype(np.int32)
    return dict(x=x, v=v, tid=tid, flag=flag, lo=lo, hi=hi, pbc=3,
                energy=energy, theta=theta, phi=phi)


def batch(nrep, seed, n_min=2000, n_max=20000, e_min=50.0, e_max=500.0, theta_max=60.0,
          temperature=300.0):
    """nrep replicas: sizes log-uniform in [n_min, n_max], energy uniform in
    [e_min, e_max] eV, polar angle uniform in [0, theta_max], azimuth uniform
    in [0, 360) degrees.  Returns the concatenated arrays, the replica offsets
    and boxes, and the per-replica draws."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), nrep)).round().astype(int)
    energy = rng.uniform(e_min, e_max, nrep)
    theta = rng.uniform(0.0, theta_max, nrep)
    phi = rng.uniform(0.0, 360.0, nrep)
    seeds = rng.integers(1 << 31, size=nrep)
    reps = [replica(int(sizes[r]), energy[r], theta[r], phi[r], int(seeds[r]), temperature)
            for r in range(nrep)]
    off = np.zeros(nrep + 1, np.int32)
    off[1:] = np.cumsum([len(p["x"]) for p in reps])
    return dict(
        x=np.vstack([p["x"] for p in reps]), v=np.vstack([p["v"] for p in reps]),
        tid=np.concatenate([p["tid"] for p in reps]),
        flag=np.concatenate([p["flag"] for p in reps]),
        off=off, lo=np.array([p["lo"] for p in reps]), hi=np.array([p["hi"] for p in reps]),
        pbc=np.array([p["pbc"] for p in reps], np.int32), energy=energy, theta=theta, phi=phi,
        mass=[tables.mass_sim(tables.MASS_SI), tables.mass_sim(tables.MASS_AR)])


def setup_engine(eng, b, sw_path, pair_path, skin=0.5, graph_steps=20, maxnbr=0):
    """Load a batch into an Engine: box of the whole set (unused by replicas
    but required), atoms, SW + Ar-Si table, modifies, replicas."""
    from . import POT_SW, TAB_PAIR
    eng.set_box(b["lo"].min(0), b["hi"].max(0), 3)
    eng.upload_atoms(b["x"], b["v"], b["tid"], b["flag"], b["mass"])
    eng.load_table(TAB_PAIR, 0, 1, os.path.abspath(pair_path), "")
    eng.set_sw(sw_params(sw_path))
    eng.set_potential(POT_SW)
    eng.set_neighbor(skin, graph_steps, maxnbr)
    eng.set_modifies(DT_MAX, MAXPATH, UPDATE, 7, DAMP_FACTOR, RANGE_BORDER)
    eng.set_replicas(b["off"], b["lo"], b["hi"], b["pbc"])
