"""Crystal builders and initial velocities for synthetic inputs.

fcc100 and diamond100 produce the atoms in the order of the reference's
builders (FCC::Create100, src/gadget/FCC.cpp:142-242; Diamond::Create100,
include/crystal/Diamond.h:17-58) and the box of Crystalline::CalcBox
(include/crystal/Crystalline.h:53-74): the bounding box widened by half a
monolayer on every side.
"""
import numpy as np

from .tables import MD_BOLTZMANN


def fcc100(cells_x, cells_y, cells_z, a):
    """fcc crystal of cells_* unit cells (FCC("100", 2*cells_x, ...)).
    Returns positions (n,3) and box (lo[3], hi[3])."""
    h = a / 2.0
    xml, yml, zml = 2 * cells_x, 2 * cells_y, 2 * cells_z
    x1 = h * (xml - 1.0)
    y1 = h * (yml - 1.0)
    z0 = 0.0 - h * (zml - 1.0)
    rel = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]], np.float64) * h
    nz, ny, nx = (zml + 1) // 2, (yml + 1) // 2, (xml + 1) // 2
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    org = np.stack([i.ravel() * (2 * h), j.ravel() * (2 * h), k.ravel() * (2 * h) + z0], 1)
    pos = (org[:, None, :] + rel[None, :, :]).reshape(-1, 3)
    dev = 0.0001 * h
    keep = (pos[:, 0] <= x1 + dev) & (pos[:, 1] <= y1 + dev) & (pos[:, 2] <= 0.0 + dev)
    pos = pos[keep]
    lo = pos.min(0) - 0.5 * h
    hi = pos.max(0) + 0.5 * h
    return pos, (lo, hi)


def diamond100(cells_x, cells_y, cells_z, a):
    """Diamond crystal of cells_* unit cells (Diamond("100", 4*(cells-1), ...))."""
    base = np.array([
        [0, 0, 0], [0.5, 0.5, 0], [0.25, 0.25, -0.25], [0.75, 0.75, -0.25],
        [0, 0.5, -0.5], [0.5, 0, -0.5], [0.25, 0.75, -0.75], [0.75, 0.25, -0.75]]) * a
    i, j, k = np.meshgrid(np.arange(cells_x), np.arange(cells_y), np.arange(cells_z),
                          indexing="ij")
    org = np.stack([i.ravel() * a, j.ravel() * a, -k.ravel() * a], 1)
    pos = (org[:, None, :] + base[None, :, :]).reshape(-1, 3)
    lo = pos.min(0) - 0.125 * a
    hi = pos.max(0) + 0.125 * a
    return pos, (lo, hi)


def maxwell_boltzmann(n, mass, temperature, seed):
    """Velocities (A/ps) drawn from the Maxwell-Boltzmann distribution at
    temperature (K) with zero total momentum.  mass in eV ps^2/A^2 (scalar or
    per atom)."""
    rng = np.random.default_rng(seed)
    m = np.broadcast_to(np.asarray(mass, np.float64), (n,))
    v = rng.standard_normal((n, 3)) * np.sqrt(MD_BOLTZMANN * temperature / m)[:, None]
    p = (m[:, None] * v).sum(0)
    v -= p / m.sum()
    return v


def kinetic_temperature(v, mass):
    """T = 2/3 <Ek> / kB (include/omd/unit.h:119-128)."""
    m = np.broadcast_to(np.asarray(mass, np.float64), (len(v),))
    ek = 0.5 * (m[:, None] * v * v).sum()
    return 2.0 / 3.0 * ek / len(v) / MD_BOLTZMANN
