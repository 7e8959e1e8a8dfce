"""Pins the oracle's lattice builders and neighbour lists."""
import os

import numpy as np
import pytest

from objectmd_amd import lattice
from tests.common import GOLDEN, read_xyz


def test_fcc100_matches_reference_builder(orc):
    # FCC("100", 20, 20, 20) of config 0: 10^3 cells, 4000 atoms
    xyz, box = orc.fcc100(20, 20, 20, 5.26)
    assert xyz.shape == (4000, 3)
    pos, (lo, hi) = lattice.fcc100(10, 10, 10, 5.26)
    assert np.array_equal(pos, xyz)               # same atoms, same order
    assert np.allclose(box, [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], rtol=0, atol=0)
    assert hi[0] - lo[0] == pytest.approx(52.6)   # L = cells * a
    d = np.linalg.norm(xyz[1] - xyz[0])
    assert d == pytest.approx(5.26 / np.sqrt(2))
    # crystal grows downward from z = 0 (to rounding of i*a + z0)
    assert xyz[:, 2].max() == pytest.approx(0.0, abs=1e-12)


def test_fcc_odd_monolayers(orc):
    xyz, _ = orc.fcc100(3, 3, 3, 4.0)
    # monolayers 0, h, 2h in each direction: 14 atoms of a cube + face centres
    assert len(xyz) == 14


def test_diamond100(orc):
    xyz, box = orc.diamond100(196, 196, 196, 5.431)   # 196/4+1 = 50 cells
    assert len(xyz) == 50 ** 3 * 8
    small, _ = orc.diamond100(4, 4, 4, 5.431)
    pos, (lo, hi) = lattice.diamond100(2, 2, 2, 5.431)
    assert np.allclose(np.sort(small, 0), np.sort(pos, 0))
    d = np.linalg.norm(small[2] - small[0])
    assert d == pytest.approx(5.431 * np.sqrt(3) / 4)


def _pairs(cnt, code):
    return [tuple(code[i, :cnt[i]]) for i in range(len(cnt))]


@pytest.mark.parametrize("pbc", [0, 7, 3, 5])
def test_cells_equal_brute_force(orc, pbc):
    rng = np.random.default_rng(11 + pbc)
    L = np.array([13.0, 11.5, 12.25])
    lo = np.array([-1.0, 0.5, -3.0])
    x = lo + rng.random((300, 3)) * L
    a = orc.neighbors(x, lo, L, pbc, 3.3, maxn=64)
    b = orc.neighbors(x, lo, L, pbc, 3.3, maxn=64, brute=True)
    assert _pairs(*a) == _pairs(*b)
    # symmetry: j in list(i) with image s  <=>  i in list(j) with image -s
    cnt, code = a
    s = set()
    for i in range(len(x)):
        for c in code[i, :cnt[i]]:
            s.add((i, c // 27, c % 27))
    for (i, j, im) in s:
        assert (j, i, 26 - im) in s


def test_inclusive_membership(orc):
    # two atoms exactly rlist apart: r^2 <= rlist^2 keeps them (VerletList.cpp:184)
    x = np.array([[0.0, 0.0, 0.0], [2.5, 0.0, 0.0]])
    cnt, code = orc.neighbors(x, np.zeros(3), np.ones(3) * 10, 0, 2.5, brute=True)
    assert list(cnt) == [1, 1]
    cnt, _ = orc.neighbors(x, np.zeros(3), np.ones(3) * 10, 0, np.nextafter(2.5, 0),
                           brute=True)
    assert list(cnt) == [0, 0]


def test_fcc_shell_counts(orc):
    pos, (lo, hi) = lattice.fcc100(6, 6, 6, 4.05)
    cnt, _ = orc.neighbors(pos, lo, hi - lo, 7, 5.5 + 0.3)
    assert np.all(cnt == 54)          # 12 + 6 + 24 + 12
    pos, (lo, hi) = lattice.fcc100(10, 10, 10, 5.26)
    cnt, _ = orc.neighbors(pos, lo, hi - lo, 7, 8.5 + 0.3)
    assert np.all(cnt == 78)          # 12 + 6 + 24 + 12 + 24


def test_small_periodic_box_images(orc):
    # box shorter than 2 rlist: an atom sees its own images
    x = np.array([[0.5, 0.5, 0.5]])
    cnt, code = orc.neighbors(x, np.zeros(3), np.array([3.0, 3.0, 3.0]), 7, 3.0,
                              brute=True)
    assert cnt[0] == 6


@pytest.mark.parametrize("name,count", [("dimer", 2), ("trimer", 3), ("tetra", 5),
                                        ("1x1x1", 8)])
def test_reference_clusters(orc, name, count):
    """test/check-verlet-list clusters, list radius 5.0 as its CheckForce:
    every pair within 5 A, both directions."""
    x = read_xyz(os.path.join(GOLDEN, "ref", name))
    assert len(x) == count
    cnt, code = orc.neighbors(x, np.zeros(3), np.ones(3), 0, 5.0, brute=True)
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    np.fill_diagonal(d, 1e9)
    assert np.array_equal(cnt, (d <= 5.0).sum(1))
