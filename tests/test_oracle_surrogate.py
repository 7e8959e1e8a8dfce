"""The surrogate's oracle (oracle/surrogate.py) pinned by finite differences
and invariances, and the training descriptor (PyTorch) checked against it."""
import numpy as np
import pytest

from objectmd_amd import lattice
from objectmd_amd import surrogate as S
from oracle import surrogate as OS


def _model(seed=3):
    rng = np.random.default_rng(seed)
    return dict(W1=rng.normal(0, 0.3, (64, 32)), b1=rng.normal(0, 0.1, 64),
                W2=rng.normal(0, 0.2, (64, 64)), b2=rng.normal(0, 0.1, 64),
                w3=rng.normal(0, 0.2, 64), b3=-4.0, mean=np.full(32, 0.5),
                std=np.full(32, 0.7), lo=np.zeros(32), hi=np.ones(32))


def _cell(seed=1, disp=0.15, cells=2):
    rng = np.random.default_rng(seed)
    pos, (lo, hi) = lattice.diamond100(cells, cells, cells, 5.431)
    return pos + rng.normal(0, disp, pos.shape), lo, hi


@pytest.mark.parametrize("kind", [1, 0])
def test_forces_are_minus_gradient(kind):
    x, lo, hi = _cell()
    m = _model()
    d = OS.default_desc()
    if kind == 0:  # Behler's (1 + lambda cos)^zeta form
        d = dict(d, kind=0, pq=[(1.0, 1.0), (-1.0, 1.0), (1.0, 2.0), (-1.0, 2.0),
                                (1.0, 4.0), (-1.0, 4.0), (1.0, 16.0), (-1.0, 16.0)])
    E, Ei, F, G = OS.energy_forces(x, lo, hi, 7, m, d, bf=False)
    assert np.abs(F).max() > 0.5
    h = 1e-5
    for a in range(0, len(x), 9):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            fd = -(OS.energy_forces(xp, lo, hi, 7, m, d, bf=False)[0] -
                   OS.energy_forces(xm, lo, hi, 7, m, d, bf=False)[0]) / (2 * h)
            assert fd == pytest.approx(F[a, c], abs=1e-7)


def test_invariances():
    x, lo, hi = _cell(seed=5)
    m = _model()
    E, Ei, F, G = OS.energy_forces(x, lo, hi, 7, m, bf=False)
    assert np.abs(F.sum(0)).max() < 1e-9          # no net force
    # a rigid shift inside the periodic box leaves everything unchanged
    E2, _, F2, G2 = OS.energy_forces(x + np.array([0.7, -0.3, 1.1]), lo, hi, 7, m, bf=False)
    assert E2 == pytest.approx(E, rel=1e-12)
    assert np.allclose(np.sort(G2, 0), np.sort(G, 0), atol=1e-10)


def test_descriptor_known_values():
    """Ideal diamond: 4 neighbours at a sqrt(3)/4 (inside 3 A), tetrahedral
    angles cos = -1/3, so every G has a closed form."""
    a = 5.431
    pos, (lo, hi) = lattice.diamond100(2, 2, 2, a)
    d = OS.default_desc()
    E, Ei, F, G = OS.energy_forces(pos, lo, hi, 7, _model(), d, bf=False)
    r = a * np.sqrt(3) / 4
    fc = 0.5 * (np.cos(np.pi * r / d["rc"]) + 1)
    mu = d["mu0"] + d["dmu"] * np.arange(24)
    assert np.allclose(G[:, :24], 4 * np.exp(-d["eta_r"] * (r - mu) ** 2) * fc, rtol=1e-12)
    for m, (p, q) in enumerate(d["pq"]):
        g = 6 * np.exp(-q * (-1.0 / 3.0 - p) ** 2) * fc * fc
        assert np.allclose(G[:, 24 + m], g, rtol=1e-12)
    assert np.abs(F).max() < 1e-9


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    v = np.random.default_rng(0).normal(0, 3, 10000)
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OS.bf16(v), ref)


def test_training_descriptor_matches_oracle():
    torch = pytest.importorskip("torch")
    x, lo, hi = _cell(seed=9, cells=3)
    d = S.default_desc()
    assert S.desc_vector(d).tolist() == OS.desc_vector(OS.default_desc())
    ii, jj, sh = S.neighbour_pairs(x, lo, hi, 3, d["rc"])
    tu, tw = S.triplets(ii, len(x))
    t = lambda a: torch.as_tensor(a)
    G = S.torch_descriptor(t(x), t(ii), t(jj), t(sh), t(tu), t(tw), d, len(x)).numpy()
    Go = OS.energy_forces(x, lo, hi, 3, _model(), OS.default_desc(), bf=False)[3]
    assert np.allclose(G, Go, rtol=1e-10, atol=1e-12)


def test_api_blocks_layout():
    m = _model()
    desc, norm, w = S.api_blocks(m, S.default_desc())
    assert desc.size == 22 and norm.size == 128 and w.size == 6337
    assert w[0] == np.float32(m["W1"][0, 0]) and w[-1] == np.float32(m["b3"])
    assert w[2048 + 64] == np.float32(m["W2"][0, 0])
