"""The C++ Object-MD host API: programs written like the reference's
test/example/example.cpp build against objectmd_amd/host/include and run on
the device.  CPU tests check the program side; GPU tests compare a whole
run's md.out with the FP64 oracle started from the program's own init.dat."""
import os
import subprocess
import sys

import numpy as np
import pytest

from objectmd_amd import tables
from tests.common import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "objectmd_amd", "host", "bin")


def _program(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built (python -m objectmd_amd.build)")
    return p


def _setup_eam(tmp, cells, steps, dt=0.002, every=20, temperature=150.0):
    tables.JohnsonEAM().write(str(tmp / "eam.johnson-al"))
    (tmp / "def.al-johnson").write_text(
        "element Aluminum\nmass 26.98154\nnumber 13\nstructure fcc\nlattice_constant 4.05\n")
    (tmp / "omd-parameter").write_text(
        f"material al-johnson\neam johnson-al\ncells {cells}\ntemperature {temperature}\n"
        f"time.step {dt}\ntime.max {(steps - 0.5) * dt}\nmonitor.every {every}\n"
        "comm.tolerance 0.4\n")


def _run(exe, tmp, timeout=300):
    env = dict(os.environ, OMD_TABLE=str(tmp))
    print(f"[host-api] running {os.path.basename(exe)} in {tmp}", file=sys.stderr, flush=True)
    return subprocess.run([exe], cwd=str(tmp), env=env, capture_output=True, text=True,
                          timeout=timeout)


def _md_out(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    return np.array([[float(v) for v in r] for r in rows])


def _init_dat(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    a = np.array([[float(v) for v in r] for r in rows])
    return a[:, :3], a[:, 3:6]


def test_program_fails_loudly_without_device(tmp_path):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    _setup_eam(tmp_path, 3, 10)
    r = _run(_program("eam_crystal"), tmp_path)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr


def test_missing_table_is_an_error(tmp_path):
    (tmp_path / "def.al-johnson").write_text("mass 26.98154\nnumber 13\nlattice_constant 4.05\n")
    (tmp_path / "omd-parameter").write_text(
        "material al-johnson\neam missing\ncells 2\ntime.max 0.01\n")
    r = _run(_program("eam_crystal"), tmp_path)
    assert r.returncode == 1
    assert "error" in r.stderr


@pytest.mark.gpu
def test_eam_program_matches_oracle(tmp_path):
    from oracle import oracle as O
    steps, every, dt = 200, 20, 0.002
    _setup_eam(tmp_path, 8, steps, dt=dt, every=every)
    r = _run(_program("eam_crystal"), tmp_path)
    assert r.returncode == 0, r.stderr
    md = _md_out(tmp_path / "md.out")
    assert list(md[:, 0].astype(int)) == list(range(every, steps, every))
    x, v = _init_dat(tmp_path / "init.dat")
    n = len(x)
    a = 4.05
    lo = x.min(0) - a / 4
    hi = x.max(0) + a / 4
    t = tables.JohnsonEAM().tables(5000)
    sp = [O.Spline.from_table(t[k]) for k in ("PAIR", "EDENS", "EMBED")]
    m = 26.98154 / tables.NORMAL_AMU
    _, _, en = O.run(1, x, v, m, lo, hi - lo, 7, sp, 5.5, 0.4, 5, steps, dt, maxn=96)
    for row in md:
        s = int(row[0])
        ek, ep, e = row[3], row[4], row[5]
        print(f"[host-api] step {s}: Ek {ek:.8f} oracle {en[s, 0]:.8f} "
              f"E {e:.8f} oracle {en[s, 2]:.8f}", file=sys.stderr, flush=True)
        assert abs(ek - en[s, 0]) / n < 2e-6
        assert abs(e - en[s, 2]) / n < 2e-6


@pytest.mark.gpu
def test_lj_program_with_inline_function(tmp_path):
    """InlineFunction reads Atoms(0) every step: the host mirror is synced
    every step and the trajectory file has one row per step."""
    t = tables.lj_table()
    tables.write_table_file(str(tmp_path / "pair.ar-cfg0"), {"": t})
    (tmp_path / "def.argon").write_text("element Argon\nmass 39.948\nnumber 18\n"
                                        "lattice_constant 5.26\n")
    steps = 50
    (tmp_path / "omd-parameter").write_text(
        f"material argon\npair ar-cfg0\ncells 4\ntemperature 40\ntime.step 0.001\n"
        f"time.max {(steps - 0.5) * 0.001}\nmonitor.every 10\ncomm.tolerance 0.3\n")
    r = _run(_program("lj_argon"), tmp_path)
    assert r.returncode == 0, r.stderr
    traj = np.loadtxt(tmp_path / "traj.dat")
    assert traj.shape == (steps, 3)
    md = _md_out(tmp_path / "md.out")
    e = md[:, 5] / 256
    assert np.ptp(e) < 1e-4


@pytest.mark.gpu
def test_tersoff_program_runs(tmp_path):
    import shutil
    shutil.copy(os.path.join(GOLDEN, "ref", "def.silicon"), tmp_path / "def.silicon")
    src = open(os.path.join(GOLDEN, "ref", "tersoff.silicon")).read()
    src = src.replace("tersoff.R 3.0", "tersoff.R 2.7").replace("tersoff.D 0.2", "tersoff.D 0.15")
    (tmp_path / "tersoff.si-c2").write_text(src)
    steps = 100
    (tmp_path / "omd-parameter").write_text(
        f"material silicon\ntersoff si-c2\ncells 5\ntemperature 300\ntime.step 0.001\n"
        f"time.max {(steps - 0.5) * 0.001}\nmonitor.every 20\ncomm.tolerance 0.3\n")
    r = _run(_program("si_tersoff"), tmp_path)
    assert r.returncode == 0, r.stderr
    md = _md_out(tmp_path / "md.out")
    e = md[:, 5] / (5 ** 3 * 8)
    assert np.ptp(e) < 2e-4
