"""The C host driver (pic_run) runs the reference's step loop through the
C-ABI; the routine-by-routine loop and the fused pic_step agree, and a run
suspended and resumed (or stopped by a finish file and resumed) prints the
same diagnostics as an uninterrupted one."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "plasmasim_amd", "driver", "pic_run")


def run(*args, cwd=None):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120,
                       cwd=cwd)
    assert r.returncode == 0, r.stderr
    rows = []
    for line in r.stdout.splitlines():
        f = line.split()
        vals = [float(f[i + 1]) for i, t in enumerate(f) if t in ("Ee", "Eb1", "Ki", "Ke", "total")]
        rows.append(vals)
    return np.array(rows)


@pytest.mark.parametrize("form", ["--fp64", "--fp32"])
def test_routine_loop_matches_fused_step(form):
    a = run("--preset", "0", "--steps", "12", form, "--diag", "4")
    b = run("--preset", "0", "--steps", "12", form, "--diag", "4", "--fused")
    assert a.shape == b.shape == (3, 5)
    tol = 1e-10 if form == "--fp64" else 1e-4
    assert np.allclose(a, b, rtol=tol, atol=0), (a, b)


@pytest.mark.parametrize("mode", [[], ["--fused"]])
def test_suspend_and_resume_match_uninterrupted_run(tmp_path, mode):
    base = ["--preset", "1", "--fp32", "--det", "--diag", "5", *mode]
    whole = run(*base, "--steps", "15")
    ck = str(tmp_path / "s.ckpt")
    first = run(*base, "--steps", "10", "--suspend", ck, "--suspend-every", "5")
    rest = run(*base, "--steps", "15", "--resume", ck)
    assert np.array_equal(np.vstack([first, rest]), whole), (first, rest, whole)


def test_finish_file_stops_and_suspends(tmp_path):
    base = ["--preset", "1", "--fp32", "--det", "--diag", "5"]
    whole = run(*base, "--steps", "15")
    ck = str(tmp_path / "f.ckpt")
    (tmp_path / "finish").write_text("")
    first = run(*base, "--steps", "15", "--suspend", ck, cwd=str(tmp_path))
    assert first.shape[0] == 1  # stopped at the first diagnostic step
    os.remove(tmp_path / "finish")
    rest = run(*base, "--steps", "15", "--resume", ck, cwd=str(tmp_path))
    assert np.array_equal(np.vstack([first, rest]), whole)


def test_output_and_thrust_files(tmp_path):
    import plasmasim_amd as ps
    args = ["--preset", "1", "--fp32", "--det", "--fused", "--steps", "7", "--diag", "0",
            "--output-every", "3", "--thrust-every", "2", "--thrust-ave", "--version", "4"]
    run(*args, cwd=str(tmp_path))
    thrust = np.loadtxt(tmp_path / "thrust4.txt")
    ave = np.loadtxt(tmp_path / "thrust4_ave.txt")
    assert thrust.shape == (4, 1 + 27 + 6) and thrust[:, 0].tolist() == [0, 2, 4, 6]
    assert ave.shape == (4, 1 + 27) and ave[:, 0].tolist() == [0, 2, 4, 6]
    assert np.all(thrust[:, 16:28] == 0) and np.all(thrust[:, 32:] == 0)
    # the same run from Python: diagnostics after every step, dumps at 0, 3, 6
    c = ps.preset(1)
    c.precision = ps.PIC_FP32
    c.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
    s = ps.Simulation(c)
    s.load_particles()
    s.begin()
    F, E = [], []
    for step in range(7):
        s.step(1)
        d = s.diagnose()
        F.append(np.array(d["force"]))
        E.append(np.array(d["energy_box"]))
        if step % 3 == 0:
            mine = tmp_path / f"mine{step}.dat"
            s.output_fields(str(mine))
            assert mine.read_bytes() == (tmp_path / f"fdtd4-0-0-{step}.dat").read_bytes()
    for i, step in enumerate((0, 2, 4, 6)):
        assert np.allclose(thrust[i, 1:16], F[step], rtol=6e-4, atol=1e-300)
        assert np.allclose(thrust[i, 28:32], E[step][:4], rtol=1e-8, atol=1e-300)
        window = [F[step]] if step == 0 else [F[step - 1], F[step]]
        assert np.allclose(ave[i, 1:16], sum(window) / 2, rtol=6e-4, atol=1e-300)
