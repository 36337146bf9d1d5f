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
