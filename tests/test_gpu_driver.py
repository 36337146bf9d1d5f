"""The C host driver (pic_run) runs the reference's step loop through the
C-ABI; the routine-by-routine loop and the fused pic_step agree."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "plasmasim_amd", "driver", "pic_run")


def run(*args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
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
