"""thrust_fave (PIC/3d_pic458_domain.c:926, :5124-5285) summed over the
steps of pic_step (pic_thrust_ave_enable / pic_thrust_ave_get).

- FP64 form against the oracle: the oracle's diagnostics force (thrust_f's
  force over the same central box) taken at the end of every step and
  summed, relative to the largest component 1e-9.  On the 32^3 grid the
  force is added inside the next step's particle pass; on the 8x8x16 grid
  (nz < 20: the box reaches the inflow slab) it is taken before the inflow
  by the separate gather, and a read in the middle of the run takes the
  pending step the same way.
- FP32 form against the FP64 form from the same load: 1e-4 of the largest
  component.
- The sum does not perturb the run: with and without it, fields and
  particles are bit-identical (FP32 deterministic mode) or agree to 1e-12
  (FP64, whose float atomics round in varying order)."""
import numpy as np
import pytest

import plasmasim_amd as ps
from helpers import rel_l2
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
FIELDS = ["ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex", "jey",
          "jez", "ni", "ne"]


def started(c):
    s = ps.Simulation(c)
    s.load_particles()
    s.begin()
    return s


def oracle_sum(c, steps, threads=4):
    o = Oracle(c, threads=threads)
    o.load_particles()
    o.begin()
    acc = np.zeros(15)
    for _ in range(steps):
        o.step(1)
        acc += np.asarray(o.diagnose()["force"])
    o.close()
    return acc


@pytest.mark.parametrize("preset,steps", [(1, 12), (0, 11)])
def test_fp64_sum_matches_oracle(preset, steps):
    c = ps.preset(preset)
    c.precision = ps.PIC_FP64
    g = started(c)
    g.thrust_ave(True)
    if preset == 0:
        # a read in the middle takes the pending step by the separate gather
        g.step(4)
        F4, n4 = g.thrust_ave_get()
        assert n4 == 4
        g.step(steps - 4)
    else:
        g.step(steps)
    F, n = g.thrust_ave_get()
    g.close()
    assert n == steps
    ref = oracle_sum(c, steps)
    scale = np.abs(ref).max()
    print(f"preset {preset}: GPU {F[:15]}\n  oracle {ref}")
    assert scale > 0
    assert np.abs(F[:15] - ref).max() <= 1e-9 * scale, np.abs(F[:15] - ref) / scale
    assert np.all(F[15:] == 0.0)
    if preset == 0:
        assert np.abs(F4[:15]).max() > 0


def test_reset_and_reenable():
    c = ps.preset(1)
    c.precision = ps.PIC_FP64
    g = started(c)
    g.thrust_ave(True)
    g.step(3)
    F3, n3 = g.thrust_ave_get(reset=True)
    assert n3 == 3
    g.step(2)
    F2, n2 = g.thrust_ave_get()
    assert n2 == 2
    g.thrust_ave(True)  # re-enabling starts a new sum
    Fz, nz = g.thrust_ave_get()
    assert nz == 0 and np.all(Fz == 0.0)
    g.close()


def test_fp32_sum_near_fp64():
    out = {}
    for prec in (ps.PIC_FP32, ps.PIC_FP64):
        c = ps.preset(1)
        c.precision = prec
        c.tile[0], c.tile[1], c.tile[2] = 4, 4, 8
        g = started(c)
        g.thrust_ave(True)
        g.step(12)
        out[prec] = g.thrust_ave_get()
        g.close()
    (F32, n32), (F64, n64) = out[ps.PIC_FP32], out[ps.PIC_FP64]
    assert n32 == n64 == 12
    scale = np.abs(F64).max()
    err = np.abs(F32 - F64).max() / scale
    print(f"FP32 vs FP64 thrust sum: rel to the largest component {err:.2e}")
    assert err <= 1e-4, err


@pytest.mark.parametrize("precision", [ps.PIC_FP32, ps.PIC_FP64])
def test_sum_does_not_change_the_run(precision):
    sims = []
    for on in (False, True):
        c = ps.preset(1)
        c.precision = precision
        if precision == ps.PIC_FP32:
            c.tile[0], c.tile[1], c.tile[2] = 4, 4, 8
            c.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
        g = started(c)
        g.thrust_ave(on)
        g.step(12)
        sims.append(g)
    a, b = sims
    for n in FIELDS:
        x, y = a.download_field(n), b.download_field(n)
        if precision == ps.PIC_FP32:
            assert np.array_equal(x, y), n
        else:
            assert rel_l2(x, y) <= 1e-12, n
    for sp in (0, 1):
        pa, pb = a.download_particles(sp), b.download_particles(sp)
        assert pa.shape == pb.shape
        if precision == ps.PIC_FP32:
            assert np.array_equal(pa, pb, equal_nan=True), sp
    a.close()
    b.close()
