"""The 3D Poisson correction on the GPU (pic_poisson; pic_step with
poisson_every) against the oracle: phi and E to relative L2 <= 1e-12 in the
FP64 form for the routine and over whole steps; the FP32 form against the
FP64 form to 1e-5; the routine loop of the C driver against pic_step."""
import numpy as np
import pytest

import plasmasim_amd as ps
from helpers import rel_l2
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
NAMES = ["ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex", "jey",
         "jez", "ni", "ne"]


def cfg(precision=ps.PIC_FP64, every=1, margin=0, preset=0):
    c = ps.preset(preset)
    c.precision = precision
    c.poisson_every = every
    c.phi_margin = margin
    c.inflow_every = 3
    return c


def seeded(sims, shape, seed):
    rng = np.random.default_rng(seed)
    for n in ("ex", "ey", "ez", "ni", "ne"):
        a = rng.standard_normal(shape) * (1e-2 if n[0] == "e" else 1.0)
        if n in ("ni", "ne"):
            a = np.abs(a) * 0.04
        for s in sims:
            s.upload_field(n, a)


@pytest.mark.parametrize("margin", [0, 2])
def test_routine_matches_oracle(margin):
    c = cfg(margin=margin)
    g, o = ps.Simulation(c), Oracle(c)
    seeded((g, o), o.shape, 11)
    g.poisson()
    o.poisson()
    for n in ("phi", "ex", "ey", "ez"):
        e = rel_l2(g.download_field(n), o.download_field(n))
        assert e <= TOL, (n, e)


@pytest.mark.parametrize("every", [1, 2])
def test_whole_steps_match_oracle(every):
    c = cfg(every=every)
    g, o = ps.Simulation(c), Oracle(c, threads=4)
    for s in (g, o):
        s.load_particles()
        s.begin()
        s.step(7)
    for n in NAMES + ["phi"]:
        e = rel_l2(g.download_field(n), o.download_field(n))
        assert e <= TOL, (n, e)
    for sp in (0, 1):
        assert g.particle_count(sp) == o.particle_count(sp)
        assert rel_l2(g.download_particles(sp), o.download_particles(sp)) <= TOL


def test_fp32_form_against_fp64():
    c64, c32 = cfg(preset=1), cfg(ps.PIC_FP32, preset=1)
    a, b = ps.Simulation(c64), ps.Simulation(c32)
    seeded((a, b), (c64.nx + 4, c64.ny + 4, c64.nz + 4), 5)
    a.poisson()
    b.poisson()
    for n in ("phi", "ex", "ey", "ez"):
        e = rel_l2(b.download_field(n), a.download_field(n))
        assert e <= 1e-5, (n, e)


def test_gauss_residual_drops_with_the_correction():
    res = {}
    for every in (0, 1):
        c = cfg(ps.PIC_FP32, every=every, preset=1)
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(20)
        res[every] = s.diagnose()["gauss_l2"]
    print("Gauss L2 after 20 steps at 32^3 without / with the correction:", res)
    assert res[1] < 0.7 * res[0], res


def test_disabled_and_invalid_configs():
    s = ps.Simulation(cfg(every=0))
    with pytest.raises(ps.PicError) as e:
        s.poisson()
    assert e.value.code == -5
    assert np.all(s.download_field("phi") == 0)
    bad = cfg()
    bad.nz = 12  # not a power of two
    with pytest.raises(ps.PicError):
        ps.Simulation(bad)
