"""pic_output_fields against the oracle's output() / output_tecplot():
byte-identical files on an identical FP64 state, the same values (to the
printed precision) after steps of the FP64 form, and the FP32 form's dump
of the same state within its rounding."""
import io

import numpy as np
import pytest

import plasmasim_amd as ps
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
NAMES = ["ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z", "jix", "jiy",
         "jiz", "jex", "jey", "jez", "ni", "ne"]


def rows(path):
    return np.loadtxt(io.StringIO("\n".join(
        s for s in open(path).read().split("\n") if s and s[0] in "-0123456789")))


@pytest.mark.parametrize("tec", [False, True])
def test_identical_state_gives_identical_file(tmp_path, tec):
    c = ps.preset(0)
    g, o = ps.Simulation(c), Oracle(c)
    rng = np.random.default_rng(7)
    for n in NAMES:
        a = rng.standard_normal(o.shape) * 10.0 ** rng.integers(-6, 2)
        g.upload_field(n, a)
        o.upload_field(n, a)
    pg, po = tmp_path / "g.dat", tmp_path / "o.dat"
    g.output_fields(str(pg), tecplot=tec, step=9)
    o.output_fields(str(po), tecplot=tec, step=9)
    assert pg.read_bytes() == po.read_bytes()


@pytest.mark.parametrize("poisson", [0, 1])
def test_after_steps_matches_oracle_to_printed_precision(tmp_path, poisson):
    c = ps.preset(0)
    c.inflow_every = 3
    c.poisson_every = poisson  # 1: the phi column carries the potential
    c.phi_margin = 0
    g, o = ps.Simulation(c), Oracle(c, threads=4)
    for s in (g, o):
        s.load_particles()
        s.begin()
        s.step(8)
    pg, po = tmp_path / "g.dat", tmp_path / "o.dat"
    g.output_fields(str(pg))
    o.output_fields(str(po))
    a, b = rows(pg), rows(po)
    assert a.shape == b.shape == (c.nx * c.ny * c.nz, 23)
    # values agree to ~1e-13 of each column's scale: a printed digit differs
    # only next to a rounding boundary or for values near zero
    scale = np.abs(b).max(axis=0)
    assert np.all(np.abs(a - b) <= 2e-3 * np.abs(b) + 1e-11 * scale)
    # most printed tokens are identical (with the Poisson correction, near-
    # zero phi values from the FFT and from the oracle's direct sums differ
    # in the last printed digit more often)
    assert np.mean(a == b) > 0.99
    assert (np.abs(a[:, 18]).max() > 0) == bool(poisson)


def test_fp32_form_dump(tmp_path):
    c = ps.preset(1)
    c64 = c.copy()
    c64.precision = ps.PIC_FP64
    c32 = c.copy()
    c32.precision = ps.PIC_FP32
    a, b = ps.Simulation(c64), ps.Simulation(c32)
    for s in (a, b):
        s.load_particles()
        s.begin()
        s.step(2)
    pa, pb = tmp_path / "a.dat", tmp_path / "b.dat"
    a.output_fields(str(pa), tecplot=True, step=2)
    b.output_fields(str(pb), tecplot=True, step=2)
    x, y = rows(pa), rows(pb)
    assert x.shape == y.shape == ((c.nx + 1) * (c.ny + 1) * c.nz, 24)
    # columns 0-11 (positions, densities, E, total B): FP32 vs FP64 after two
    # steps, relative to each column's scale
    scale = np.abs(x).max(axis=0) + 1e-300
    assert np.all(np.abs(x - y)[:, :12].max(axis=0) <= 2e-3 * scale[:12])
