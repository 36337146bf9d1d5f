"""Edge cases of the whole step on the GPU, through the C-ABI: no particles at
all, a species whose particles all leave through the flow-axis faces, and
particle counts that fill neither the pass's 4-particle vector groups nor a
tile.  FP64 form against the oracle (relative L2 <= 1e-12, particle order and
counts exact); FP32 form against the oracle's FP64 run (fields to 1e-5
relative L2, particles matched one to one within 1e-5, counts exact)."""
import numpy as np
import pytest

import plasmasim_amd as ps
from helpers import random_fields, rel_l2, upload_fields
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
FIELDS = ["ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex", "jey",
          "jez", "ni", "ne"]
FORMS = [ps.PIC_FP64, ps.PIC_FP32]


def cfg(precision, **over):
    c = ps.preset(0)
    c.precision = precision
    c.inflow_every = 1 << 30  # no inflow: the counts stay what was uploaded
    c.sort_every = 1
    if precision == ps.PIC_FP32:
        # 8 tiles of 4 x 4 x 8 cells: ragged ranges per tile and tail chunks;
        # the FP64 form keeps the reference's single-key order
        c.tile[0], c.tile[1], c.tile[2] = 4, 4, 8
    for k, v in over.items():
        setattr(c, k, v)
    return c


def tol(precision):
    return 1e-12 if precision == ps.PIC_FP64 else 1e-5


def pair(precision, seed):
    """A library context and an FP64 oracle with the same seeded fields."""
    g = ps.Simulation(cfg(precision))
    o = Oracle(cfg(ps.PIC_FP64))
    b0 = [o.download_field(n) for n in ("b0x", "b0y", "b0z")]
    f = random_fields(o.shape, seed, b0=b0)
    upload_fields(g, f)
    upload_fields(o, f)
    return g, o


def interior_particles(c, n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(2, c.nx + 2, n) * c.dx,
                            rng.uniform(2, c.ny + 2, n) * c.dx,
                            rng.uniform(4.5, c.nz - 0.5, n) * c.dx,
                            rng.standard_normal((n, 3)) * 0.05])


def compare_fields(g, o, t):
    for n in FIELDS:
        e = rel_l2(g.download_field(n), o.download_field(n))
        assert e <= t, (n, e)


def compare_particles(g, o, sp, precision):
    a, b = g.download_particles(sp), o.download_particles(sp)
    assert a.shape == b.shape
    if precision == ps.PIC_FP64:
        assert rel_l2(a, b) <= 1e-12
        return
    # the FP32 form orders by its tiles: match each particle to its nearest
    if len(a) == 0:
        return
    d = np.abs(a[:, None, :] - b[None, :, :]).max(-1)
    near = d.argmin(1)
    assert d.min(1).max() <= 1e-5
    assert len(set(near.tolist())) == len(a)


@pytest.mark.parametrize("precision", FORMS)
def test_run_without_particles(precision):
    g, o = pair(precision, 1)
    g.step(5)
    o.step(5)
    compare_fields(g, o, tol(precision))
    assert g.particle_count(0) == g.particle_count(1) == 0


@pytest.mark.parametrize("precision", FORMS)
def test_species_whose_particles_all_leave(precision):
    g, o = pair(precision, 3)
    c = cfg(precision)
    rng = np.random.default_rng(5)
    n = 300
    # electrons in the low guard layers of the flow axis: the particle
    # boundary drops every one of them in the first step
    gone = np.column_stack([rng.uniform(2, c.nx + 2, n) * c.dx,
                            rng.uniform(2, c.ny + 2, n) * c.dx,
                            rng.uniform(0.2, 1.9, n) * c.dx,
                            rng.standard_normal((n, 3)) * 0.05])
    ions = interior_particles(c, 77, 4)
    for s in (g, o):
        s.upload_particles(0, ions)
        s.upload_particles(1, gone)
    g.step(3)
    o.step(3)
    assert g.particle_count(1) == o.particle_count(1) == 0
    assert g.particle_count(0) == o.particle_count(0) == 77
    compare_fields(g, o, tol(precision))
    compare_particles(g, o, 0, precision)


@pytest.mark.parametrize("precision", FORMS)
@pytest.mark.parametrize("counts", [(1, 3), (5, 2), (257, 63)])
def test_ragged_counts(precision, counts):
    g, o = pair(precision, 2)
    c = cfg(precision)
    for sp, n in enumerate(counts):
        p = interior_particles(c, n, 10 + sp)
        for s in (g, o):
            s.upload_particles(sp, p)
    g.step(3)
    o.step(3)
    compare_fields(g, o, tol(precision))
    for sp, n in enumerate(counts):
        assert g.particle_count(sp) == o.particle_count(sp)
        compare_particles(g, o, sp, precision)


@pytest.mark.parametrize("n", [3, 4, 1027])
def test_inflow_kill_with_ragged_counts(n):
    """Inflow on uploaded particles spread over the whole flow axis, counts
    that are not (or are just) a multiple of the kill kernel's four slots
    per thread: the FP32 form against the oracle.  The survivors (between
    the slabs) and the refilled slabs hold the same particles."""
    c32 = cfg(ps.PIC_FP32)
    g = ps.Simulation(c32)
    o = Oracle(cfg(ps.PIC_FP64))
    rng = np.random.default_rng(n)
    for sp in (0, 1):
        m = n + 5 * sp
        p = np.column_stack([rng.uniform(2, c32.nx + 2, m) * c32.dx,
                             rng.uniform(2, c32.ny + 2, m) * c32.dx,
                             rng.uniform(2.01, c32.nz + 1.99, m) * c32.dx,
                             rng.standard_normal((m, 3)) * 0.05])
        # keep clear of the kill zone's edges (z = 4 dx and z = nz dx), where
        # the fp32 offsets could round across
        p = p[(np.abs(p[:, 2] - 4 * c32.dx) > 1e-3) & (np.abs(p[:, 2] - c32.nz * c32.dx) > 1e-3)]
        for s in (g, o):
            s.upload_particles(sp, p)
    g.inflow()
    o.inflow()
    for sp in (0, 1):
        assert g.particle_count(sp) == o.particle_count(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape
        lo, hi = 4 * c32.dx, c32.nz * c32.dx
        for part in (lambda q: (q[:, 2] >= lo) & (q[:, 2] <= hi),
                     lambda q: (q[:, 2] < lo) | (q[:, 2] > hi)):
            pa, pb = a[part(a)], b[part(b)]
            assert pa.shape == pb.shape, sp
            for j in range(6):
                assert np.abs(np.sort(pa[:, j]) - np.sort(pb[:, j])).max(initial=0) < 1e-5, (sp, j)
