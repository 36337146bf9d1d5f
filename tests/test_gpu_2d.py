"""The 2D3V step on the GPU through the C-ABI, against the CPU oracle.

Plain FP64 form: every routine of the 2D fdtd() and whole steps match the
oracle to relative L2 <= 1e-12 (fields, particle phase space), integer
work (loaded and injected positions, boundary wrap and delete, sort order,
particle counts) bit-exact.  FP32 form: per-step relative error against the
FP64 form on the same inputs, and a multi-step run with sort, inflow and
the Poisson correction."""
import numpy as np
import pytest

import plasmasim_amd as ps
from helpers import rel_l2
from oracle_lib import Oracle2D

pytestmark = pytest.mark.gpu
TOL = 1e-12
EB = ["ex", "ey", "ez", "bx", "by", "bz"]
JN = ["jix", "jiy", "jiz", "jex", "jey", "jez", "ni", "ne"]


def cfg2(preset=0, precision=ps.PIC_FP64, **over):
    c = ps.preset2d(preset)
    c.precision = precision
    for k, v in over.items():
        setattr(c, k, v)
    return c


def pair(**over):
    c = cfg2(**over)
    return ps.Simulation2D(c), Oracle2D(c), c


def seed_fields(sims, shape, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    f = {n: rng.standard_normal(shape) * scale for n in EB}
    for s in sims:
        for n, a in f.items():
            s.upload_field(n, a)
    return f


def test_create_fields_and_sigma_identical():
    g, o, c = pair()
    for n in ("b0x", "b0y", "b0z", "bx", "ex", "phi"):
        assert np.array_equal(g.download_field(n), o.download_field(n)), n


@pytest.mark.parametrize("routine", ["renew_b_half", "renew_e", "halo",
                                     "fold_current", "fold_density", "poisson"])
def test_field_routine_matches_oracle(routine):
    g, o, c = pair()
    seed_fields((g, o), o.shape, 1)
    rng = np.random.default_rng(2)
    for n in JN:
        a = rng.standard_normal(o.shape) * 1e-4 + (0.04 if n in ("ni", "ne") else 0)
        g.upload_field(n, a)
        o.upload_field(n, a)
    getattr(g, routine)()
    getattr(o, routine)()
    for n in EB + JN + ["phi"]:
        a, b = g.download_field(n), o.download_field(n)
        assert rel_l2(a, b) <= TOL, (routine, n, rel_l2(a, b))


def test_external_current_ramp():
    g, o, c = pair()
    for step in (0, 7, c.ext_ramp + 3):
        g.external_current(step)
        o.external_current(step)
        assert np.array_equal(g.download_field("jiy"), o.download_field("jiy"))


def test_load_bit_exact_positions():
    g, o, c = pair()
    g.load_particles()
    o.load_particles()
    for sp in (0, 1):
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape == (c.nx * c.nz * c.ppc, 5)
        assert np.array_equal(a[:, :2], b[:, :2])
        assert rel_l2(a[:, 2:], b[:, 2:]) < 1e-14


def test_push_deposit_and_drift_match_oracle():
    g, o, c = pair()
    for s in (g, o):
        s.load_particles()
    seed_fields((g, o), o.shape, 3, scale=1e-2)
    for sp in (0, 1):
        for s in (g, o):
            s.renew_grid()
            s.push_deposit_j(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert rel_l2(a, b) <= TOL
        for n in JN[:6]:
            assert rel_l2(g.download_field(n), o.download_field(n)) <= TOL, n
        for s in (g, o):
            s.drift_deposit_n(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert rel_l2(a, b) <= TOL
        nn = "ni" if sp == 0 else "ne"
        assert rel_l2(g.download_field(nn), o.download_field(nn)) <= TOL


def test_particle_boundary_wraps_and_deletes_bit_exact():
    g, o, c = pair()
    rng = np.random.default_rng(4)
    n = 5000
    p = np.column_stack([rng.uniform(0.5, c.nx + 3.5, n), rng.uniform(1.0, c.nz + 3.0, n),
                         rng.standard_normal((n, 3)) * 0.01])
    p[:10, 0] = 1.0  # x == dx: deleted, not wrapped
    for s in (g, o):
        s.upload_particles(1, p)
        s.particle_boundary()
    a, b = g.download_particles(1), o.download_particles(1)
    assert len(a) < n and np.array_equal(a, b)


def test_inflow_and_sort_bit_exact():
    g, o, c = pair(tile=(8, 16))
    rng = np.random.default_rng(5)
    n = 20000
    p = np.column_stack([rng.uniform(0.5, c.nx + 3.5, n), rng.uniform(0.5, c.nz + 3.5, n),
                         rng.standard_normal((n, 3))])
    for s in (g, o):
        s.upload_particles(0, p)
        s.upload_particles(1, p[::-1].copy())
        s.inflow()
    # positions are integer-exact arithmetic on the same draws; velocities
    # pass through the device's log and cos, which may differ from the
    # host's in the last bit
    for sp in (0, 1):
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape
        assert np.array_equal(a[:, :2], b[:, :2])
        assert rel_l2(a[:, 2:], b[:, 2:]) < 1e-14
    for s in (g, o):
        s.sort(0)
        s.sort(1)
    for sp in (0, 1):
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert g.particle_count(sp) == o.particle_count(sp)
        assert np.array_equal(a[:, :2], b[:, :2])
        assert rel_l2(a[:, 2:], b[:, 2:]) < 1e-14


def test_fp32_tile_sort_matches_oracle():
    """FP32 form: the sort after 6 steps (the tiles of the load's sort still
    current, the inflows of begin and of steps 3 and 6 appended as a tail)
    takes the tile sort, and its result equals the oracle's stable sort of
    the same particles (cell-major within 16 x 16 tiles, slot order within a
    cell; the dropped ones removed)."""
    c = cfg2(1, ps.PIC_FP32, sort_every=1000, inflow_every=3)
    g = ps.Simulation2D(c)
    g.load_particles()
    g.begin()
    g.step(6)
    pre = [g.download_particles(sp) for sp in (0, 1)]
    t0, r0 = g.sort_stats()
    g.sort(0)
    g.sort(1)
    t1, r1 = g.sort_stats()
    assert (t1 - t0, r1 - r0) == (2, 0)
    o = Oracle2D(cfg2(1, ps.PIC_FP64, sort_every=1000, inflow_every=3))
    for sp in (0, 1):
        o.upload_particles(sp, pre[sp])
        o.sort(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape and np.array_equal(a, b), sp
    g.close()


@pytest.mark.parametrize("sort_every", [0, 3])
def test_whole_steps_match_oracle(sort_every):
    g, o, c = pair(sort_every=sort_every, inflow_every=4)
    for s in (g, o):
        s.load_particles()
        s.begin()
    for block in range(3):
        for s in (g, o):
            s.step(4)
        for n in EB + ["ni", "ne", "phi"]:
            a, b = g.download_field(n), o.download_field(n)
            assert rel_l2(a, b) <= TOL, (block, n, rel_l2(a, b))
        for sp in (0, 1):
            assert g.particle_count(sp) == o.particle_count(sp)
            a, b = g.download_particles(sp), o.download_particles(sp)
            assert rel_l2(a, b) <= TOL, (block, sp, rel_l2(a, b))
    dg, do = g.diagnose(), o.diagnose()
    for k in ("field_energy_e", "field_energy_b", "field_energy_b1", "charge",
              "gauss_l2", "divb_l2"):
        assert abs(dg[k] - do[k]) <= 1e-10 * max(abs(do[k]), 1e-300), k
    assert np.allclose(dg["kinetic"], do["kinetic"], rtol=1e-12, atol=0)
    assert np.array_equal(dg["count"], do["count"])
    # probe columns: point values after 12 steps, so a pointwise tolerance
    # a little looser than the L2 bar; the last two columns are relative
    # residuals |div B|/|B| and |div E - rho/eps0|/|E| (dimensionless), whose
    # exact value after the Poisson correction is 0: an absolute floor of
    # 1e-12 covers their rounding residue
    k, m = c.ext_k, c.ext_m[1]
    pg, po = g.probe(k, m), o.probe(k, m)
    print("probe GPU", pg, "oracle", po)
    assert np.allclose(pg[:6], po[:6], rtol=1e-9, atol=1e-11 * np.abs(po[:6]).max())
    assert np.allclose(pg[6:], po[6:], rtol=1e-8, atol=1e-12)


def test_fused_step_matches_routine_sequence():
    """pic2d_step's fused pass against the routine-by-routine loop of fdtd
    on the GPU itself (plain FP64: identical to rounding)."""
    c = cfg2(inflow_every=3)
    a, b = ps.Simulation2D(c), ps.Simulation2D(c)
    for s in (a, b):
        s.load_particles()
        s.begin()
    a.step(5)
    for step in range(5):
        b.renew_grid()
        if step == 3:
            b.inflow()
        b.renew_b_half()
        b.halo()
        for sp in (0, 1):
            b.push_deposit_j(sp)
        b.fold_current()
        b.external_current(step)
        b.renew_b_half()
        b.halo()
        for sp in (0, 1):
            b.drift_deposit_n(sp)
        b.fold_density()
        b.renew_e()
        b.halo()
        b.poisson()
        b.halo()
        b.particle_boundary()
    for n in EB + ["ni", "ne", "phi"]:
        assert rel_l2(a.download_field(n), b.download_field(n)) <= TOL, n
    for sp in (0, 1):
        assert rel_l2(a.download_particles(sp), b.download_particles(sp)) <= TOL


# ---------------------------------------------------------------- FP32 form
def test_fp32_one_step_error_vs_fp64():
    c64 = cfg2(1, ps.PIC_FP64)
    c32 = cfg2(1, ps.PIC_FP32)
    a, b = ps.Simulation2D(c64), ps.Simulation2D(c32)
    for s in (a, b):
        s.load_particles()
        s.begin()
        s.step(3)
    errs = {}
    for n in EB + JN + ["phi"]:
        errs[n] = rel_l2(b.download_field(n), a.download_field(n))
    print("2D FP32 vs FP64 after 3 steps, relative L2:",
          {k: f"{v:.2e}" for k, v in errs.items()})
    for n in EB + JN:
        assert errs[n] <= 1e-5, (n, errs[n])
    for sp in (0, 1):
        assert a.particle_count(sp) == b.particle_count(sp)


def test_fp32_step_sort_applied_by_the_pass_matches_copying_sort():
    """pic2d_step's FP32 tile sort leaves its permutation to the species'
    fused pass, which reads the particles through it.  A run sorting inside
    each step that way and a run sorting with the C-ABI sort (which copies
    the particles) before the same steps hold the same particles in the same
    slots after every step: slot-by-slot relative L2 <= 1e-6 (the passes'
    float atomics add in no fixed order), counts exact, and the fields to
    the same bar; a wrong permutation moves particles between slots, an
    error of order one."""
    runs = []
    for in_step in (True, False):
        c = cfg2(1, ps.PIC_FP32, inflow_every=1000, sort_every=1 if in_step else 1000)
        s = ps.Simulation2D(c)
        s.load_particles()
        s.begin()
        states = []
        for k in range(5):
            # pic2d_step sorts when c_sort >= sort_every: from the second step
            if not in_step and k >= 1:
                s.sort(0)
                s.sort(1)
            s.step(1)
            states.append([s.download_particles(sp) for sp in (0, 1)])
        runs.append((states, {n: s.download_field(n) for n in EB + ["ni", "ne"]},
                     s.sort_stats()))
        s.close()
    (pa, fa, sa), (pb, fb, sb) = runs
    assert sa == sb == (8, 0), (sa, sb)
    for k in range(5):
        for sp in (0, 1):
            a, b = pa[k][sp], pb[k][sp]
            assert a.shape == b.shape and a.shape[0] > 0, (k, sp)
            assert rel_l2(a, b) <= 1e-6, (k, sp, rel_l2(a, b))
    for n in fa:
        assert rel_l2(fa[n], fb[n]) <= 1e-6, (n, rel_l2(fa[n], fb[n]))


def test_fp32_multistep_with_sort_inflow_poisson():
    c = cfg2(1, ps.PIC_FP32, inflow_every=7, sort_every=5)
    s = ps.Simulation2D(c)
    s.load_particles()
    s.begin()
    n0 = [s.particle_count(sp) for sp in (0, 1)]
    s.step(30)
    d = s.diagnose()
    assert all(np.isfinite([d["field_energy_e"], d["field_energy_b"], d["gauss_l2"]]))
    for sp in (0, 1):
        assert abs(d["count"][sp] - n0[sp]) < 0.05 * n0[sp]


def test_fp32_out_of_window_particles_match_fp64():
    """Unsorted particles in chunks of 2048 slots, each chunk inside a 4 x 4
    cell box except 2 % scattered anywhere: most chunks keep an LDS window
    and deposit their strays through the wave-cooperative global path; a
    chunk whose sampled box takes in a stray has no window and deposits lane
    by lane.  One step against the FP64 form on the same upload: J, n, E, B
    relative L2 <= 1e-5, the same particles."""
    c64, c32 = cfg2(1, ps.PIC_FP64), cfg2(1, ps.PIC_FP32)
    rng = np.random.default_rng(29)
    parts = []
    for sp in (0, 1):
        vth = 0.005 if sp == 0 else 0.05
        chunks = []
        for _ in range(48):
            x0 = rng.integers(2, c64.nx - 2)
            z0 = rng.integers(10, c64.nz - 10)
            p = np.column_stack([x0 + rng.uniform(0, 4, 2048), z0 + rng.uniform(0, 4, 2048),
                                 rng.standard_normal((2048, 3)) * vth])
            stray = rng.random(2048) < 0.02
            p[stray, 0] = rng.uniform(2, c64.nx + 2, stray.sum())
            p[stray, 1] = rng.uniform(10, c64.nz - 6, stray.sum())
            chunks.append(p)
        parts.append(np.concatenate(chunks))
    sims = [ps.Simulation2D(c) for c in (c64, c32)]
    for s in sims:
        for sp in (0, 1):
            s.upload_particles(sp, parts[sp])
        s.step(1)
    a, b = sims
    for n in EB + JN:
        e = rel_l2(b.download_field(n), a.download_field(n))
        assert e <= 1e-5, (n, e)
    for sp in (0, 1):
        pa, pb = a.download_particles(sp), b.download_particles(sp)
        assert pa.shape == pb.shape
        # the same particles: each coordinate's sorted values agree (pairing
        # particles by sorting rows is ambiguous at 128 per cell, where two
        # can be closer than the fp32 rounding)
        for j in range(5):
            assert np.abs(np.sort(pa[:, j]) - np.sort(pb[:, j])).max() < 1e-5, (sp, j)
