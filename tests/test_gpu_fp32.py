"""The FP32 chosen form against the FP64 plain form and the oracle.

Bars (tolerances stated per assertion):
  - loaded/inflow positions: cell exact, offset within fp32 rounding;
  - tile-order sort: same permutation as the oracle's stable sort (integer);
  - one step from an identical developed state: relative L2 of J, E and the
    induced field B - B0 against the FP64 form <= 1e-5;
  - deterministic deposit: bit-identical results run to run.
"""
import numpy as np
import pytest

import plasmasim_amd as ps
from helpers import rel_l2
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
FIELDS = ["ex", "ey", "ez", "bx", "by", "bz"]
J = ["jix", "jiy", "jiz", "jex", "jey", "jez", "ni", "ne"]


def cfg(preset=1, precision=ps.PIC_FP32, tile=(4, 4, 8), deposit=ps.PIC_DEPOSIT_ATOMIC):
    c = ps.preset(preset)
    c.precision = precision
    c.tile[0], c.tile[1], c.tile[2] = tile
    c.deposit = deposit
    return c


def test_load_matches_fp64_stream():
    c = cfg()
    g = ps.Simulation(c)
    g.load_particles()
    o = Oracle(cfg(precision=ps.PIC_FP64))
    o.load_particles()
    o.sort(0)
    o.sort(1)
    for sp in (0, 1):
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape
        assert np.array_equal(np.floor(a[:, :3]), np.floor(b[:, :3]))
        assert np.abs(a[:, :3] - b[:, :3]).max() < 1.2e-7 * 2
        assert np.abs(a[:, 3:] - b[:, 3:]).max() <= 6e-8 * np.abs(b[:, 3:]).max()


def test_tile_sort_permutation_matches_oracle():
    c = cfg(1)
    g = ps.Simulation(c)
    c64 = cfg(1, precision=ps.PIC_FP64)
    o = Oracle(c64)
    rng = np.random.default_rng(3)
    n = 300000
    p = np.column_stack([rng.uniform(0.5, c.nx + 3.5, n), rng.uniform(0.5, c.ny + 3.5, n),
                         rng.uniform(0.5, c.nz + 3.5, n), rng.standard_normal((n, 3)) * 0.01])
    # snap positions to fp32-representable offsets so both sides hold the
    # same particle and the same cell
    cell = np.floor(p[:, :3])
    p[:, :3] = cell + np.minimum(np.float32(p[:, :3] - cell).astype(np.float64), 0.99999994)
    p[:, 3:] = p[:, 3:].astype(np.float32)
    g.upload_particles(1, p)
    o.upload_particles(1, p)
    g.sort(1)
    o.sort(1)
    a, b = g.download_particles(1), o.download_particles(1)
    assert a.shape == b.shape
    assert np.array_equal(a, b)


@pytest.mark.parametrize("tile", [(4, 4, 8), (8, 4, 4), (2, 2, 2)])
@pytest.mark.parametrize("precision", [ps.PIC_FP32, ps.PIC_FP64])
def test_tile_sort_after_steps_matches_oracle(tile, precision):
    """The tile sort meets what a periodic sort meets: particles tile-sorted
    by the last sort, most still in their tile, some moved to other tiles,
    dead slots and an inflow tail.  The live particles it leaves are the
    oracle's stable sort of the same state, bit for bit and in order, and
    the sort took the tile path (not the radix sort).  The FP64 form runs
    the same passes on its key array."""
    c = cfg(1, precision=precision, tile=tile)
    c.sort_every, c.inflow_every = 1000, 5
    g = ps.Simulation(c)
    g.load_particles()
    g.begin()
    g.step(12)
    pre = [g.download_particles(sp) for sp in (0, 1)]
    t0, r0 = g.sort_stats()
    g.sort(0)
    g.sort(1)
    t1, r1 = g.sort_stats()
    assert (t1 - t0, r1 - r0) == (2, 0)
    o = Oracle(cfg(1, precision=ps.PIC_FP64, tile=tile))
    for sp in (0, 1):
        o.upload_particles(sp, pre[sp])
        o.sort(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape and len(a) > 0
        assert np.array_equal(a, b), sp
    # the steps after it run on the new tile ranges
    g.step(3)
    assert g.particle_count(0) > 0
    g.close()


@pytest.mark.parametrize("precision", [ps.PIC_FP32, ps.PIC_FP64])
def test_tile_sort_tiles_not_powers_of_two(precision):
    """Tiles of 3 x 4 x 6 cells on a 24^3 grid: the tile keys, the tile
    sort's destinations and the FP64 form's key decoding take their division
    path (the shifts serve power-of-two tiles only).  After 12 steps the sort
    takes the tile path, equals the oracle's stable sort of the same
    particles, and the steps after it run on the new ranges."""
    def conf(prec):
        c = ps.preset(1)
        c.precision = prec
        c.nx = c.ny = c.nz = 24
        c.tile[0], c.tile[1], c.tile[2] = 3, 4, 6
        c.sort_every, c.inflow_every = 1000, 5
        return c
    g = ps.Simulation(conf(precision))
    g.load_particles()
    g.begin()
    g.step(12)
    pre = [g.download_particles(sp) for sp in (0, 1)]
    t0, r0 = g.sort_stats()
    g.sort(0)
    g.sort(1)
    t1, r1 = g.sort_stats()
    assert (t1 - t0, r1 - r0) == (2, 0)
    o = Oracle(conf(ps.PIC_FP64))
    for sp in (0, 1):
        o.upload_particles(sp, pre[sp])
        o.sort(sp)
        a, b = g.download_particles(sp), o.download_particles(sp)
        assert a.shape == b.shape and len(a) > 0
        assert np.array_equal(a, b), sp
    g.step(3)
    assert g.particle_count(0) > 0 and g.particle_count(1) > 0
    g.close()


def cluster(seed, vel):
    """20000 electrons in one 2 x 2 x 2 tile of the ballistic layer (drift
    only), velocities from vel(rng, n)."""
    rng = np.random.default_rng(seed)
    n = 20000
    p = np.column_stack([rng.uniform(4, 6, n), rng.uniform(4, 6, n), rng.uniform(4, 6, n),
                         vel(rng, n)])
    cell = np.floor(p[:, :3])
    p[:, :3] = cell + np.minimum(np.float32(p[:, :3] - cell).astype(np.float64), 0.99999994)
    return p


def sort_after(p, steps):
    """(tile sorts, radix sorts) of the sort after `steps` steps from the
    cluster's radix sort, and whether its result is the oracle's stable sort
    of the same particles."""
    c = cfg(1, tile=(2, 2, 2))
    c.sort_every, c.inflow_every = 1000, 1000
    g = ps.Simulation(c)
    g.upload_particles(1, p)
    g.sort(1)  # uploaded, unsorted: radix sort; the tiles are current after it
    g.step(steps)
    pre = g.download_particles(1)
    t0, r0 = g.sort_stats()
    g.sort(1)
    t1, r1 = g.sort_stats()
    o = Oracle(cfg(1, precision=ps.PIC_FP64, tile=(2, 2, 2)))
    o.upload_particles(1, pre)
    o.sort(1)
    same = np.array_equal(g.download_particles(1), o.download_particles(1))
    g.close()
    return (t1 - t0, r1 - r0), same


def test_tile_sort_takes_a_crowd_of_movers():
    """The cluster moves at 0.9 c along z for three steps: all 20000 enter
    the next tile.  The movers share one pool sized by the capacity, so the
    tile sort takes them, and the result is the oracle's stable sort."""
    p = cluster(9, lambda rng, n: np.column_stack([np.zeros(n), np.zeros(n),
                                                   np.full(n, 0.9)]))
    counts, same = sort_after(p, 3)
    assert counts == (1, 0) and same


def test_tile_sort_falls_back_when_a_block_scatters():
    """The cluster moves in random directions at up to 0.9 c for 16 steps:
    its tile's particles reach more destination tiles than a source block's
    table holds (64), so the sort falls back to the radix sort, and the
    result is still the oracle's stable sort."""
    def vel(rng, n):
        d = rng.normal(size=(n, 3))
        return 0.9 * d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.5, 1, (n, 1))
    counts, same = sort_after(cluster(11, vel), 16)
    assert counts == (0, 1) and same


def sims_with_state(f, p=None):
    out = []
    for prec in (ps.PIC_FP32, ps.PIC_FP64):
        s = ps.Simulation(cfg(1, precision=prec))
        for n, a in f.items():
            s.upload_field(n, a)
        if p is not None:
            for sp in (0, 1):
                s.upload_particles(sp, p[sp])
        out.append(s)
    return out


@pytest.mark.parametrize("routine", ["renew_b_half", "renew_e", "halo_b", "halo_e",
                                     "fold_current"])
def test_field_routine_vs_fp64(routine):
    c = cfg(1, precision=ps.PIC_FP64)
    sh = (c.nx + 4, c.ny + 4, c.nz + 4)
    rng = np.random.default_rng(21)
    probe = ps.Simulation(c)
    b0 = [probe.download_field(n) for n in ("b0x", "b0y", "b0z")]
    f = {n: rng.standard_normal(sh) * 1e-2 for n in ("ex", "ey", "ez")}
    for i, n in enumerate(("bx", "by", "bz")):
        f[n] = b0[i] + rng.standard_normal(sh) * 1e-2
    for n in J:
        f[n] = rng.standard_normal(sh) * 1e-3
    s32, s64 = sims_with_state(f)
    getattr(s32, routine)()
    getattr(s64, routine)()
    for n in ("ex", "ey", "ez") + tuple(J):
        e = rel_l2(s32.download_field(n), s64.download_field(n))
        assert e < 3e-7, (routine, n, e)
    for i, n in enumerate(("bx", "by", "bz")):
        e = rel_l2(s32.download_field(n) - b0[i], s64.download_field(n) - b0[i])
        assert e < 3e-7, (routine, n, e)


@pytest.mark.parametrize("sp", [0, 1])
@pytest.mark.parametrize("routine", ["push_deposit_j", "drift_deposit_n"])
def test_particle_routine_vs_fp64(state, routine, sp):
    f, p = state
    s32, s64 = sims_with_state(f, p)
    for s in (s32, s64):
        s.renew_grid()
        getattr(s, routine)(sp)
    errs = {}
    for n in J:
        a, b = s32.download_field(n), s64.download_field(n)
        if np.linalg.norm(b) == 0:
            continue
        errs[n] = rel_l2(a, b)
    print(routine, sp, {k: f"{v:.1e}" for k, v in errs.items()})
    for n, e in errs.items():
        assert e < 1e-5, (routine, sp, n, e)
    a, b = s32.download_particles(sp), s64.download_particles(sp)
    assert a.shape == b.shape
    assert np.abs(a[:, :3] - b[:, :3]).max() < 1e-5
    assert rel_l2(a[:, 3:], b[:, 3:]) < 1e-6



@pytest.mark.parametrize("sp", [0, 1])
def test_ballistic_bounds_on_every_axis_vs_fp64(state, sp):
    """The default workloads bound the ballistic zone (no gather, no push)
    on the flow axis only; here its x and y bounds lie inside the array too,
    so the pass tests all three axes.  One push against the FP64 form: J to
    1e-5, particles as in the push test, and the particles inside the x/y
    bounds' complement keep their velocities exactly."""
    f, p = state
    sims = []
    for prec in (ps.PIC_FP32, ps.PIC_FP64):
        c = cfg(1, precision=prec)
        c.ballistic_lo[0], c.ballistic_hi[0] = 9.5 * c.dx, (c.nx - 5.5) * c.dx
        c.ballistic_lo[1], c.ballistic_hi[1] = 7.25 * c.dx, (c.ny - 3.75) * c.dx
        s = ps.Simulation(c)
        for n, a in f.items():
            s.upload_field(n, a)
        for q in (0, 1):
            s.upload_particles(q, p[q])
        s.renew_grid()
        s.push_deposit_j(sp)
        sims.append(s)
    s32, s64 = sims
    for n in J:
        a, b = s32.download_field(n), s64.download_field(n)
        if np.linalg.norm(b) == 0:
            continue
        assert rel_l2(a, b) < 1e-5, (sp, n)
    a, b = s32.download_particles(sp), s64.download_particles(sp)
    assert a.shape == b.shape
    assert np.abs(a[:, :3] - b[:, :3]).max() < 1e-5
    assert rel_l2(a[:, 3:], b[:, 3:]) < 1e-6
    c = cfg(1)
    x, y = p[sp][:, 0], p[sp][:, 1]
    out = ((x < 9.5 * c.dx) | (x > (c.nx - 5.5) * c.dx) | (y < 7.25 * c.dx) |
           (y > (c.ny - 3.75) * c.dx))
    assert 0.2 < out.mean() < 0.9
    # the FP64 form keeps their velocities exactly; the FP32 form to fp32
    # rounding of the uploaded values
    assert np.array_equal(b[out, 3:], p[sp][out, 3:])
    assert np.abs(a[out, 3:] - p[sp][out, 3:]).max() <= 1e-7 * np.abs(p[sp][:, 3:]).max()

def developed_state(steps=20):
    """Oracle state after `steps` FP64 steps of the 32^3 workload."""
    c64 = cfg(1, precision=ps.PIC_FP64)
    o = Oracle(c64)
    o.load_particles()
    o.begin()
    o.step(steps)
    f = {n: o.download_field(n) for n in FIELDS}
    p = [o.download_particles(sp) for sp in (0, 1)]
    return f, p


@pytest.fixture(scope="module")
def state():
    return developed_state()


def run_one(precision, f, p, deposit=ps.PIC_DEPOSIT_ATOMIC):
    c = cfg(1, precision=precision, deposit=deposit)
    s = ps.Simulation(c)
    for n, a in f.items():
        s.upload_field(n, a)
    for sp in (0, 1):
        s.upload_particles(sp, p[sp])
    s.step(1)
    out = {n: s.download_field(n) for n in FIELDS + J}
    out["b0"] = [s.download_field(n) for n in ("b0x", "b0y", "b0z")]
    out["count"] = [s.particle_count(sp) for sp in (0, 1)]
    s.close()
    return out


def test_one_step_error_vs_fp64(state):
    f, p = state
    r64 = run_one(ps.PIC_FP64, f, p)
    r32 = run_one(ps.PIC_FP32, f, p)
    errs = {}
    for n in J + ["ex", "ey", "ez"]:
        errs[n] = rel_l2(r32[n], r64[n])
    for i, n in enumerate(("bx", "by", "bz")):
        # induced field against the exact B0 (the FP32 form's B0 plane is a
        # rounded copy of the same host B0)
        errs[n + "-b0"] = rel_l2(r32[n] - r64["b0"][i], r64[n] - r64["b0"][i])
        errs[n] = rel_l2(r32[n], r64[n])
    print("FP32 vs FP64 one-step relative L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v <= 1e-5, (k, v)
    for sp in (0, 1):
        assert abs(r32["count"][sp] - r64["count"][sp]) <= 2


def test_deterministic_mode_is_bit_reproducible(state):
    f, p = state
    a = run_one(ps.PIC_FP32, f, p, deposit=ps.PIC_DEPOSIT_DETERMINISTIC)
    b = run_one(ps.PIC_FP32, f, p, deposit=ps.PIC_DEPOSIT_DETERMINISTIC)
    for n in FIELDS + J:
        assert np.array_equal(a[n], b[n]), n
    r64 = run_one(ps.PIC_FP64, f, p)
    for n in J:
        assert rel_l2(a[n], r64[n]) <= 1e-5, n


def test_sort_applied_by_the_pass_matches_the_copying_sort():
    """pic_step's FP32 tile sort leaves its permutation to the species' next
    fused pass, which reads the particles through it, and when the sort and
    the inflow fall on one step the inflow runs first.  In the deterministic
    deposit mode (fixed-point sums, independent of the particles' order and
    of the windows) a run with tiles (every tile sort applied by the pass)
    and a run with one tile (every sort the radix sort and its copy, the
    plain pass) give bit-identical fields and the same particles, over steps
    with a sort only, an inflow only, and both."""
    out = []
    for tile in ((4, 4, 8), (32, 32, 32)):
        c = cfg(1, tile=tile, deposit=ps.PIC_DEPOSIT_DETERMINISTIC)
        c.sort_every, c.inflow_every = 2, 4
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(9)
        parts = []
        for sp in (0, 1):
            p = s.download_particles(sp)
            parts.append(p[np.lexsort(p[:, ::-1].T)])
        out.append(({n: s.download_field(n) for n in FIELDS + J}, parts, s.sort_stats()))
        s.close()
    (fa, pa, sa), (fb, pb, sb) = out
    assert sa[0] >= 4 and sb[0] == 0, (sa, sb)
    for n in FIELDS + J:
        assert np.array_equal(fa[n], fb[n]), n
    for sp in (0, 1):
        assert pa[sp].shape == pb[sp].shape and pa[sp].shape[0] > 0, sp
        assert np.array_equal(pa[sp], pb[sp]), sp


def test_sort_applied_by_the_pass_at_the_128cubed_workload():
    """The 128^3 x 64 ppc workload (268,435,456 particles) in the
    deterministic deposit mode: 21 steps of pic_step, whose step 21 runs the
    inflow and then the tile sort that the particle pass applies, against
    20 steps, the C-ABI sort of both species (which copies the particles)
    and a 21st step with the inflow only.  Every field plane bit-identical,
    the counts equal, and the FP64 sums over the particles (which sit in
    other slots) equal to their summation rounding (relative 1e-12)."""
    out = []
    for in_step in (True, False):
        c = ps.preset(2)
        c.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
        if not in_step:
            c.sort_every = 1000
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        if in_step:
            s.step(21)
        else:
            s.step(20)
            s.sort(0)
            s.sort(1)
            s.step(1)
        out.append(({n: s.download_field(n) for n in FIELDS + J}, s.diagnose(),
                    s.sort_stats()))
        s.close()
    (fa, da, sa), (fb, db, sb) = out
    assert sa == sb and sa[0] == 2, (sa, sb)
    for n in FIELDS + J:
        assert np.array_equal(fa[n], fb[n]), n
    assert np.array_equal(da["count"], db["count"])
    for k in ("kinetic", "kinetic_nr"):
        assert np.allclose(da[k], db[k], rtol=1e-12, atol=0), (k, da[k], db[k])
    scale = np.abs(da["momentum"]).max()
    assert np.allclose(da["momentum"], db["momentum"], rtol=0, atol=1e-12 * scale)


def test_multi_step_runs_with_sort_and_inflow():
    c = cfg(1)
    s = ps.Simulation(c)
    s.load_particles()
    s.begin()
    n0 = [s.particle_count(sp) for sp in (0, 1)]
    s.step(45)
    d = s.diagnose()
    for sp in (0, 1):
        n1 = s.particle_count(sp)
        assert abs(n1 - n0[sp]) < 0.05 * n0[sp]
    assert np.isfinite(d["field_energy_e"]) and d["field_energy_e"] > 0


def test_config5_largest_workload_runs():
    """256^3 x 32 ppc x 2 species (1,073,741,824 particles, m_i/m_e = 400,
    drift 0.05, dipole tilted 30 degrees) in the FP32 form: load, pre-loop
    inflow and 21 steps (an inflow every 5, the sort at step 20) keep the
    particle counts within the inflow/outflow balance and every diagnostic
    finite."""
    import time
    c = ps.preset(5)
    s = ps.Simulation(c)
    s.load_particles()
    s.begin()
    n0 = [s.particle_count(sp) for sp in (0, 1)]
    assert n0[0] == n0[1] >= 256**3 * 32
    s.step(1)
    s.sync()
    t0 = time.perf_counter()
    s.step(20)
    s.sync()
    dt = (time.perf_counter() - t0) / 20
    d = s.diagnose()
    print(f"config 5: {sum(n0)} particles, {dt * 1e3:.1f} ms per step, "
          f"{sum(n0) / dt:.3e} particle-steps/s")
    for k in ("field_energy_e", "field_energy_b1", "charge", "gauss_l2"):
        assert np.isfinite(d[k]), k
    assert np.all(np.isfinite(d["kinetic"]))
    for sp in (0, 1):
        assert abs(d["count"][sp] - n0[sp]) < 0.02 * n0[sp]
    s.close()


def test_config5_deterministic_runs_bit_identical():
    """256^3 x 32 ppc x 2 species (1,073,741,824 particles) in the
    deterministic sorted-tile deposit mode: two runs of 21 steps (inflows at
    steps 5, 10, 15 and 20, the tile sort at step 20) end with every field
    plane bit-identical, the FP64 diagnostics (sums over all particles, so
    any differing particle bit would show) identical and the counts equal."""
    import hashlib
    out = []
    for run in range(2):
        c = ps.preset(5)
        c.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(21)
        h = {n: hashlib.sha256(s.download_field(n).tobytes()).hexdigest()
             for n in ("ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex",
                       "jey", "jez", "ni", "ne")}
        d = s.diagnose()
        stats = s.sort_stats()
        s.close()
        out.append((h, d, stats))
    (ha, da, sa), (hb, db, sb) = out
    assert ha == hb
    for k in da:
        assert np.array_equal(np.asarray(da[k]), np.asarray(db[k])), k
    assert sa == sb and sa[0] >= 2  # both species sorted by the tile sort at step 20


def drifted_state(deposit, steps=8, seed=17):
    """FP32 state in which a few percent of particles have left the LDS window
    of the tile they were sorted into: a thermal population (0.05 c) with 2 %
    of particles, and 40 % of those in the first tile column, moving at 0.9 c;
    sorted once, then stepped without re-sorting.  Returns the simulation and
    the fraction of particles whose cell lies more than one cell outside their
    tile (the window's valid box)."""
    c = cfg(1, deposit=deposit)
    c.sort_every = c.inflow_every = 1000
    s = ps.Simulation(c)
    rng = np.random.default_rng(seed)
    n = c.nx * c.ny * c.nz * 8
    for sp in (0, 1):
        p = np.column_stack([rng.uniform(2, c.nx + 2, n), rng.uniform(2, c.ny + 2, n),
                             rng.uniform(6, c.nz - 2, n), rng.standard_normal((n, 3)) * 0.05])
        fast = (rng.random(n) < 0.02) | ((p[:, 0] < 6) & (rng.random(n) < 0.4))
        p[fast, 3:] = 0.52 * rng.choice([-1.0, 1.0], (int(fast.sum()), 3))
        s.upload_particles(sp, p)
        s.sort(sp)
    tiles = [np.floor(s.download_particles(sp)[:, :3]) for sp in (0, 1)]
    s.step(steps)
    out = []
    for sp in (0, 1):
        cell = np.floor(s.download_particles(sp)[:, :3])
        assert cell.shape == tiles[sp].shape  # nobody left: same slots
        t = np.asarray(c.tile[:3])
        lo = 2 + ((tiles[sp] - 2) // t) * t
        out.append(np.mean(np.any((cell < lo - 1) | (cell > lo + t), axis=1)))
    return s, c, out


@pytest.mark.parametrize("deposit", [ps.PIC_DEPOSIT_ATOMIC, ps.PIC_DEPOSIT_DETERMINISTIC])
def test_out_of_window_particles_match_fp64(deposit):
    """Particles outside their tile's LDS window deposit straight into the
    global planes (one at a time by their whole wave, or each by its own lane
    where many share a wave): one step from such a state agrees with the FP64
    form to the per-step bar (relative L2 <= 1e-5), and in deterministic mode
    is bit-identical run to run."""
    s32, c, frac = drifted_state(deposit)
    print("out-of-window fraction (ion, electron):", frac)
    assert all(0.003 < f < 0.2 for f in frac), frac
    c64 = c.copy()
    c64.precision = ps.PIC_FP64
    s64 = ps.Simulation(c64)
    for n in FIELDS:
        s64.upload_field(n, s32.download_field(n))
    for sp in (0, 1):
        s64.upload_particles(sp, s32.download_particles(sp))
    s64.set_counters(*s32.counters())
    runs = [s32]
    if deposit == ps.PIC_DEPOSIT_DETERMINISTIC:
        runs.append(drifted_state(deposit)[0])
    for s in runs + [s64]:
        s.step(1)
    b0 = [s64.download_field(n) for n in ("b0x", "b0y", "b0z")]
    for n in J + ["ex", "ey", "ez"]:
        e = rel_l2(s32.download_field(n), s64.download_field(n))
        assert e <= 1e-5, (n, e)
    for i, n in enumerate(("bx", "by", "bz")):
        e = rel_l2(s32.download_field(n) - b0[i], s64.download_field(n) - b0[i])
        assert e <= 1e-5, (n, e)
    if len(runs) == 2:
        for n in FIELDS + J:
            assert np.array_equal(runs[0].download_field(n), runs[1].download_field(n)), n


@pytest.mark.parametrize("preset", [1, 2])
def test_inflow_batch_matches_fp64_as_set(preset):
    """The FP32 form stores an inflow batch in blocks of 4 x 4 x 2 cells; each
    particle is still drawn from the counter of its reference-order index, so
    the batch is the FP64 form's batch as a set: cells exact, offsets and
    velocities within fp32 rounding."""
    out = {}
    for prec in (ps.PIC_FP32, ps.PIC_FP64):
        c = cfg(preset, precision=prec)
        if preset == 2:
            c.nx = c.ny = 32  # a 32 x 32 face of the 128-cell-deep workload
        s = ps.Simulation(c)
        s.inflow()
        out[prec] = [s.download_particles(sp) for sp in (0, 1)]
        s.close()
    for sp in (0, 1):
        a, b = out[ps.PIC_FP32][sp], out[ps.PIC_FP64][sp]
        assert a.shape == b.shape and len(a) > 0
        a = a[np.lexsort(np.column_stack([np.floor(a[:, :3]), a[:, 3:]])[:, ::-1].T)]
        b = b[np.lexsort(np.column_stack([np.floor(b[:, :3]), b[:, 3:]])[:, ::-1].T)]
        assert np.array_equal(np.floor(a[:, :3]), np.floor(b[:, :3])), sp
        assert np.abs(a[:, :3] - b[:, :3]).max() < 1.2e-7 * 2
        assert np.abs(a[:, 3:] - b[:, 3:]).max() <= 6e-8 * np.abs(b[:, 3:]).max()


@pytest.mark.parametrize("deposit", [ps.PIC_DEPOSIT_ATOMIC, ps.PIC_DEPOSIT_DETERMINISTIC])
def test_inflow_into_dead_slots_matches_fp64_as_set(deposit):
    """With current tiles the FP32 form writes an inflow batch into the dead
    slots of the slab tiles it belongs to (appending only what does not
    fit): after load, sort and an inflow the live particles are the FP64
    form's (which appends the batch) as a set, the same in both deposit
    modes, and a second identical run places them in the same slots."""
    out = {}
    raws = []
    for prec, dep in ((ps.PIC_FP32, deposit), (ps.PIC_FP64, ps.PIC_DEPOSIT_ATOMIC),
                      (ps.PIC_FP32, deposit)):
        s = ps.Simulation(cfg(1, precision=prec, deposit=dep))
        s.load_particles()
        s.sort(0)
        s.sort(1)
        n_sorted = [s.particle_count(sp) for sp in (0, 1)]
        s.inflow()
        if prec == ps.PIC_FP32:
            raws.append([s.download_raw32(sp) for sp in (0, 1)])
        out.setdefault(prec, ([s.download_particles(sp) for sp in (0, 1)], n_sorted))
        s.close()
    for sp in (0, 1):
        a, b = out[ps.PIC_FP32][0][sp], out[ps.PIC_FP64][0][sp]
        assert a.shape == b.shape
        # every cell holds the same number of particles
        ca, cb = (np.unique(np.floor(x[:, :3]), axis=0, return_counts=True) for x in (a, b))
        assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1]), sp
        for j in range(6):
            tol = 2.4e-7 if j < 3 else 6e-8 * np.abs(b[:, j]).max()
            assert np.abs(np.sort(a[:, j]) - np.sort(b[:, j])).max() <= tol, (sp, j)
        for x, y in zip(raws[0][sp], raws[1][sp]):
            assert np.array_equal(x, y), sp


@pytest.mark.parametrize("tile", [(4, 4, 0), (4, 4, 32), (8, 4, 4), (8, 4, 16)])
def test_inflow_set_for_other_tile_shapes(tile):
    """The inflow batch after a sort is the FP64 form's as a set for tilings
    with one z tile layer (both faces' slabs then lie in the same tiles: the
    batch is appended, not written into dead slots by two workgroups at
    once) and for the 8 x 4 x 4 tiles of the 128^3 workload (tiles longer in
    x than in y), with both faces refilled.  Counts per cell are exact,
    positions agree to fp32 rounding of the offsets."""
    out = {}
    for prec in (ps.PIC_FP32, ps.PIC_FP64):
        c = cfg(1, precision=prec, tile=tile)
        c.refill_far = 1
        s = ps.Simulation(c)
        s.load_particles()
        # a few steps first, so that particles left the slabs and the
        # boundary marked dead slots, then the sort and two inflow events
        s.begin()
        s.step(3)
        s.sort(0)
        s.sort(1)
        s.inflow()
        s.inflow()
        out[prec] = [s.download_particles(sp) for sp in (0, 1)]
        s.close()
    for sp in (0, 1):
        a, b = out[ps.PIC_FP32][sp], out[ps.PIC_FP64][sp]
        assert a.shape == b.shape, (tile, sp)
        # the batch particles (slab layers z in [2, 4) and [NZ, NZ + 2))
        # carry no push and match per cell exactly
        for x in (a, b):
            assert np.isfinite(x).all()
        za, zb = a[:, 2], b[:, 2]
        sa = (za < 4) | (za >= 32)
        sb = (zb < 4) | (zb >= 32)
        assert sa.sum() == sb.sum(), (tile, sp)
        ca, cb = (np.unique(np.floor(x[m, :3]), axis=0, return_counts=True)
                  for x, m in ((a, sa), (b, sb)))
        assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1]), (tile, sp)
        for j in range(6):
            tol = 2.4e-7 if j < 3 else 6e-8 * np.abs(b[sb, j]).max()
            assert np.abs(np.sort(a[sa, j]) - np.sort(b[sb, j])).max() <= tol, (tile, sp, j)
