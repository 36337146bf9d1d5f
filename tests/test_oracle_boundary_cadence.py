"""Pin the oracle's particle boundary and the step cadences (CPU).

The reference has no tests or fixtures for these parts, so they are pinned
by what its code fixes, written here from the reference's text rather than
from the oracle's:

- sr_particle (PIC/3d_pic458_domain.c:2041-2253) keeps a particle only if
  2dx <= z < (Nz+2)dx; a kept particle with x >= (Nx+2)dx goes to the next
  rank, which stores x - Nx dx (sr_particle_x :2372-2376), one with x < 2dx
  is stored with x + Nx dx; then y likewise (sr_particle_y).  On one rank
  that is the periodic image: x' = x (mod Nx dx) with 2dx <= x' < (Nx+2)dx.
- fdtd (:661-720, :943-948): add_*_flat once before the loop; at the top of
  step c the inflow runs when c_add > 4 and the sort when c_thrust > 499,
  each counter reset there and incremented at the bottom of every step.
  Steps are numbered from 0, so the inflow runs at steps 5, 10, 15, ...
  and the sort at steps S, 2S, ... for a cadence S (the reference's 500).
- add_ion_flat (:6016-6024, :6036-6060): delete z < 4dx or z > Nz dx, then
  add Np particles per cell in the two cell layers z in [2dx, 4dx).

Each expectation is a closed form or a coarse model of those rules, so a
wrong shift, an inclusive z bound, a dropped non-periodic crosser or a
cadence off by one fails here even though the GPU parity tests (which
compare with the oracle) would not see it.
"""
import numpy as np
import pytest

import plasmasim_amd as ps
from oracle_lib import Oracle


def quiet_cfg(nx=8, ny=6, nz=10, dx=1.5, ppc=2):
    """No forces: q = 0, no dipole, no IMF, no absorber; particles move on
    straight lines and deposit nothing."""
    c = ps.preset(0)
    c.nx, c.ny, c.nz = nx, ny, nz
    c.ppc = ppc
    c.dx = dx
    c.dt = 0.5
    c.c = 10.0
    c.q = 0.0
    c.dipole_m = 0.0
    c.imf[0] = c.imf[1] = c.imf[2] = 0.0
    c.absorb[0] = c.absorb[1] = c.absorb[2] = 0
    c.sigma_amp = 0.0
    c.vth[0] = c.vth[1] = 0.0
    c.vdrift[0] = c.vdrift[1] = c.vdrift[2] = 0.0
    c.periodic[0] = c.periodic[1] = 1
    c.inflow_every = 5
    c.sort_every = 1000
    c.refill_far = 1
    c.poisson_every = 0
    for a in range(3):
        c.ballistic_lo[a], c.ballistic_hi[a] = -1e30, 1e30
    return c


def ids(parts, unit=1e-4, base=0.0):
    """Particles are tagged by a unique vx = base + unit * id (the push may
    round vx by an ulp, so the tag is rounded back to an integer)."""
    return np.rint((parts[:, 3] - base) / unit).astype(np.int64)


# ------------------------------------------------------- sr_particle pins
def boundary_case(periodic=(1, 1)):
    """dx = 1.5, Nx = 8, Ny = 6, Nz = 10: x in [3, 15) (period 12),
    y in [3, 12) (period 9), z kept in [3, 18)."""
    c = quiet_cfg()
    c.periodic[0], c.periodic[1] = periodic
    o = Oracle(c)
    # (x, y, z) -> expected (x', y') or None when dropped; values written out
    # by hand from the rules in the module docstring
    mid = 7.25
    cases = [
        ((15.0, mid, mid), (3.0, mid)),          # x = (Nx+2)dx wraps to 2dx
        ((14.75, mid, mid), (14.75, mid)),       # just inside: unchanged
        ((3.0, mid, mid), (3.0, mid)),           # x = 2dx stays
        ((2.75, mid, mid), (14.75, mid)),        # x < 2dx: + Nx dx
        ((16.5, mid, mid), (4.5, mid)),          # a cell beyond the face
        ((0.25, mid, mid), (12.25, mid)),        # deep in the low guard
        ((mid, 12.0, mid), (mid, 3.0)),          # y = (Ny+2)dx wraps to 2dx
        ((mid, 11.75, mid), (mid, 11.75)),
        ((mid, 2.75, mid), (mid, 11.75)),        # y < 2dx: + Ny dx
        ((mid, 13.25, mid), (mid, 4.25)),
        ((15.5, 2.5, mid), (3.5, 11.5)),         # both axes at once
        ((2.5, 12.5, mid), (14.5, 3.5)),
        ((mid, mid, 18.0), None),                # z = (Nz+2)dx is dropped
        ((mid, mid, 17.75), (mid, mid)),         # just below: kept
        ((mid, mid, 3.0), (mid, mid)),           # z = 2dx is kept
        ((mid, mid, 2.75), None),                # z < 2dx is dropped
        ((15.0, mid, 18.5), None),               # z decides before any wrap
        ((2.0, 2.0, 1.0), None),
    ]
    parts = []
    for i, (pos, _) in enumerate(cases):
        parts.append([pos[0], pos[1], pos[2], 0.001 * (i + 1), 0.0, 0.0])
    o.upload_particles(0, np.array(parts))
    o.upload_particles(1, np.array(parts))
    o.particle_boundary()
    return o, cases


def test_boundary_periodic_wrap_and_z_window():
    o, cases = boundary_case()
    for sp in (0, 1):
        got = o.download_particles(sp)
        by_id = {t: row for t, row in zip(ids(got, 1e-3), got)}
        for i, (pos, exp) in enumerate(cases):
            key = i + 1
            if exp is None:
                assert key not in by_id, f"case {i} {pos} should be dropped"
                continue
            assert key in by_id, f"case {i} {pos} was dropped"
            row = by_id[key]
            assert abs(row[0] - exp[0]) < 1e-12, (i, pos, row[0], exp)
            assert abs(row[1] - exp[1]) < 1e-12, (i, pos, row[1], exp)
            assert row[2] == pos[2]
        assert len(got) == sum(e is not None for _, e in cases)


@pytest.mark.parametrize("periodic", [(0, 1), (1, 0), (0, 0)])
def test_boundary_drops_crossers_on_non_periodic_axes(periodic):
    """At the reference's global edges (rank 0 / rank P-1) the crossers of
    that axis are discarded instead of sent (:2138-2148, :2235-2245)."""
    o, cases = boundary_case(periodic)
    got = o.download_particles(0)
    kept = {t - 1 for t in ids(got, 1e-3)}
    for i, (pos, exp) in enumerate(cases):
        crosses_x = not (3.0 <= pos[0] < 15.0)
        crosses_y = not (3.0 <= pos[1] < 12.0)
        alive = exp is not None and not (crosses_x and not periodic[0]) \
            and not (crosses_y and not periodic[1])
        assert (i in kept) == alive, (i, pos, periodic)


def test_boundary_random_particles_are_periodic_images():
    """Random positions across the guard-inclusive box: survivors are the
    ones inside the z window, each x (y) moved by a multiple of the period
    into [2dx, (N+2)dx), and nothing else changes."""
    c = quiet_cfg()
    o = Oracle(c)
    rng = np.random.default_rng(7)
    n = 4000
    dx = c.dx
    p = np.empty((n, 6))
    p[:, 0] = rng.uniform(dx, (c.nx + 3) * dx, n)
    p[:, 1] = rng.uniform(dx, (c.ny + 3) * dx, n)
    p[:, 2] = rng.uniform(dx, (c.nz + 3) * dx, n)
    p[:, 3] = np.arange(1, n + 1) * 1e-4
    p[:, 4] = rng.standard_normal(n)
    p[:, 5] = rng.standard_normal(n)
    o.upload_particles(0, p)
    o.upload_particles(1, p[::-1].copy())
    o.particle_boundary()
    alive = (p[:, 2] >= 2 * dx) & (p[:, 2] < (c.nz + 2) * dx)
    for sp, src in ((0, p), (1, p[::-1])):
        got = o.download_particles(sp)
        a = (src[:, 2] >= 2 * dx) & (src[:, 2] < (c.nz + 2) * dx)
        assert len(got) == int(alive.sum())
        ref = src[a]
        # order of the survivors is kept (in-place wrap, no migration)
        np.testing.assert_array_equal(ids(got), ids(ref))
        for ax, N in ((0, c.nx), (1, c.ny)):
            L = N * dx
            shift = (got[:, ax] - ref[:, ax]) / L
            assert np.all(np.abs(shift - np.round(shift)) < 1e-12)
            assert np.all(np.isin(np.round(shift), (-1.0, 0.0, 1.0)))
            assert np.all(got[:, ax] >= 2 * dx) and np.all(got[:, ax] < (N + 2) * dx)
            inside = (ref[:, ax] >= 2 * dx) & (ref[:, ax] < (N + 2) * dx)
            assert np.all(got[inside, ax] == ref[inside, ax])
        np.testing.assert_array_equal(got[:, 2:], ref[:, 2:])


# ------------------------------------------------------------ cadences
def inflows_after(k, every=5):
    """Inflow events after k steps: one before the loop, then one at each
    step index s >= 1 with s % every == 0 (c_add > every - 1 at the top of
    step s, c_add counting steps since the last reset)."""
    return 1 + (k - 1) // every if k > 0 else 1


@pytest.mark.parametrize("every", [5, 3])
def test_inflow_events_follow_the_reference_cadence(every):
    """Particles at rest: after k steps the state equals a fresh run's state
    after begin() and (inflows_after(k) - 1) more inflow calls, so the k-step
    run did exactly that many inflows, each drawing the next Philox event."""
    c = quiet_cfg(nx=4, ny=4, nz=10, dx=1.0, ppc=2)
    c.inflow_every = every
    a = Oracle(c)
    a.load_particles()
    a.begin()
    done = 0
    for k in range(1, 17):
        a.step(1)
        done += 1
        b = Oracle(c)
        b.load_particles()
        b.begin()
        for _ in range(inflows_after(done, every) - 1):
            b.inflow()
        for sp in (0, 1):
            np.testing.assert_array_equal(a.download_particles(sp),
                                          b.download_particles(sp),
                                          err_msg=f"after {done} steps")
        b.close()


def layer_model(nz, steps, every, refill_far):
    """Occupied z cell layers (m = 0..nz-1) when every particle moves one
    cell downstream per step: the reference's order within a step is inflow
    (when due), push and drift, then sr_particle."""
    occ = set(range(nz))

    def inflow(occ):
        occ = {m for m in occ if 2 <= m < nz - 2}  # z < 4dx or z > Nz dx
        occ |= {0, 1}
        if refill_far:
            occ |= {nz - 2, nz - 1}
        return occ

    occ = inflow(occ)  # before the loop
    counts = []
    c_add = 0
    for _ in range(steps):
        if c_add > every - 1:
            occ = inflow(occ)
            c_add = 0
        occ = {m + 1 for m in occ}
        occ = {m for m in occ if m < nz}  # z >= (Nz+2)dx leaves
        c_add += 1
        counts.append(len(occ))
    return counts


@pytest.mark.parametrize("refill_far", [0, 1])
def test_particle_counts_follow_outflow_and_inflow_cadence(refill_far):
    """Cold beam along z at one cell per step (v dt = dx): the live count
    after each step is (occupied layers) x Nx Ny ppc."""
    c = quiet_cfg(nx=2, ny=2, nz=12, dx=1.0, ppc=3)
    c.vdrift[2] = 2.0  # dt = 0.5: two half drifts of 0.5 cell
    c.refill_far = refill_far
    o = Oracle(c)
    o.load_particles()
    o.begin()
    per_layer = c.nx * c.ny * c.ppc
    model = layer_model(c.nz, 24, c.inflow_every, refill_far)
    for k in range(24):
        o.step(1)
        for sp in (0, 1):
            assert o.particle_count(sp) == model[k] * per_layer, (k, sp)


def test_sort_runs_at_the_top_of_step_S_and_every_S_after():
    """Particles uploaded in a shuffled order: the stored order is the
    upload order until the first sort, which runs at the top of step S (the
    step that sees c_thrust > S - 1), and is then the stable key order of
    the positions at that moment; the next sort comes S steps later."""
    S = 3
    c = quiet_cfg(nx=4, ny=4, nz=8, dx=1.0)
    c.sort_every = S
    c.inflow_every = 10 ** 6
    c.tile[0] = c.tile[1] = c.tile[2] = 0  # the reference's (k,l,m) key
    o = Oracle(c)
    rng = np.random.default_rng(3)
    n = 600
    p = np.empty((n, 6))
    p[:, 0] = rng.uniform(2.0, 6.0, n)
    p[:, 1] = rng.uniform(2.0, 6.0, n)
    p[:, 2] = rng.uniform(4.0, 6.0, n)
    p[:, 3] = (np.arange(n) + 1) * 1e-4 + 0.5  # id and an x speed
    p[:, 4] = rng.uniform(-0.6, 0.6, n)
    p[:, 5] = rng.uniform(-0.1, 0.1, n)
    for sp in (0, 1):
        o.upload_particles(sp, p)

    def key(q):
        k = np.floor(q[:, 0]).astype(int) - 2
        l = np.floor(q[:, 1]).astype(int) - 2
        m = np.floor(q[:, 2]).astype(int) - 2
        return (k * c.ny + l) * c.nz + m

    order = ids(p, base=0.5)
    at_sort = None
    for s in range(2 * S + 2):
        before = o.download_particles(0)
        if s % S == 0 and s > 0:
            at_sort = before
            order = ids(before, base=0.5)[np.argsort(key(before), kind="stable")]
        o.step(1)
        got = o.download_particles(0)
        np.testing.assert_array_equal(ids(got, base=0.5), order,
                                      err_msg=f"after step {s}")
        if s == S:
            # the sort moved particles (the upload order was shuffled)
            assert not np.array_equal(ids(at_sort, base=0.5), order)
