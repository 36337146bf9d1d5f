"""Pin the CPU oracle against known answers and closed forms.

The reference ships no tests, fixtures or golden vectors, so each oracle
function is checked against what the mathematics fixes: published Philox
known-answer vectors, exactness of trilinear interpolation, the closed-form
rotation of the reference's Boris variant, the (2,4) leapfrog dispersion
relation, the sigma damping factors, brute-force deposits/sorts written
independently of the oracle's index arithmetic, the point dipole (its
softened core included) and closed-form diagnostics.
"""
import math

import numpy as np
import pytest

import plasmasim_amd as ps
from oracle_lib import Oracle, boris, philox


def cfg_fields(nx=8, ny=8, nz=16):
    c = ps.preset(0)
    c.nx, c.ny, c.nz = nx, ny, nz
    c.dipole_m = 0.0
    c.absorb[0] = c.absorb[1] = c.absorb[2] = 0
    c.ballistic_lo[2], c.ballistic_hi[2] = -1e30, 1e30
    return c


# ---------------------------------------------------------------- Philox
# Random123 known-answer vectors for philox4x32-10 (ctr, key -> out).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2,
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(philox(ctr, key)) == out


# ---------------------------------------------------------------- gather
def hat(d):
    return np.maximum(0.0, 1.0 - np.abs(d))


# staggering offsets (x, y, z) of each gathered component, Appendix of the
# reference's shape_func_v2 (:3104-3139): ex ey ez bx by bz
OFFS = {"ex": (0, .5, .5), "ey": (.5, 0, .5), "ez": (.5, .5, 0),
        "bx": (.5, 0, 0), "by": (0, .5, 0), "bz": (0, 0, .5)}
ORDER = ["ex", "ey", "ez", "bx", "by", "bz"]


def hat_interp(F, off, x, y, z):
    """Independent CIC interpolation: sum over nodes of F * hat products."""
    S = F.shape
    k = np.arange(S[0])[:, None, None] + off[0]
    l = np.arange(S[1])[None, :, None] + off[1]
    m = np.arange(S[2])[None, None, :] + off[2]
    w = hat(x - k) * hat(y - l) * hat(z - m)
    return float((F * w).sum())


def test_gather_matches_independent_cic():
    c = cfg_fields()
    o = Oracle(c)
    rng = np.random.default_rng(1)
    fl = {n: rng.standard_normal(o.shape) for n in ORDER}
    for n in ORDER:
        o.upload_field(n, fl[n])
    for _ in range(40):
        p = rng.uniform(2.6, 9.4, 3) * np.array([1, 1, 1.8])
        got = o.gather(*p)
        for i, n in enumerate(ORDER):
            ref = hat_interp(fl[n], OFFS[n], *p)
            assert abs(got[i] - ref) < 1e-12 * (1 + abs(ref)), (n, p)


def test_gather_constant_and_linear_exact():
    c = cfg_fields()
    o = Oracle(c)
    K, L, M = np.meshgrid(*[np.arange(s, dtype=float) for s in o.shape], indexing="ij")
    for i, n in enumerate(ORDER):
        off = OFFS[n]
        o.upload_field(n, 0.7 + 0.3 * (K + off[0]) - 0.2 * (L + off[1]) + 0.1 * (M + off[2]))
    rng = np.random.default_rng(2)
    for _ in range(50):
        x, y, z = rng.uniform(2.5, 9.5), rng.uniform(2.5, 9.5), rng.uniform(2.5, 17.5)
        got = o.gather(x, y, z)
        assert np.allclose(got, 0.7 + 0.3 * x - 0.2 * y + 0.1 * z, atol=1e-13, rtol=0)


# ---------------------------------------------------------------- push
def test_boris_pure_rotation_closed_form():
    c = ps.preset(0)
    m = c.mass[0]
    B = 0.3
    v0 = np.array([0.02, -0.01, 0.005])
    p = np.array([5.0, 5.0, 5.0, *v0])
    out = boris(c, m, np.array([0, 0, 0, 0, 0, B]), p)
    gam = 1 / math.sqrt(1 - v0 @ v0)
    T = 0.5 * c.q * c.dt / m
    th = 2 * math.atan(T * B / gam)  # b = B / gamma(u^{n-1/2}) (:2830-2836)
    R = np.array([[math.cos(th), math.sin(th), 0], [-math.sin(th), math.cos(th), 0], [0, 0, 1]])
    v1 = R @ v0
    assert np.allclose(out[3:], v1, rtol=0, atol=1e-16)
    assert np.allclose(out[:3], p[:3] + v1 * c.dt * 0.5, rtol=0, atol=1e-15)


def test_boris_electric_energy_gain():
    c = ps.preset(0)
    m = -c.mass[1]  # electron: the push uses -m_e (:784)
    E = np.array([0.01, -0.02, 0.03])
    v0 = np.array([0.1, 0.0, -0.05])
    out = boris(c, m, np.array([*E, 0, 0, 0]), np.array([3, 3, 3, *v0]))
    T = 0.5 * c.q * c.dt / m
    u1 = v0 / math.sqrt(1 - v0 @ v0) + 2 * T * E
    v1 = u1 / math.sqrt(1 + u1 @ u1)
    assert np.allclose(out[3:], v1, rtol=1e-14, atol=0)


def test_boris_superluminal_input_resets_momentum():
    c = ps.preset(0)
    E = np.array([0.0, 0.05, 0.0])
    out = boris(c, c.mass[0], np.array([*E, 0, 0, 0]), np.array([3, 3, 3, 1.2, 0, 0]))
    T = 0.5 * c.q * c.dt / c.mass[0]
    u1 = 2 * T * E  # gamma1 is NaN -> 0, so u^{n-1/2} = 0 (:2822-2824)
    assert np.allclose(out[3:], u1 / math.sqrt(1 + u1 @ u1), rtol=1e-15, atol=0)


def test_exb_drift():
    c = ps.preset(0)
    m = c.mass[1]
    E = np.array([0.0, 0.001, 0.0])
    B = np.array([0.0, 0.0, 0.2])
    p = np.array([5, 5, 5, 0, 0, 0.0])
    vs = []
    for _ in range(4000):
        p = boris(c, m, np.concatenate([E, B]), p)
        vs.append(p[3:].copy())
    vd = np.mean(vs, axis=0)
    assert np.allclose(vd, np.cross(E, B) / (B @ B), atol=2e-5)


# ---------------------------------------------------------------- deposit
def test_single_particle_deposits_match_cic():
    c = cfg_fields()
    o = Oracle(c)
    rng = np.random.default_rng(3)
    for sp in (0, 1):
        for _ in range(6):
            x, y, z = rng.uniform(3, 9), rng.uniform(3, 9), rng.uniform(4, 16)
            v = rng.uniform(-0.2, 0.2, 3)
            # place the particle so that after the half-step drift it is at (x,y,z)
            dth = c.dt * 0.5
            o.upload_particles(sp, np.array([[x - v[0] * dth, y - v[1] * dth,
                                              z - v[2] * dth, *v]]))
            o.renew_grid()
            o.drift_deposit_n(sp)
            nname = "ni" if sp == 0 else "ne"
            n = o.download_field(nname)
            S = o.shape
            K, L, M = np.meshgrid(*[np.arange(s, dtype=float) for s in S], indexing="ij")
            w = hat(x - (K + .5)) * hat(y - (L + .5)) * hat(z - (M + .5))
            assert np.allclose(n, c.weight[sp] * w, atol=1e-15)
            assert abs(n.sum() - c.weight[sp]) < 1e-15
            # J at the pushed position: fields are zero and the particle is
            # outside the ballistic zone -> pure drift by v dt/2 through Boris
            o.upload_particles(sp, np.array([[x - v[0] * dth, y - v[1] * dth,
                                              z - v[2] * dth, *v]]))
            o.renew_grid()
            o.push_deposit_j(sp)
            p = o.download_particles(sp)[0]
            pre = "ji" if sp == 0 else "je"
            for a, comp in enumerate("xyz"):
                J = o.download_field(pre + comp)
                off = [0.5, 0.5, 0.5]
                off[a] = 0.0
                w = (hat(p[0] - (K + off[0])) * hat(p[1] - (L + off[1]))
                     * hat(p[2] - (M + off[2])))
                assert np.allclose(J, c.q * c.weight[sp] * p[3 + a] * w, atol=1e-16)


# ---------------------------------------------------------------- FDTD
def run_fields(o, steps):
    for _ in range(steps):
        o.renew_b_half()
        o.halo_b()
        o.renew_b_half()
        o.halo_b()
        o.renew_e()
        o.halo_e()


# (axis, E component, B component, sign of B so that E x B points along +axis)
WAVES = [(0, "ey", "bz", 1), (0, "ez", "by", -1), (1, "ez", "bx", 1),
         (1, "ex", "bz", -1), (2, "ex", "by", 1), (2, "ey", "bx", -1)]


@pytest.mark.parametrize("axis,ecomp,bcomp,sgn", WAVES)
def test_plane_wave_follows_24_dispersion(axis, ecomp, bcomp, sgn):
    n = [8, 8, 8]
    n[axis] = 32
    # z is not periodic and its guards stay frozen: disturbances enter from
    # the z edges at up to 4 cells per step, so z is made long and only a
    # core far from both edges is checked
    n[2] = 64 if axis == 2 else 48
    c = cfg_fields(*n)
    o = Oracle(c)
    N = n[axis]
    j = 3
    kap = 2 * math.pi * j / 32
    K = (27 * math.sin(kap / 2) - math.sin(3 * kap / 2)) / 12.0
    w = 2 / c.dt * math.asin(c.dt * K / 2)  # sin(w dt/2) = c dt K / 2
    idx = np.arange(o.shape[axis], dtype=float)
    shp = [1, 1, 1]
    shp[axis] = -1
    A = 1e-3
    E0 = np.broadcast_to((A * np.cos(kap * (idx + 0.5))).reshape(shp), o.shape)
    B0 = np.broadcast_to((sgn * A * np.cos(kap * idx + w * c.dt / 2)).reshape(shp), o.shape)
    o.upload_field(ecomp, E0)
    o.upload_field(bcomp, B0)
    steps = 3
    run_fields(o, steps)
    E = o.download_field(ecomp)
    B = o.download_field(bcomp)
    Eref = A * np.cos(kap * (idx + 0.5) - w * steps * c.dt)
    Bref = sgn * A * np.cos(kap * idx - w * (steps - 0.5) * c.dt)
    core = [slice(2, s - 2) for s in o.shape]
    zc = o.shape[2] // 2
    core[2] = slice(zc - 4, zc + 4) if axis != 2 else slice(24, 44)
    sl = [slice(None)] * 3
    Ec, Bc = E[tuple(core)], B[tuple(core)]
    ii = np.arange(o.shape[axis])[core[axis]]
    shp2 = [1, 1, 1]
    shp2[axis] = -1
    Er = np.broadcast_to(Eref[ii].reshape(shp2), Ec.shape)
    Br = np.broadcast_to(Bref[ii].reshape(shp2), Bc.shape)
    assert np.max(np.abs(Ec - Er)) < 1e-15 * 1e3 * A
    assert np.max(np.abs(Bc - Br)) < 1e-15 * 1e3 * A


def test_cfl_detector_at_dt_half():
    """dt = 0.5 exceeds the (2,4) 3D limit 6/(7 sqrt 3): noise grows, as the
    growth factor 4/3 of the Nyquist-corner mode predicts; dt = 0.45 does not."""
    growth = {}
    for dt in (0.45, 0.5):
        c = cfg_fields(16, 16, 16)
        c.dt = dt
        o = Oracle(c)
        rng = np.random.default_rng(5)
        e = rng.standard_normal(o.shape) * 1e-6
        o.upload_field("ex", e)
        run_fields(o, 1)
        a0 = np.abs(o.download_field("ex")).max()
        run_fields(o, 40)
        growth[dt] = np.abs(o.download_field("ex")).max() / a0
    assert growth[0.45] < 10
    assert growth[0.5] > 1e2


def test_sigma_damping_factors():
    c = cfg_fields(8, 8, 32)
    c.absorb[2] = 8
    c.sigma_amp = 0.05
    o = Oracle(c)
    one = np.ones(o.shape)
    o.upload_field("ex", one)
    o.renew_e()
    ex = o.download_field("ex")
    o2 = Oracle(c)
    o2.upload_field("bx", one)  # B0 = 0: B - B0 = 1, uniform
    o2.renew_b_half()
    bx = o2.download_field("bx")

    def depth(i2, N=32, A=8):  # reference init_grid (:332-366), width A -> 32
        if i2 - 4 <= 2 * A:
            d = (2 * A - (i2 - 4)) / 4.0
        elif i2 - 4 >= 2 * N - 2 * A:
            d = (i2 - 4 - (2 * N - 2 * A)) / 4.0
        else:
            d = 0.0
        return d * 32.0 / A

    def sig(m2):
        d = depth(m2)
        return c.sigma_amp * (0.1 * math.sqrt(d ** 4) + 0.01 * math.sqrt(d ** 12))

    for m in range(4, 32):  # away from the frozen z guards
        s = sig(2 * m + 1)  # Ex sits at (k, l+1/2, m+1/2)
        se = 0.5 * c.mu0 * s * c.c ** 2 * c.dt
        assert abs(ex[5, 5, m] - (1 - se) / (1 + se)) < 1e-15
        s = sig(2 * m)      # Bx sits at (k+1/2, l, m)
        sb = s * c.dt * 0.25 / c.eps0
        assert abs(bx[5, 5, m] - (1 - sb) / (1 + sb)) < 1e-15


# ---------------------------------------------------------------- halo / fold
def test_halo_is_periodic_image():
    c = cfg_fields()
    o = Oracle(c)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(o.shape)
    o.upload_field("ey", f)
    o.halo_e()
    g = o.download_field("ey")
    N = c.nx
    assert np.array_equal(g[2:N + 2, 2:N + 2], f[2:N + 2, 2:N + 2])
    assert np.array_equal(g[0:2, 2:N + 2], f[N:N + 2, 2:N + 2])
    assert np.array_equal(g[N + 2:N + 4, 2:N + 2], f[2:4, 2:N + 2])
    assert np.array_equal(g[:, 0:2], g[:, N:N + 2])
    assert np.array_equal(g[:, N + 2:N + 4], g[:, 2:4])


def test_fold_conserves_column_sums():
    c = cfg_fields()
    o = Oracle(c)
    rng = np.random.default_rng(8)
    for name in ("jix", "ne"):
        f = rng.standard_normal(o.shape)
        o.upload_field(name, f)
        o.fold_current()
        g = o.download_field(name)
        N = c.nx
        inner = g[2:N + 2, 2:N + 2].sum(axis=(0, 1))
        assert np.allclose(inner, f.sum(axis=(0, 1)), rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- sort
def test_sort_equals_stable_key_sort():
    c = ps.preset(0)
    o = Oracle(c)
    rng = np.random.default_rng(9)
    n = 5000
    p = np.column_stack([rng.uniform(0.5, c.nx + 3.5, n), rng.uniform(0.5, c.ny + 3.5, n),
                         rng.uniform(0.5, c.nz + 3.5, n), rng.standard_normal((n, 3))])
    p[::7, :3] = np.floor(p[::7, :3]) + 0.5  # many exact ties in the key
    o.upload_particles(1, p)
    o.sort(1)
    got = o.download_particles(1)
    k = np.trunc(p[:, 0]).astype(int) - 2
    l = np.trunc(p[:, 1]).astype(int) - 2
    m = np.trunc(p[:, 2]).astype(int) - 2
    ok = (k >= 0) & (k < c.nx) & (l >= 0) & (l < c.ny) & (m >= 0) & (m < c.nz)
    key = (k * c.ny + l) * c.nz + m
    order = np.flatnonzero(ok)[np.argsort(key[ok], kind="stable")]
    assert np.array_equal(got, p[order])


# ---------------------------------------------------------------- loading / inflow
def test_load_places_ppc_per_cell_in_order():
    c = ps.preset(0)
    o = Oracle(c)
    o.load_particles()
    for sp in (0, 1):
        p = o.download_particles(sp)
        assert len(p) == c.nx * c.ny * c.nz * c.ppc
        cell = (np.floor(p[:, 0]) - 2) * c.ny * c.nz + (np.floor(p[:, 1]) - 2) * c.nz + np.floor(p[:, 2]) - 2
        assert np.array_equal(cell, np.repeat(np.arange(c.nx * c.ny * c.nz), c.ppc))


def test_inflow_refills_slabs_with_drifting_maxwellian():
    c = ps.preset(0)
    c.ppc = 64
    o = Oracle(c)
    o.load_particles()
    o.inflow()
    for sp in (0, 1):
        p = o.download_particles(sp)
        z = p[:, 2]
        low = p[z < 4]
        high = p[z >= c.nz]
        assert len(low) == len(high) == c.nx * c.ny * 2 * c.ppc
        # untouched bulk is still ppc per cell
        mid = p[(z >= 4) & (z < c.nz)]
        assert len(mid) == c.nx * c.ny * (c.nz - 4) * c.ppc
        v = np.concatenate([low, high])[:, 3:]
        nv = len(v)
        mu = np.array(c.vdrift[:])
        s = c.vth[sp]
        assert np.all(np.abs(v.mean(0) - mu) < 5 * s / math.sqrt(nv))
        assert np.all(np.abs(v.std(0) / s - 1) < 5 / math.sqrt(2 * nv))


# ---------------------------------------------------------------- dipole B0
def test_b0_matches_point_dipole():
    c = ps.preset(1)
    o = Oracle(c)
    M, a = c.dipole_m, c.dipole_alpha
    mhat = np.array([0.0, math.sin(a), math.cos(a)])
    cx = (c.nx + 4) // 2
    rng = np.random.default_rng(11)
    comps = ["b0x", "b0y", "b0z"]
    f = {n: o.download_field(n) for n in comps}
    for _ in range(50):
        k, l, m = rng.integers(0, c.nx + 4, 3)
        for i, n in enumerate(comps):
            off = np.zeros(3)
            off[i] = 0.5
            r = np.array([k, l, m]) + off - cx
            rr = np.linalg.norm(r)
            if rr < c.dipole_soft + 1e-9:
                continue
            B = c.mu0 * M / (4 * math.pi) * (3 * (mhat @ r) * r - rr ** 2 * mhat) / rr ** 5
            assert abs(f[n][k, l, m] - B[i]) < 1e-12 * (1 + abs(B[i]))


def test_b0_softened_core_is_the_field_at_the_core_radius():
    # inside dipole_soft cells of the centre, B0 at a node is the point-dipole
    # field at radius dipole_soft on the ray through that node: constant along
    # each ray and continuous at the core's edge
    c = ps.preset(1)
    o = Oracle(c)
    M, a, s = c.dipole_m, c.dipole_alpha, c.dipole_soft
    mhat = np.array([0.0, math.sin(a), math.cos(a)])
    cx = (c.nx + 4) // 2
    comps = ["b0x", "b0y", "b0z"]
    f = {n: o.download_field(n) for n in comps}
    checked = 0
    for k in range(cx - 3, cx + 4):
        for l in range(cx - 3, cx + 4):
            for m in range(cx - 3, cx + 4):
                for i, n in enumerate(comps):
                    off = np.zeros(3)
                    off[i] = 0.5
                    r = np.array([k, l, m]) + off - cx
                    rr = np.linalg.norm(r)
                    if rr >= s - 1e-9:
                        continue
                    u = r * (s / rr)
                    B = c.mu0 * M / (4 * math.pi) * (3 * (mhat @ u) * u - s ** 2 * mhat) / s ** 5
                    assert abs(f[n][k, l, m] - B[i]) < 1e-12 * (1 + abs(B[i])), (k, l, m, n)
                    checked += 1
    assert checked >= 30


# ---------------------------------------------------------------- diagnostics
def diag_cfg():
    # constants away from 1, so that a dropped or misplaced factor shows
    c = cfg_fields(10, 10, 20)
    c.dx, c.dt, c.c = 0.5, 0.05, 2.0
    c.eps0, c.mu0, c.q = 2.5, 0.7, 1.3
    c.mass[0], c.mass[1] = 30.0, 0.9
    c.weight[0], c.weight[1] = 0.02, 0.03
    c.imf[0], c.imf[1], c.imf[2] = 0.1, -0.2, 0.3
    return c


def in_box(n):
    # interior indices i with 0.2 n < i < 0.8 n (strict, as thrust_f)
    return sum(1 for i in range(n) if 0.2 * n < i < 0.8 * n)


def test_diagnose_gauss_residual_exact_for_cubic_e():
    # the (2,4) divergence is exact for cubics: E_x = a x^3 on its nodes
    # gives d/dx at the cell's half node, 3 a ((k + 1/2) dx)^2
    c = diag_cfg()
    o = Oracle(c)
    dx = c.dx
    K, L, M = np.meshgrid(*[np.arange(s, dtype=float) for s in o.shape], indexing="ij")
    a, b, g = 0.3, -0.7, 0.11
    o.upload_field("ex", a * (K * dx) ** 3)
    o.upload_field("ey", b * (L * dx) ** 3)
    o.upload_field("ez", g * (M * dx) ** 3)
    d = o.diagnose()
    k = np.arange(2, c.nx + 2)[:, None, None]
    l = np.arange(2, c.ny + 2)[None, :, None]
    m = np.arange(2, c.nz + 2)[None, None, :]
    r = (3 * a * ((k + 0.5) * dx) ** 2 + 3 * b * ((l + 0.5) * dx) ** 2
         + 3 * g * ((m + 0.5) * dx) ** 2)
    assert d["gauss_max"] == pytest.approx(np.abs(r).max(), rel=1e-12)
    assert d["gauss_l2"] == pytest.approx(math.sqrt((r ** 2).sum()), rel=1e-12)
    assert d["charge"] == 0.0


def test_diagnose_gauss_residual_cancels_against_charge():
    # E_x = a x has div E = a; a uniform ion density eps0 a / q cancels it
    c = diag_cfg()
    o = Oracle(c)
    K = np.meshgrid(*[np.arange(s, dtype=float) for s in o.shape], indexing="ij")[0]
    a = 0.9
    o.upload_field("ex", a * K * c.dx)
    ncell = c.nx * c.ny * c.nz
    d = o.diagnose()
    assert d["gauss_max"] == pytest.approx(a, rel=1e-12)
    assert d["gauss_l2"] == pytest.approx(a * math.sqrt(ncell), rel=1e-12)
    o.upload_field("ni", np.full(o.shape, c.eps0 * a / c.q))
    d = o.diagnose()
    assert d["gauss_l2"] < 1e-12 * a * math.sqrt(ncell)
    assert d["charge"] == pytest.approx(c.eps0 * a * ncell * c.dx ** 3, rel=1e-12)


def test_diagnose_uniform_field_energies_and_box():
    c = diag_cfg()
    o = Oracle(c)
    imf = np.array(c.imf[:])
    for i, n in enumerate(["b0x", "b0y", "b0z"]):
        assert np.all(o.download_field(n) == imf[i])
    E = np.array([0.4, -1.1, 0.25])
    B = np.array([-0.3, 0.8, 1.7])
    for i in range(3):
        o.upload_field("ex ey ez".split()[i], np.full(o.shape, E[i]))
        o.upload_field("bx by bz".split()[i], np.full(o.shape, B[i]))
    d = o.diagnose()
    ncell, dv = c.nx * c.ny * c.nz, c.dx ** 3
    assert d["field_energy_e"] == pytest.approx(0.5 * c.eps0 * (E @ E) * ncell * dv, rel=1e-12)
    assert d["field_energy_b"] == pytest.approx(0.5 * (B @ B) / c.mu0 * ncell * dv, rel=1e-12)
    b1 = B - imf
    assert d["field_energy_b1"] == pytest.approx(0.5 * (b1 @ b1) / c.mu0 * ncell * dv, rel=1e-12)
    # the central box holds 5 x 5 x 11 cells of the 10 x 10 x 20 interior;
    # the reference sums its energy densities there without the cell volume
    nbox = in_box(c.nx) * in_box(c.ny) * in_box(c.nz)
    assert nbox == 5 * 5 * 11
    assert d["energy_box"][0] == pytest.approx(0.5 * c.eps0 * (E @ E) * nbox, rel=1e-12)
    assert d["energy_box"][1] == pytest.approx(0.5 * (B @ B) / c.mu0 * nbox, rel=1e-12)


def test_diagnose_particle_sums_in_uniform_fields():
    c = diag_cfg()
    o = Oracle(c)
    E = np.array([0.4, -1.1, 0.25])
    B = np.array([-0.3, 0.8, 1.7])
    for i in range(3):
        o.upload_field("ex ey ez".split()[i], np.full(o.shape, E[i]))
        o.upload_field("bx by bz".split()[i], np.full(o.shape, B[i]))
    rng = np.random.default_rng(5)
    dx, dv = c.dx, c.dx ** 3
    L = np.array([c.nx, c.ny, c.nz]) * dx
    parts = {}
    for sp in (0, 1):
        # half inside the 20-80 % box, half outside it along one axis (the
        # box bounds are 0.2 and 0.8 of the interior length on the array's
        # coordinates, whose interior starts at 2 dx, as thrust_f)
        pos = rng.uniform(0.35, 0.65, (12, 3)) * L
        for j in range(6):
            pos[j, j % 3] = (0.9 * L[j % 3] + 2 * dx) if j % 2 else 0.1 * L[j % 3]
        beta = rng.uniform(0.05, 0.8, 12)
        u = rng.standard_normal((12, 3))
        vel = (u / np.linalg.norm(u, axis=1)[:, None]) * (beta * c.c)[:, None]
        parts[sp] = np.hstack([pos, vel])
        o.upload_particles(sp, parts[sp])
    d = o.diagnose()
    F = np.zeros(15)
    for sp in (0, 1):
        p = parts[sp]
        msp, w = c.mass[sp], c.weight[sp]
        v = p[:, 3:]
        b2 = (v ** 2).sum(1) / c.c ** 2
        gam = 1 / np.sqrt(1 - b2)
        # gamma - 1 in the cancellation-free form beta^2 gamma^2 / (gamma + 1)
        assert d["kinetic"][sp] == pytest.approx(
            (msp * c.c ** 2 * w * dv * b2 * gam ** 2 / (gam + 1)).sum(), rel=1e-12)
        assert d["kinetic_nr"][sp] == pytest.approx(
            (0.5 * msp * w * dv * (v ** 2).sum(1)).sum(), rel=1e-12)
        assert np.allclose(d["momentum"][sp], (gam[:, None] * msp * w * dv * v).sum(0),
                           rtol=1e-12, atol=0)
        assert d["count"][sp] == len(p)
        inside = np.all((p[:, :3] > 0.2 * L) & (p[:, :3] < 0.8 * L), axis=1)
        assert inside.sum() == 6
        qn = (1 if sp == 0 else -1) * c.q * w * dv
        fe = qn * E * inside.sum()
        fb = qn * np.cross(v[inside], B).sum(0)
        b = 3 if sp == 0 else 9
        F[b:b + 3] += fe
        F[b + 3:b + 6] += fb
        F[0:3] += fe + fb
        assert d["energy_box"][2 + sp] == pytest.approx(
            (0.5 * msp * w * dv * (v[inside] ** 2).sum(1)).sum(), rel=1e-12)
    assert np.allclose(d["force"], F, rtol=1e-12, atol=1e-15)
