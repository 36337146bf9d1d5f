"""Throughput of the EM-PIC step on one MI355X.

Default workload (3d): 128^3 cells x 64 particles/cell/species (268,435,456
particles), two species, dipole standoff 32 cells, inflow every 5 steps,
cell sort every 20 steps, synthetic seeded plasma.  A step is one full
fdtd() step (B half steps, fused gather/push/deposit pass per species,
current fold, E step, particle boundary, the amortised inflow and sort
when due, and thrust_fave's per-step FP64 force sum over the central box),
with all state resident in HBM.  The line also carries the field error of
the FP32 form: the relative L2 of J, n, E and B - B0 against the FP64 form
after one step from the same load.

--workload config5: 256^3 cells x 32 ppc x 2 species (1,073,741,824
particles, m_i/m_e = 400, drift 0.05, dipole tilted 30 degrees), FP32 form
in the deterministic sorted-tile deposit mode and in the atomic mode:
particle-steps/s and the particle pass's roofline fraction of each, and
whether two deterministic runs of 21 steps (a sort and inflows included)
end in bit-identical fields and diagnostics.

--workload adjoint: gradient of the step-averaged Maxwell-stress flux
through a sphere of 10 cells around the dipole with respect to the dipole
moment vector and the inflow speed, 64^3 cells x 32 ppc/species, 200 steps,
checkpoint every 20, FP32 forward with FP64 adjoint accumulation; checked
against the FP64-forward adjoint and central differences.

--workload 2d3v: the 2D3V variant, 2048 x 2048 cells x 100 particles/cell/
species (838,860,800 particles), line-current dipole, DST Poisson
correction every step, inflow every 250 steps, tile sort every 20.

--workload checkpoint: pic_checkpoint_save and pic_checkpoint_load of the
default workload's state after a few steps (FP32 and FP64 forms), in bytes
per second of checkpoint file, to a file under $TMPDIR (or /tmp), and the
time of the output() field dump of the FP32 state.

Prints one JSON line: value = particle-steps/s of the chosen form (FP32 by
default), with the plain FP64 form's rate, the roofline of the dominant
kernel (the fused particle pass) and a CPU baseline (the C oracle on the
host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# algorithmic bytes per particle-step of the fused pass (read + write state):
# 3D 6 doubles / int32 cell + 6 floats, 2D 5 doubles / int32 cell + 5 floats
BYTES_PER_PARTICLE = {"fp64": 96, "fp32": 56}
BYTES_PER_PARTICLE_2D = {"fp64": 80, "fp32": 48}
# HBM bytes per particle-step of the same pass measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE at each workload, scaled by the counts of a
# 1 GiB copy kernel (profiles/particle_pass_3d_128cubed_64ppc.md,
# profiles/particle_pass_2d3v_2048sq_100ppc.md)
TRAFFIC_PER_PARTICLE = {"fp64": 101.7, "fp32": 58.2}
TRAFFIC_PER_PARTICLE_2D = {"fp64": 81.4, "fp32": 49.0}


def run_form(ps, form, steps, warmup, workload="3d", poisson_every=0, tile=None,
             thrust=True, preset=2, deposit=None):
    if workload == "2d3v":
        c = ps.preset2d(3)
        c.precision = ps.PIC_FP64 if form == "fp64" else ps.PIC_FP32
        sim = ps.Simulation2D(c)
    else:
        c = ps.preset(preset)
        c.precision = ps.PIC_FP64 if form == "fp64" else ps.PIC_FP32
        c.poisson_every = poisson_every
        if deposit is not None:
            c.deposit = deposit
        if tile:
            c.tile[0], c.tile[1], c.tile[2] = tile
        sim = ps.Simulation(c)
    sim.load_particles()
    sim.begin()
    if workload != "2d3v" and thrust:
        # thrust_fave every step (:926), as the reference's step loop
        sim.thrust_ave(True)
    # live timing of the particle passes over every step of the run (HIP
    # events on the library's stream around the passes only: an event record
    # around every launch would add ~0.1 ms per step), so the pass average
    # covers the same launches a kernel trace of this command does
    sim.timing("passes")
    sim.timing_reset()
    sim.step(warmup)
    sim.sync()
    n0 = sim.particle_count(0) + sim.particle_count(1)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    t1 = time.perf_counter()
    n1 = sim.particle_count(0) + sim.particle_count(1)
    particles = 0.5 * (n0 + n1)
    rate = particles * steps / (t1 - t0)
    push_ms, push_n = sim.timing_get("push")
    # where the step's time goes: every class timed over 20 further steps
    # (one sort cadence), after the timed region
    tsteps = 20
    sim.timing(True)
    sim.timing_reset()
    sim.step(tsteps)
    sim.sync()
    tot = {k: sim.timing_get(k)
           for k in ("push", "push_sort", "field_b", "field_e", "sort", "inflow", "other")}
    sim.timing(False)
    thrust_sum = None
    if workload != "2d3v" and thrust:
        F, nsteps = sim.thrust_ave_get()
        thrust_sum = {"steps": nsteps, "F0_2": F[:3].tolist()}
    # which sort ran: tile sorts and radix sorts over the context's life (the
    # load's sort of the unsorted initial state is a radix sort)
    st = sim.sort_stats()
    sim.close()
    per_launch_particles = particles / 2.0  # one launch per species
    push_avg_s = push_ms / push_n / 1e3
    bpp = (BYTES_PER_PARTICLE_2D if workload == "2d3v" else BYTES_PER_PARTICLE)[form]
    achieved = bpp * per_launch_particles / push_avg_s
    return {
        "rate": rate, "ms_per_step": (t1 - t0) / steps * 1e3,
        "particles": particles, "push_avg_ms": push_avg_s * 1e3,
        "achieved": achieved, "kernel_ms": {k: v[0] / tsteps for k, v in tot.items()},
        "traffic": (TRAFFIC_PER_PARTICLE_2D if workload == "2d3v"
                    else TRAFFIC_PER_PARTICLE)[form] * per_launch_particles,
        "bytes_per_particle": bpp,
        "thrust": thrust_sum,
        "sorts": {"tile": st[0], "radix": st[1]},
    }


def field_error(ps, preset=2, tile=None):
    """Relative L2 of J, n, E and the induced field B - B0 of the FP32 form
    against the FP64 form after one step from the same seeded load."""
    import numpy as np

    def rel_l2(x, y):
        ny = np.linalg.norm(y)
        return float(np.linalg.norm(x - y) / ny) if ny else float(np.linalg.norm(x - y))

    sims = {}
    for form, prec in (("fp32", ps.PIC_FP32), ("fp64", ps.PIC_FP64)):
        c = ps.preset(preset)
        c.precision = prec
        if tile:
            c.tile[0], c.tile[1], c.tile[2] = tile
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(1)
        sims[form] = s
    a, b = sims["fp32"], sims["fp64"]
    out = {}
    for n in ("jix", "jiy", "jiz", "jex", "jey", "jez", "ni", "ne", "ex", "ey", "ez"):
        out[n] = rel_l2(a.download_field(n), b.download_field(n))
    for n in ("bx", "by", "bz"):
        b0 = b.download_field("b0" + n[1])
        out[n + "-b0"] = rel_l2(a.download_field(n) - b0, b.download_field(n) - b0)
    a.close()
    b.close()
    return {"after_steps": 1, "rel_l2": out, "max": max(out.values())}


def cpu_baseline(ps):
    """The C oracle (FP64, OpenMP) on the host cores: the same physics and
    particle density on a 128 x 128 x 32 slab of the workload (1/4 of the
    domain), 3 timed steps after 1 warm-up step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c = ps.preset(2)
    c.precision = ps.PIC_FP64
    c.nz = 32
    c.absorb[2] = 4
    c.ballistic_lo[2], c.ballistic_hi[2] = 8.0, 28.0
    o = Oracle(c, threads=cores)
    o.load_particles()
    o.begin()
    o.step(1)
    n = o.particle_count(0) + o.particle_count(1)
    steps = 3
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    o.close()
    return {"value": n * steps / (t1 - t0), "unit": "particle-steps/s",
            "cores": cores, "kind": "port",
            "sample": f"C oracle FP64, 128x128x32 slab x 64 ppc x 2 species "
                      f"({n} particles), {steps} steps"}


def calibrate_copy(mib):
    """One device-to-device copy of `mib` MiB: reads and writes exactly
    mib MiB, so the HBM byte counters of its dispatch give the factor that
    turns counter readings into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy calibration: {mib} MiB read, {mib} MiB written", file=sys.stderr)
    del src, dst
    torch.cuda.empty_cache()


def copy_bandwidth(nbytes=4 << 30, reps=5):
    """Bytes per second of a plain device-to-device copy (the HIP runtime's
    copy kernel, the runtime the library links), read + written, timed with
    HIP events: the achievable HBM rate beside the 8 TB/s vendor peak."""
    import ctypes
    h = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
    h.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
    h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipEventRecord.argtypes = [vp, vp]
    h.hipEventSynchronize.argtypes = [vp]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    h.hipEventDestroy.argtypes = [vp]
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    if h.hipMalloc(ctypes.byref(a), nbytes) or h.hipMalloc(ctypes.byref(b), nbytes):
        return None
    try:
        h.hipMemset(a, 1, nbytes)
        h.hipEventCreate(ctypes.byref(e0))
        h.hipEventCreate(ctypes.byref(e1))
        h.hipMemcpyAsync(b, a, nbytes, 3, None)  # warm-up
        h.hipEventRecord(e0, None)
        for _ in range(reps):
            h.hipMemcpyAsync(b, a, nbytes, 3, None)
        h.hipEventRecord(e1, None)
        h.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        h.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        h.hipEventDestroy(e0)
        h.hipEventDestroy(e1)
        return 2.0 * nbytes * reps / (ms.value * 1e-3)
    finally:
        h.hipFree(a)
        h.hipFree(b)


def cpu_baseline_2d(ps):
    """The 2D C oracle (FP64, OpenMP) on a 2048 x 64 strip of the 2D
    workload (1/32 of the domain, same density), 3 timed steps after 1."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle2D
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c = ps.preset2d(3)
    c.precision = ps.PIC_FP64
    c.nz = 64
    c.absorb[1] = 4
    c.ballistic_lo[1], c.ballistic_hi[1] = 4.0, 64.0
    c.phi_lo[1], c.phi_hi[1] = 8, 60
    c.ext_m[0], c.ext_m[1] = 31, 36
    o = Oracle2D(c, threads=cores)
    o.load_particles()
    o.begin()
    o.step(1)
    n = o.particle_count(0) + o.particle_count(1)
    steps = 3
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    o.close()
    return {"value": n * steps / (t1 - t0), "unit": "particle-steps/s",
            "cores": cores, "kind": "port",
            "sample": f"2D C oracle FP64, 2048x64 strip x 100 ppc x 2 species "
                      f"({n} particles), {steps} steps"}


def run_adjoint(ps, steps=200, ckpt=20, radius=10.0, nquad=1024,
                horizons=(20, 40, 80, 200)):
    """Forward + backward at the adjoint workload in both forms, and the
    FP64 gradient against central differences of the FP64 forward (h = 1e-4
    and 1e-7 of each parameter) over growing horizons."""
    import numpy as np

    def make(form):
        c = ps.preset(4)
        c.precision = ps.PIC_FP64 if form == "fp64" else ps.PIC_FP32
        n0 = c.nx * c.ny * c.nz * c.ppc
        add = c.nx * c.ny * 2 * c.ppc * (2 if c.refill_far else 1)
        c.capacity = n0 + (steps // c.inflow_every + 2) * add
        return c, n0

    c, n0 = make("fp64")
    M = c.dipole_m
    # the workload's moment M (0, sin a, cos a), as a vector
    p = np.array([0.0, M * np.sin(c.dipole_alpha), M * np.cos(c.dipole_alpha), c.vdrift[2]])
    scale = np.array([M, M, M, abs(c.vdrift[2])])
    out = {"params": p.tolist()}
    for form in ("fp32", "fp64"):
        cf, _ = make(form)
        sim = ps.Simulation(cf)
        sim.sync()
        t0 = time.perf_counter()
        loss = sim.adj_forward(p, steps, ckpt, radius, nquad)
        sim.sync()
        t1 = time.perf_counter()
        g = sim.adj_backward(1.0)
        sim.sync()
        t2 = time.perf_counter()
        out[form] = {"loss": loss, "grad": g.tolist(), "forward_s": t1 - t0,
                     "backward_s": t2 - t1, "particles": 2 * n0}
        sim.close()
    g32, g64 = np.array(out["fp32"]["grad"]), np.array(out["fp64"]["grad"])
    out["fp32_vs_fp64_grad_rel"] = (np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-300)).tolist()
    sim = ps.Simulation(c)
    c32, _ = make("fp32")
    sim32 = ps.Simulation(c32)
    table = []
    for N in horizons:
        loss = sim.adj_forward(p, N, ckpt, radius, nquad)
        g = sim.adj_backward(1.0)
        row = {"steps": N, "loss": loss, "adjoint": g.tolist()}
        # the FP32-forward gradient at the same horizon, against the FP64 one
        sim32.adj_forward(p, N, ckpt, radius, nquad)
        g32 = sim32.adj_backward(1.0)
        row["adjoint_fp32_forward"] = g32.tolist()
        row["fp32_vs_fp64_rel"] = [abs(a - b) / max(abs(b), 1e-300)
                                   for a, b in zip(g32.tolist(), g.tolist())]
        for h_rel in (1e-4, 1e-7):
            fd = []
            for i in range(4):
                h = h_rel * scale[i]
                pp, pm = p.copy(), p.copy()
                pp[i] += h
                pm[i] -= h
                fd.append((sim.adj_forward(pp, N, N, radius, nquad) -
                           sim.adj_forward(pm, N, N, radius, nquad)) / (2 * h))
            row[f"fd_h{h_rel:g}"] = fd
            row[f"rel_err_h{h_rel:g}"] = [abs(a - b) / max(abs(a), 1e-300)
                                          for a, b in zip(fd, g.tolist())]
        table.append(row)
        print(f"adjoint horizon {N}: rel err (h=1e-7) {row['rel_err_h1e-07']}, "
              f"(h=1e-4) {row['rel_err_h0.0001']}, FP32 forward vs FP64 "
              f"{row['fp32_vs_fp64_rel']}", file=sys.stderr, flush=True)
    sim.close()
    sim32.close()
    out["horizons"] = table
    return out


def run_config5(ps, steps, warmup):
    """256^3 x 32 ppc FP32 in the deterministic and the atomic deposit mode
    (rate and particle-pass roofline of each), and two deterministic runs of
    21 steps compared bit for bit (every field plane, the diagnostics and
    the counts)."""
    import numpy as np
    out = {}
    for mode, dep in (("deterministic", ps.PIC_DEPOSIT_DETERMINISTIC),
                      ("atomic", ps.PIC_DEPOSIT_ATOMIC)):
        r = run_form(ps, "fp32", steps, warmup, preset=5, deposit=dep)
        out[mode] = {"value": r["rate"], "ms_per_step": r["ms_per_step"],
                     "pass_avg_ms": r["push_avg_ms"],
                     "frac": r["achieved"] / HBM_PEAK,
                     "achieved_GBps": r["achieved"] / 1e9,
                     "kernel_ms_per_step": r["kernel_ms"], "particles": r["particles"]}
        print(f"config5 {mode}: {r['rate']:.3e} particle-steps/s, pass "
              f"{r['push_avg_ms']:.2f} ms", file=sys.stderr, flush=True)
    runs = []
    for _ in range(2):
        c = ps.preset(5)
        c.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(21)  # inflows at 5, 10, 15, 20 and the sort at step 20
        f = {n: s.download_field(n) for n in
             ("ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex", "jey",
              "jez", "ni", "ne")}
        d = s.diagnose()
        s.close()
        runs.append((f, d))
    (fa, da), (fb, db) = runs
    same_fields = all(np.array_equal(fa[n], fb[n]) for n in fa)
    same_diag = all(np.array_equal(np.asarray(da[k]), np.asarray(db[k])) for k in da)
    out["deterministic_bit_identical"] = {"steps": 21, "fields": bool(same_fields),
                                          "diagnostics_and_counts": bool(same_diag)}
    return out


def run_checkpoint(ps, steps=3):
    """Save and load times of the default workload's state in both forms,
    and the time of the output() field dump of the FP32 state."""
    import tempfile
    out = {}
    d = os.environ.get("TMPDIR", "/tmp")
    for form, prec in (("fp32", ps.PIC_FP32), ("fp64", ps.PIC_FP64)):
        c = ps.preset(2)
        c.precision = prec
        s = ps.Simulation(c)
        s.load_particles()
        s.begin()
        s.step(steps)
        s.sync()
        fd, path = tempfile.mkstemp(suffix=".ckpt", dir=d)
        os.close(fd)
        fd, dump = tempfile.mkstemp(suffix=".dat", dir=d)
        os.close(fd)
        try:
            t0 = time.perf_counter()
            s.checkpoint_save(path)
            t_save = time.perf_counter() - t0
            size = os.path.getsize(path)
            e0 = s.diagnose()
            s.close()
            r = ps.Simulation(c)
            t0 = time.perf_counter()
            r.checkpoint_load(path)
            r.sync()
            t_load = time.perf_counter() - t0
            e1 = r.diagnose()
            if form == "fp32":
                t0 = time.perf_counter()
                r.output_fields(dump)
                out["output_fields_s"] = time.perf_counter() - t0
                out["output_fields_bytes"] = os.path.getsize(dump)
            r.close()
        finally:
            os.remove(path)
            os.remove(dump)
        same = all(e0[k] == e1[k] for k in ("field_energy_e", "field_energy_b",
                                            "charge")) and \
            list(e0["count"]) == list(e1["count"])
        out[form] = {"bytes": size, "save_s": t_save, "load_s": t_load,
                     "save_GBps": size / t_save / 1e9, "load_GBps": size / t_load / 1e9,
                     "diagnostics_identical": bool(same)}
        print(f"{form}: {size / 1e9:.2f} GB, save {t_save:.2f} s, load {t_load:.2f} s",
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--workload", choices=["3d", "2d3v", "adjoint", "checkpoint", "config5"],
                    default="3d")
    ap.add_argument("--no-thrust", action="store_true",
                    help="3d: leave out thrust_fave's per-step force sum")
    ap.add_argument("--poisson-every", type=int, default=0, metavar="P",
                    help="3d: run the 3D Poisson correction every P steps (0: off, "
                         "as the reference)")
    ap.add_argument("--tile", type=int, nargs=3, metavar=("TX", "TY", "TZ"),
                    help="sort tile of the 3D workload in cells (default: the preset's, 8 4 4)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a plain "
                         "copy kernel of known size, to scale HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one subdomain per GPU)", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy > 0:
        calibrate_copy(a.calibrate_copy)
    import plasmasim_amd as ps
    ps.lib()
    copy_Bps = copy_bandwidth()
    if a.workload == "adjoint":
        r = run_adjoint(ps)
        f32 = r["fp32"]
        print(json.dumps({
            "metric": "adjoint particle-steps/s (forward + backward) on 1 MI355X",
            "value": f32["particles"] * 200 / (f32["forward_s"] + f32["backward_s"]),
            "unit": "particle-steps/s", "n_gpus": 1, "steps": 200, "warmup": 0,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 forward, f64 adjoint", "data": "synthetic seeded plasma",
            "config": {"workload": "adjoint: 64^3 x 32 ppc x 2 species, 200 steps, "
                                   "checkpoint every 20, stress flux through r = 10 cells, "
                                   "1024 quadrature points"},
            "adjoint": r}))
        return
    if a.workload == "config5":
        r = run_config5(ps, a.steps, a.warmup)
        det = r["deterministic"]
        print(json.dumps({
            "metric": "particle-steps/s on 1 MI355X, FP32 deterministic sorted-tile deposit",
            "value": det["value"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": det["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic seeded plasma (Philox4x32-10)",
            "config": {"workload": "3D EM-PIC step, 256^3 cells x 32 ppc x 2 species "
                                   "(1,073,741,824 particles), m_i/m_e 400, drift 0.05, "
                                   "dipole tilted 30 degrees, sort every 20, inflow every 5"},
            "roofline": {"bound": "hbm", "achieved": det["achieved_GBps"],
                         "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": det["frac"],
                         "traffic": None,
                         "kernel": "fused particle pass, deterministic mode",
                         "kernel_avg_ms": det["pass_avg_ms"], "bytes_per_particle": 56},
            "config5": r}))
        return
    if a.workload == "checkpoint":
        r = run_checkpoint(ps)
        print(json.dumps({
            "metric": "checkpoint save bytes/s on 1 MI355X (128^3 x 64 ppc state)",
            "value": r["fp32"]["save_GBps"] * 1e9, "unit": "B/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 state (fp64 alongside)",
            "data": "synthetic seeded plasma after 3 steps",
            "config": {"workload": "checkpoint of the 3D state, 128^3 cells x 64 ppc "
                                   "x 2 species, file under $TMPDIR"},
            "checkpoint": r}))
        return
    thrust = not a.no_thrust
    main_r = run_form(ps, a.form, a.steps, a.warmup, a.workload, a.poisson_every, a.tile,
                      thrust)
    other = None
    if a.form == "fp32" and not a.no_fp64:
        other = run_form(ps, "fp64", a.steps, a.warmup, a.workload, a.poisson_every, a.tile,
                         thrust)
    ferr = None
    if a.workload == "3d" and a.form == "fp32" and not a.no_fp64:
        ferr = field_error(ps, tile=a.tile)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(ps) if a.workload == "3d" else cpu_baseline_2d(ps)
    if a.workload == "2d3v":
        workload = ("2D3V EM-PIC step, 2048^2 cells x 100 ppc x 2 species "
                    "(838,860,800 particles), line-current dipole, dt 0.5, "
                    "Poisson every step, sort every 20, inflow every 250")
    else:
        workload = ("3D EM-PIC step, 128^3 cells x 64 ppc x 2 species "
                    "(268,435,456 particles), dipole standoff 32, dt 0.45, "
                    "sort every 20, inflow every 5")
        workload += (", thrust_fave force sum every step" if thrust
                     else ", without thrust_fave")
        if a.poisson_every > 0:
            workload += f", Poisson correction every {a.poisson_every}"
        if a.tile:
            workload += ", sort tiles %d x %d x %d" % tuple(a.tile)
    out = {
        "metric": "particle-steps/s on 1 MI355X, FP32 vs FP64 form at stated field error",
        "value": main_r["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.form == "fp32" else "f64",
        "data": "synthetic seeded plasma (Philox4x32-10), random-free fields at start",
        "config": {"workload": workload, "form": a.form},
        "roofline": {"bound": "hbm", "achieved": main_r["achieved"] / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": main_r["achieved"] / HBM_PEAK,
                     "traffic": main_r["traffic"],
                     "traffic_source": "profile constant: FETCH_SIZE + WRITE_SIZE bytes "
                                       "per particle from a rocprofv3 --pmc run of this "
                                       "workload, times the particles of one launch",
                     "kernel": "fused particle pass (gather, push, J and n deposit, boundary)",
                     "kernel_avg_ms": main_r["push_avg_ms"],
                     "bytes_per_particle": main_r["bytes_per_particle"],
                     "copy_measured": copy_Bps / 1e9 if copy_Bps else None,
                     "frac_of_copy": main_r["achieved"] / copy_Bps if copy_Bps else None},
        "cpu_baseline": cpu,
        "kernel_ms_per_step": main_r["kernel_ms"],
        "kernel_ms_per_step_over": "20 steps after the timed ones, every class timed",
        "sorts": main_r["sorts"],
    }
    if ferr:
        out["field_error_vs_fp64"] = ferr
    if main_r.get("thrust"):
        out["thrust_fave"] = main_r["thrust"]
    if other:
        out["fp64_value"] = other["rate"]
        out["fp64_ms_per_step"] = other["ms_per_step"]
        out["fp32_over_fp64"] = main_r["rate"] / other["rate"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
