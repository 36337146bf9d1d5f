"""The FP32 form against the plain FP64 form on identical seeded inputs.

Both forms load the same Philox-seeded plasma (the FP32 form stores the same
draws as cell + fp32 offset and fp32 velocity), do the same pre-loop inflow
and run the same step loop (sort and inflow cadences of the config).  The
report gives

- after ``--first`` steps: the relative L2 difference of every field plane
  (E, the induced field B - B0, J and n of both species) and the relative
  difference of every diagnostic of ``pic_diagnose`` (field energies,
  kinetic energy and momentum per species, charge, Gauss residual, the
  reference's force and box-energy sums);
- every ``--every`` steps up to ``--steps``: the total energy of each form
  (electric + induced magnetic + kinetic of both species, each times dx^3),
  its drift from the first sample, and the relative difference between the
  two forms.

Usage: python -m plasmasim_amd.fp32_report [--preset 2] [--steps 1000]
           [--every 50] [--first 1] [--out FILE]
Prints one JSON object (and writes it to FILE if given).
"""
import argparse
import json
import sys
import time

import numpy as np

import plasmasim_amd as ps

PLANES = ["ex", "ey", "ez", "bx", "by", "bz", "jix", "jiy", "jiz", "jex", "jey",
          "jez", "ni", "ne"]


def _rel_l2(a, b):
    d = float(np.sqrt(np.sum((a - b) ** 2)))
    n = float(np.sqrt(np.sum(b ** 2)))
    return d / n if n > 0 else d


def _rel(a, b):
    return abs(a - b) / abs(b) if b != 0 else abs(a - b)


def total_energy(d):
    return (d["field_energy_e"] + d["field_energy_b1"] + d["kinetic"][0] +
            d["kinetic"][1])


def field_errors(s32, s64):
    b0 = [s64.download_field(n) for n in ("b0x", "b0y", "b0z")]
    out = {}
    for n in PLANES:
        a, b = s32.download_field(n), s64.download_field(n)
        if n in ("bx", "by", "bz"):
            i = "xyz".index(n[1])
            a, b = a - b0[i], b - b0[i]
            n = n + "-b0"
        out[n] = _rel_l2(a, b)
    return out


def diag_errors(d32, d64):
    out = {}
    for k, v in d64.items():
        a = np.atleast_1d(np.asarray(d32[k], dtype=float)).ravel()
        b = np.atleast_1d(np.asarray(v, dtype=float)).ravel()
        if a.size == 1:
            out[k] = _rel(float(a[0]), float(b[0]))
        else:
            scale = float(np.max(np.abs(b))) or 1.0
            out[k] = float(np.max(np.abs(a - b))) / scale
    return out


def run(preset=2, steps=1000, every=50, first=1):
    cfgs = {}
    for name, prec in (("fp64", ps.PIC_FP64), ("fp32", ps.PIC_FP32)):
        c = ps.preset(preset)
        c.precision = prec
        c.deposit = ps.PIC_DEPOSIT_ATOMIC
        cfgs[name] = c
    sims = {k: ps.Simulation(c) for k, c in cfgs.items()}
    for s in sims.values():
        s.load_particles()
        s.begin()
    t0 = time.time()
    for s in sims.values():
        s.step(first)
    s32, s64 = sims["fp32"], sims["fp64"]
    rep = {"preset": preset, "steps": steps, "every": every, "first": first,
           "particles": int(sum(s64.particle_count(sp) for sp in (0, 1)))}
    rep["fields_after_first"] = field_errors(s32, s64)
    d32, d64 = s32.diagnose(), s64.diagnose()
    rep["diagnostics_after_first"] = diag_errors(d32, d64)
    traj = []
    e0 = {}
    done = first
    while True:
        d = {k: s.diagnose() for k, s in sims.items()}
        e = {k: total_energy(v) for k, v in d.items()}
        if not e0:
            e0 = dict(e)
        traj.append({"step": done, "e64": e["fp64"], "e32": e["fp32"],
                     "drift64": (e["fp64"] - e0["fp64"]) / e0["fp64"],
                     "drift32": (e["fp32"] - e0["fp32"]) / e0["fp32"],
                     "rel_diff": _rel(e["fp32"], e["fp64"]),
                     "gauss64": d["fp64"]["gauss_l2"],
                     "gauss32": d["fp32"]["gauss_l2"]})
        if done >= steps:
            break
        n = min(every, steps - done)
        for s in sims.values():
            s.step(n)
        done += n
    rep["energy"] = traj
    rep["max_energy_rel_diff"] = max(t["rel_diff"] for t in traj)
    rep["final_drift64"] = traj[-1]["drift64"]
    rep["final_drift32"] = traj[-1]["drift32"]
    rep["final_diagnostics_rel"] = diag_errors(s32.diagnose(), s64.diagnose())
    rep["wall_s"] = time.time() - t0
    for s in sims.values():
        s.close()
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--first", type=int, default=1)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    rep = run(a.preset, a.steps, a.every, a.first)
    txt = json.dumps(rep)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
