"""FP32 form against the FP64 form over many steps (32^3 x 16 ppc, with
inflow, outflow and the dipole): total energy (E, induced B, kinetic) of the
two forms within 1 % of each other at every sample over 1000 steps, and after
the first step every field plane within relative L2 1e-5 of the FP64 form.
The same comparison at 128^3 x 64 ppc is `python -m plasmasim_amd.fp32_report`."""
import pytest

from plasmasim_amd import fp32_report

pytestmark = pytest.mark.gpu


def test_fp32_energy_tracks_fp64_over_1000_steps():
    rep = fp32_report.run(preset=1, steps=1000, every=100, first=1)
    print("fields after one step:", {k: f"{v:.2e}" for k, v in rep["fields_after_first"].items()})
    print("energy samples:", [(t["step"], f"{t['rel_diff']:.2e}", f"{t['drift32']:.3e}",
                               f"{t['drift64']:.3e}") for t in rep["energy"]])
    for k, v in rep["fields_after_first"].items():
        assert v <= 1e-5, (k, v)
    assert rep["max_energy_rel_diff"] <= 1e-2, rep["max_energy_rel_diff"]
    assert abs(rep["final_drift32"] - rep["final_drift64"]) <= 1e-2
