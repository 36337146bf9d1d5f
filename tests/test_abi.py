"""The C-ABI library loads and exports every function include/plasmasim.h
and include/plasmasim2d.h declare; struct layouts agree with the binding.
No GPU calls."""
import ctypes
import os
import re

import pytest

import plasmasim_amd
from plasmasim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(pic(?:2d)?_\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(declared_functions()) == bound


def test_struct_layout_matches():
    lib = plasmasim_amd.lib()
    assert lib.pic_sizeof_config() == ctypes.sizeof(_lib.PicConfig)
    assert lib.pic_sizeof_diag() == ctypes.sizeof(_lib.PicDiag)
    assert lib.pic2d_sizeof_config() == ctypes.sizeof(_lib.Pic2dConfig)
    assert lib.pic2d_sizeof_diag() == ctypes.sizeof(_lib.Pic2dDiag)


def test_2d_presets_fill_config():
    for n in (0, 1, 3):
        c = plasmasim_amd.preset2d(n)
        assert c.nx > 0 and c.ppc > 0 and c.dt < 6 / (7 * 2 ** 0.5)
        wpe = (c.ppc * c.weight[1] * c.q ** 2 / (c.eps0 * c.mass[1])) ** 0.5
        assert abs(wpe * c.dt - 0.1) < 1e-12
        # line currents straddle the centre, 5 cells apart along z
        assert c.ext_m[1] - c.ext_m[0] == 5
        assert 2 <= c.phi_lo[1] < c.phi_hi[1] <= c.nz + 2
    c3 = plasmasim_amd.preset2d(3)
    assert (c3.nx, c3.nz, c3.ppc) == (2048, 2048, 100)


def test_presets_fill_config():
    for n in (0, 1, 2, 4, 5):
        c = plasmasim_amd.preset(n)
        assert c.nx > 0 and c.ppc > 0 and c.dt < 6 / (7 * 3 ** 0.5)
        # omega_pe * dt = 0.1 with n0 = ppc * weight
        wpe = (c.ppc * c.weight[1] * c.q ** 2 / (c.eps0 * c.mass[1])) ** 0.5
        assert abs(wpe * c.dt - 0.1) < 1e-12
    c2 = plasmasim_amd.preset(2)
    assert (c2.nx, c2.ppc, c2.precision) == (128, 64, plasmasim_amd.PIC_FP32)


def test_standoff_formula():
    import math
    for n, L in ((1, 8.0), (2, 32.0)):
        c = plasmasim_amd.preset(n)
        n0 = c.ppc * c.weight[0]
        V = c.vdrift[2]
        got = (c.mu0 * c.dipole_m ** 2 /
               (8 * math.pi ** 2 * n0 * c.mass[0] * V * V)) ** (1 / 6)
        assert abs(got - L) < 1e-9


def test_autograd_bridge_checks_arguments_on_host():
    import torch
    from plasmasim_amd.autograd import thrust_loss
    with pytest.raises(ValueError):
        thrust_loss(None, torch.zeros(3, dtype=torch.float64), 4)
