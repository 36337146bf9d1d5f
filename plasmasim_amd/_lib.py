"""ctypes binding of the C-ABI declared in include/plasmasim.h.

The library is loaded from this package directory (built in-tree by
plasmasim_amd/build.py).  There is no fallback: if libplasmasim.so is missing
or cannot be loaded, importing raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplasmasim.so")

# constants mirrored from include/plasmasim.h
PIC_OK = 0
PIC_E_ARG, PIC_E_NOMEM, PIC_E_HIP, PIC_E_CAPACITY, PIC_E_STATE = -1, -2, -3, -4, -5
PIC_FP64, PIC_FP32 = 0, 1
PIC_DEPOSIT_ATOMIC, PIC_DEPOSIT_DETERMINISTIC = 0, 1
PIC_ION, PIC_ELECTRON = 0, 1
FIELDS = ["ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z",
          "jix", "jiy", "jiz", "jex", "jey", "jez", "ni", "ne"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
FIELD_ID["phi"] = len(FIELDS)  # PIC_PHI: download only
TIMING = {"push": 0, "field_b": 1, "field_e": 2, "sort": 3, "inflow": 4,
          "other": 5, "push_sort": 6}


class PicConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ppc", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("c", ctypes.c_double), ("eps0", ctypes.c_double),
        ("mu0", ctypes.c_double), ("q", ctypes.c_double),
        ("mass", ctypes.c_double * 2), ("weight", ctypes.c_double * 2),
        ("vth", ctypes.c_double * 2), ("vdrift", ctypes.c_double * 3),
        ("imf", ctypes.c_double * 3),
        ("dipole_m", ctypes.c_double), ("dipole_alpha", ctypes.c_double),
        ("dipole_soft", ctypes.c_double), ("dipole_rmax", ctypes.c_double),
        ("absorb", ctypes.c_int32 * 3), ("sigma_amp", ctypes.c_double),
        ("ballistic_lo", ctypes.c_double * 3),
        ("ballistic_hi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int32 * 2), ("inflow_every", ctypes.c_int32),
        ("sort_every", ctypes.c_int32), ("refill_far", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("deposit", ctypes.c_int32),
        ("tile", ctypes.c_int32 * 3),
        ("seed", ctypes.c_uint64), ("capacity", ctypes.c_int64),
        ("poisson_every", ctypes.c_int32), ("phi_margin", ctypes.c_int32),
    ]

    def copy(self):
        c = PicConfig()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(self))
        return c


class PicDiag(ctypes.Structure):
    _fields_ = [
        ("field_energy_e", ctypes.c_double), ("field_energy_b", ctypes.c_double),
        ("field_energy_b1", ctypes.c_double),
        ("kinetic", ctypes.c_double * 2), ("kinetic_nr", ctypes.c_double * 2),
        ("momentum", (ctypes.c_double * 3) * 2),
        ("charge", ctypes.c_double), ("gauss_l2", ctypes.c_double),
        ("gauss_max", ctypes.c_double),
        ("force", ctypes.c_double * 15), ("energy_box", ctypes.c_double * 6),
        ("count", ctypes.c_int64 * 2),
    ]

    def to_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, ctypes.Array):
                v = np.array([list(x) if isinstance(x, ctypes.Array) else x
                              for x in v])
            d[name] = v
        return d


class Pic2dConfig(ctypes.Structure):
    """pic2d_config of include/plasmasim2d.h."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("nz", ctypes.c_int32), ("ppc", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("c", ctypes.c_double), ("eps0", ctypes.c_double),
        ("mu0", ctypes.c_double), ("q", ctypes.c_double),
        ("mass", ctypes.c_double * 2), ("weight", ctypes.c_double * 2),
        ("vth", ctypes.c_double * 2), ("vdrift", ctypes.c_double),
        ("imf", ctypes.c_double * 3),
        ("ext_current", ctypes.c_double), ("ext_ramp", ctypes.c_int32),
        ("ext_k", ctypes.c_int32), ("ext_m", ctypes.c_int32 * 2),
        ("absorb", ctypes.c_int32 * 2), ("sigma_amp", ctypes.c_double),
        ("ballistic_lo", ctypes.c_double * 2),
        ("ballistic_hi", ctypes.c_double * 2),
        ("phi_lo", ctypes.c_int32 * 2), ("phi_hi", ctypes.c_int32 * 2),
        ("poisson_every", ctypes.c_int32), ("inflow_every", ctypes.c_int32),
        ("sort_every", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("deposit", ctypes.c_int32), ("tile", ctypes.c_int32 * 2),
        ("seed", ctypes.c_uint64), ("capacity", ctypes.c_int64),
    ]

    def copy(self):
        c = Pic2dConfig()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(self))
        return c


class Pic2dDiag(ctypes.Structure):
    """pic2d_diag of include/plasmasim2d.h."""
    _fields_ = [
        ("field_energy_e", ctypes.c_double), ("field_energy_b", ctypes.c_double),
        ("field_energy_b1", ctypes.c_double),
        ("kinetic", ctypes.c_double * 2),
        ("momentum", (ctypes.c_double * 3) * 2),
        ("charge", ctypes.c_double), ("gauss_l2", ctypes.c_double),
        ("gauss_max", ctypes.c_double), ("divb_l2", ctypes.c_double),
        ("divb_max", ctypes.c_double), ("count", ctypes.c_int64 * 2),
    ]

    to_dict = PicDiag.to_dict


FIELDS2D = FIELDS + ["phi"]
FIELD2D_ID = {name: i for i, name in enumerate(FIELDS2D)}

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64
_I64P = ctypes.POINTER(ctypes.c_int64)

# (name, restype, argtypes) of every symbol the header declares
SIGNATURES = [
    ("pic_sizeof_config", ctypes.c_size_t, []),
    ("pic_sizeof_diag", ctypes.c_size_t, []),
    ("pic_config_preset", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PicConfig)]),
    ("pic_create", ctypes.c_int, [ctypes.POINTER(PicConfig), ctypes.POINTER(_P)]),
    ("pic_destroy", None, [_P]),
    ("pic_last_error", ctypes.c_char_p, [_P]),
    ("pic_get_config", ctypes.c_int, [_P, ctypes.POINTER(PicConfig)]),
    ("pic_stream", _P, [_P]),
    ("pic_sync", ctypes.c_int, [_P]),
    ("pic_upload_field", ctypes.c_int, [_P, ctypes.c_int, _D]),
    ("pic_download_field", ctypes.c_int, [_P, ctypes.c_int, _D]),
    ("pic_upload_particles", ctypes.c_int, [_P, ctypes.c_int, _I64] + [_D] * 6),
    ("pic_download_particles", ctypes.c_int,
     [_P, ctypes.c_int, _I64, _I64P] + [_D] * 6),
    ("pic_particle_count", ctypes.c_int, [_P, ctypes.c_int, _I64P]),
    ("pic_download_particles_raw32", ctypes.c_int,
     [_P, ctypes.c_int, _I64, _I64P, ctypes.POINTER(ctypes.c_int32)]
     + [ctypes.POINTER(ctypes.c_float)] * 6),
    ("pic_load_particles", ctypes.c_int, [_P]),
    ("pic_renew_grid", ctypes.c_int, [_P]),
    ("pic_renew_b_half", ctypes.c_int, [_P]),
    ("pic_halo_b", ctypes.c_int, [_P]),
    ("pic_halo_e", ctypes.c_int, [_P]),
    ("pic_push_deposit_j", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_drift_deposit_n", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_fold_current", ctypes.c_int, [_P]),
    ("pic_renew_e", ctypes.c_int, [_P]),
    ("pic_particle_boundary", ctypes.c_int, [_P]),
    ("pic_inflow", ctypes.c_int, [_P]),
    ("pic_sort", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_diagnose", ctypes.c_int, [_P, ctypes.POINTER(PicDiag)]),
    ("pic_diagnose_device", ctypes.c_int, [_P, _P]),
    ("pic_step", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_thrust_ave_enable", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_thrust_ave_get", ctypes.c_int, [_P, _D, _I64P, ctypes.c_int]),
    ("pic_begin", ctypes.c_int, [_P]),
    ("pic_sort_stats", ctypes.c_int, [_P, _I64P, _I64P]),
    ("pic_get_counters", ctypes.c_int,
     [_P, _I64P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    ("pic_set_counters", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
    ("pic_timing_enable", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic_timing_get", ctypes.c_int, [_P, ctypes.c_int, _D, _I64P]),
    ("pic_timing_reset", ctypes.c_int, [_P]),
    ("pic_adj_forward", ctypes.c_int, [_P, _D, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, _D]),
    ("pic_adj_backward", ctypes.c_int, [_P, ctypes.c_double, _D]),
    ("pic_output_fields", ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_int64]),
    ("pic_poisson", ctypes.c_int, [_P]),
    ("pic_checkpoint_save", ctypes.c_int, [_P, ctypes.c_char_p]),
    ("pic_checkpoint_load", ctypes.c_int, [_P, ctypes.c_char_p]),
    # include/plasmasim2d.h
    ("pic2d_sizeof_config", ctypes.c_size_t, []),
    ("pic2d_sizeof_diag", ctypes.c_size_t, []),
    ("pic2d_config_preset", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Pic2dConfig)]),
    ("pic2d_create", ctypes.c_int, [ctypes.POINTER(Pic2dConfig), ctypes.POINTER(_P)]),
    ("pic2d_destroy", None, [_P]),
    ("pic2d_last_error", ctypes.c_char_p, [_P]),
    ("pic2d_get_config", ctypes.c_int, [_P, ctypes.POINTER(Pic2dConfig)]),
    ("pic2d_stream", _P, [_P]),
    ("pic2d_sync", ctypes.c_int, [_P]),
    ("pic2d_upload_field", ctypes.c_int, [_P, ctypes.c_int, _D]),
    ("pic2d_download_field", ctypes.c_int, [_P, ctypes.c_int, _D]),
    ("pic2d_upload_particles", ctypes.c_int, [_P, ctypes.c_int, _I64] + [_D] * 5),
    ("pic2d_download_particles", ctypes.c_int,
     [_P, ctypes.c_int, _I64, _I64P] + [_D] * 5),
    ("pic2d_particle_count", ctypes.c_int, [_P, ctypes.c_int, _I64P]),
    ("pic2d_load_particles", ctypes.c_int, [_P]),
    ("pic2d_begin", ctypes.c_int, [_P]),
    ("pic2d_renew_grid", ctypes.c_int, [_P]),
    ("pic2d_inflow", ctypes.c_int, [_P]),
    ("pic2d_renew_b_half", ctypes.c_int, [_P]),
    ("pic2d_halo", ctypes.c_int, [_P]),
    ("pic2d_push_deposit_j", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic2d_fold_current", ctypes.c_int, [_P]),
    ("pic2d_external_current", ctypes.c_int, [_P, _I64]),
    ("pic2d_drift_deposit_n", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic2d_fold_density", ctypes.c_int, [_P]),
    ("pic2d_renew_e", ctypes.c_int, [_P]),
    ("pic2d_poisson", ctypes.c_int, [_P]),
    ("pic2d_particle_boundary", ctypes.c_int, [_P]),
    ("pic2d_sort", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic2d_sort_stats", ctypes.c_int, [_P, _I64P, _I64P]),
    ("pic2d_step", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic2d_step_index", ctypes.c_int, [_P, _I64P]),
    ("pic2d_diagnose", ctypes.c_int, [_P, ctypes.POINTER(Pic2dDiag)]),
    ("pic2d_probe", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    ("pic2d_timing_enable", ctypes.c_int, [_P, ctypes.c_int]),
    ("pic2d_timing_get", ctypes.c_int, [_P, ctypes.c_int, _D, _I64P]),
    ("pic2d_timing_reset", ctypes.c_int, [_P]),
]


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m plasmasim_amd.build`")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pic_sizeof_config() != ctypes.sizeof(PicConfig):
        raise ImportError("pic_config layout mismatch between header and binding")
    if lib.pic_sizeof_diag() != ctypes.sizeof(PicDiag):
        raise ImportError("pic_diag layout mismatch between header and binding")
    if lib.pic2d_sizeof_config() != ctypes.sizeof(Pic2dConfig):
        raise ImportError("pic2d_config layout mismatch between header and binding")
    if lib.pic2d_sizeof_diag() != ctypes.sizeof(Pic2dDiag):
        raise ImportError("pic2d_diag layout mismatch between header and binding")
    return lib


def dptr(a):
    return a.ctypes.data_as(_D)
