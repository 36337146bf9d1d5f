"""MI355X-native 3D electromagnetic PIC time step (host-side Python mirror).

The product is the C-ABI library libplasmasim.so (include/plasmasim.h) over
hand-written HIP kernels for gfx950.  This module is a thin ctypes mirror of
that interface: one method per reference routine of fdtd()
(ashida-lab/PLASMAsim PIC/3d_pic458_domain.c :627-952), same names, same
order, errors raised as exceptions.  It performs no computation of its own.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (FIELDS, FIELD_ID, PIC_DEPOSIT_ATOMIC,  # noqa: F401
                   PIC_DEPOSIT_DETERMINISTIC, PIC_ELECTRON, PIC_FP32, PIC_FP64,
                   PIC_ION, PicConfig, PicDiag)

_LIB = None


def lib():
    """The loaded libplasmasim.so (raises ImportError when it is not built)."""
    global _LIB
    if _LIB is None:
        _LIB = _lib.load()
    return _LIB


class PicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pic error {code}: {msg}")
        self.code = code


def _timing_mode(on):
    return 2 if on == "passes" else (1 if on else 0)


def preset(n):
    """Workload preset n (0 tiny, 1 = 32^3, 2 = 128^3, 4 = 64^3, 5 = 256^3)."""
    c = PicConfig()
    rc = lib().pic_config_preset(int(n), ctypes.byref(c))
    if rc:
        raise PicError(rc, f"unknown preset {n}")
    return c


class Simulation:
    """One pic_ctx: device-resident fields and particles of one subdomain."""

    def __init__(self, cfg):
        self._L = lib()
        self.cfg = cfg.copy()
        ptr = ctypes.c_void_p()
        rc = self._L.pic_create(ctypes.byref(self.cfg), ctypes.byref(ptr))
        if rc:
            raise PicError(rc, "pic_create failed (see stderr)")
        self._ctx = ptr
        self.shape = (cfg.nx + 4, cfg.ny + 4, cfg.nz + 4)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._L.pic_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            msg = self._L.pic_last_error(self._ctx)
            raise PicError(rc, msg.decode() if msg else "")

    # -- data movement ----------------------------------------------------
    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        self._check(self._L.pic_upload_field(self._ctx, FIELD_ID[name], _lib.dptr(a)))

    def download_field(self, name):
        a = np.empty(self.shape, dtype=np.float64)
        self._check(self._L.pic_download_field(self._ctx, FIELD_ID[name], _lib.dptr(a)))
        return a

    def upload_particles(self, sp, parts):
        """parts: (n, 6) array of x, y, z, vx, vy, vz."""
        p = np.asarray(parts, dtype=np.float64)
        cols = [np.ascontiguousarray(p[:, j]) for j in range(6)]
        self._check(self._L.pic_upload_particles(
            self._ctx, sp, len(p), *[_lib.dptr(c) for c in cols]))

    def download_particles(self, sp):
        n = ctypes.c_int64()
        self._check(self._L.pic_particle_count(self._ctx, sp, ctypes.byref(n)))
        cap = max(n.value, 1)
        cols = [np.empty(cap) for _ in range(6)]
        self._check(self._L.pic_download_particles(
            self._ctx, sp, cap, ctypes.byref(n), *[_lib.dptr(c) for c in cols]))
        return np.stack([c[:n.value] for c in cols], axis=1)

    def download_raw32(self, sp):
        n = ctypes.c_int64()
        self._check(self._L.pic_particle_count(self._ctx, sp, ctypes.byref(n)))
        cap = max(n.value, 1)
        cell = np.empty(cap, np.int32)
        f = [np.empty(cap, np.float32) for _ in range(6)]
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.pic_download_particles_raw32(
            self._ctx, sp, cap, ctypes.byref(n),
            cell.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            *[a.ctypes.data_as(fp) for a in f]))
        k = n.value
        return cell[:k], np.stack([a[:k] for a in f], axis=1)

    def particle_count(self, sp):
        n = ctypes.c_int64()
        self._check(self._L.pic_particle_count(self._ctx, sp, ctypes.byref(n)))
        return n.value

    # -- step routines, in fdtd() order -------------------------------------
    def load_particles(self):
        self._check(self._L.pic_load_particles(self._ctx))

    def renew_grid(self):
        self._check(self._L.pic_renew_grid(self._ctx))

    def renew_b_half(self):
        self._check(self._L.pic_renew_b_half(self._ctx))

    def halo_b(self):
        self._check(self._L.pic_halo_b(self._ctx))

    def halo_e(self):
        self._check(self._L.pic_halo_e(self._ctx))

    def push_deposit_j(self, sp):
        self._check(self._L.pic_push_deposit_j(self._ctx, sp))

    def drift_deposit_n(self, sp):
        self._check(self._L.pic_drift_deposit_n(self._ctx, sp))

    def fold_current(self):
        self._check(self._L.pic_fold_current(self._ctx))

    def renew_e(self):
        self._check(self._L.pic_renew_e(self._ctx))

    def particle_boundary(self):
        self._check(self._L.pic_particle_boundary(self._ctx))

    def inflow(self):
        self._check(self._L.pic_inflow(self._ctx))

    def sort(self, sp):
        self._check(self._L.pic_sort(self._ctx, sp))

    def poisson(self):
        self._check(self._L.pic_poisson(self._ctx))

    def begin(self):
        self._check(self._L.pic_begin(self._ctx))

    def step(self, n=1):
        self._check(self._L.pic_step(self._ctx, int(n)))

    def diagnose(self):
        d = PicDiag()
        self._check(self._L.pic_diagnose(self._ctx, ctypes.byref(d)))
        return d.to_dict()

    def sort_stats(self):
        """(sorts by the tile sort, sorts by the radix sort) so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.pic_sort_stats(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def thrust_ave(self, on=True):
        """pic_thrust_ave_enable: sum thrust_fave's force over pic_step's steps."""
        self._check(self._L.pic_thrust_ave_enable(self._ctx, 1 if on else 0))

    def thrust_ave_get(self, reset=False):
        """(F[27] summed over the steps, number of steps)."""
        F = np.zeros(27)
        n = ctypes.c_int64()
        self._check(self._L.pic_thrust_ave_get(self._ctx, _lib.dptr(F), ctypes.byref(n),
                                               1 if reset else 0))
        return F, n.value

    def diagnose_device(self, dev_ptr):
        """pic_diagnose_device: the diagnostics record into device memory at
        dev_ptr (an integer address of pic_sizeof_diag() bytes), queued on
        the context stream without synchronising."""
        self._check(self._L.pic_diagnose_device(self._ctx, ctypes.c_void_p(int(dev_ptr))))

    def sync(self):
        self._check(self._L.pic_sync(self._ctx))

    def stream(self):
        return self._L.pic_stream(self._ctx)

    # -- adjoint ------------------------------------------------------------
    def adj_forward(self, params, nsteps, ckpt_every, radius, nquad):
        """Loss for p = (Mx, My, Mz, V); keeps checkpoints for adj_backward."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        loss = ctypes.c_double()
        self._check(self._L.pic_adj_forward(self._ctx, _lib.dptr(p), int(nsteps),
                                            int(ckpt_every), float(radius), int(nquad),
                                            ctypes.byref(loss)))
        return loss.value

    def adj_backward(self, dloss=1.0):
        g = np.zeros(4)
        self._check(self._L.pic_adj_backward(self._ctx, float(dloss), _lib.dptr(g)))
        return g

    def counters(self):
        """(steps done, inflow cadence counter, sort cadence counter)."""
        st, ca, cs = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.pic_get_counters(self._ctx, ctypes.byref(st),
                                             ctypes.byref(ca), ctypes.byref(cs)))
        return st.value, ca.value, cs.value

    def set_counters(self, step, c_add, c_sort):
        """Set the step and cadence counters (as a restored run would)."""
        self._check(self._L.pic_set_counters(self._ctx, int(step), int(c_add),
                                             int(c_sort)))

    def output_fields(self, path, tecplot=False, step=0):
        """Cell-centred text dump (output / output_tecplot of the reference)."""
        self._check(self._L.pic_output_fields(self._ctx, os.fsencode(path),
                                              1 if tecplot else 0, int(step)))

    # -- checkpoint and resume ---------------------------------------------
    def checkpoint_save(self, path):
        """Write the complete state between steps (pic_checkpoint_save)."""
        self._check(self._L.pic_checkpoint_save(self._ctx, os.fsencode(path)))

    def checkpoint_load(self, path):
        """Resume from a checkpoint written with the same config."""
        self._check(self._L.pic_checkpoint_load(self._ctx, os.fsencode(path)))

    # -- timing -------------------------------------------------------------
    def timing(self, on=True):
        """HIP-event timing per class: True every class, "passes" the
        particle passes only (fewer packets in the stream), False off."""
        self._check(self._L.pic_timing_enable(self._ctx, _timing_mode(on)))

    def timing_reset(self):
        self._check(self._L.pic_timing_reset(self._ctx))

    def timing_get(self, which):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.pic_timing_get(self._ctx, _lib.TIMING[which],
                                           ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


# ---------------------------------------------------------------------------
# 2D3V variant (include/plasmasim2d.h; reference PIC2D/2d_pic278_mpi.c)

from ._lib import FIELDS2D, FIELD2D_ID, Pic2dConfig, Pic2dDiag  # noqa: E402


def preset2d(n):
    """2D workload preset n (0 = 32 x 64 parity case, 1 = 256^2, 3 = 2048^2)."""
    c = Pic2dConfig()
    rc = lib().pic2d_config_preset(int(n), ctypes.byref(c))
    if rc:
        raise PicError(rc, f"unknown 2D preset {n}")
    return c


class Simulation2D:
    """One pic2d_ctx: one method per routine of the 2D fdtd() (:527-1019)."""

    _ROUTINES = ("load_particles", "begin", "renew_grid", "inflow",
                 "renew_b_half", "halo", "fold_current", "fold_density",
                 "renew_e", "poisson", "particle_boundary", "sync")

    def __init__(self, cfg):
        self._L = lib()
        self.cfg = cfg.copy()
        ptr = ctypes.c_void_p()
        rc = self._L.pic2d_create(ctypes.byref(self.cfg), ctypes.byref(ptr))
        if rc:
            raise PicError(rc, "pic2d_create failed (see stderr)")
        self._ctx = ptr
        self.shape = (cfg.nx + 4, cfg.nz + 4)

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.pic2d_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            msg = self._L.pic2d_last_error(self._ctx)
            raise PicError(rc, msg.decode() if msg else "")

    def __getattr__(self, name):
        if name in Simulation2D._ROUTINES:
            fn = getattr(self._L, "pic2d_" + name)
            return lambda: self._check(fn(self._ctx))
        raise AttributeError(name)

    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        self._check(self._L.pic2d_upload_field(self._ctx, FIELD2D_ID[name], _lib.dptr(a)))

    def download_field(self, name):
        a = np.empty(self.shape, dtype=np.float64)
        self._check(self._L.pic2d_download_field(self._ctx, FIELD2D_ID[name], _lib.dptr(a)))
        return a

    def upload_particles(self, sp, parts):
        """parts: (n, 5) array of x, z, vx, vy, vz."""
        p = np.asarray(parts, dtype=np.float64)
        cols = [np.ascontiguousarray(p[:, j]) for j in range(5)]
        self._check(self._L.pic2d_upload_particles(
            self._ctx, sp, len(p), *[_lib.dptr(c) for c in cols]))

    def particle_count(self, sp):
        n = ctypes.c_int64()
        self._check(self._L.pic2d_particle_count(self._ctx, sp, ctypes.byref(n)))
        return n.value

    def download_particles(self, sp):
        cap = max(self.particle_count(sp), 1)
        n = ctypes.c_int64()
        cols = [np.empty(cap) for _ in range(5)]
        self._check(self._L.pic2d_download_particles(
            self._ctx, sp, cap, ctypes.byref(n), *[_lib.dptr(c) for c in cols]))
        return np.stack([c[:n.value] for c in cols], axis=1)

    def push_deposit_j(self, sp):
        self._check(self._L.pic2d_push_deposit_j(self._ctx, sp))

    def drift_deposit_n(self, sp):
        self._check(self._L.pic2d_drift_deposit_n(self._ctx, sp))

    def external_current(self, step):
        self._check(self._L.pic2d_external_current(self._ctx, int(step)))

    def sort(self, sp):
        self._check(self._L.pic2d_sort(self._ctx, sp))

    def sort_stats(self):
        """(sorts by the tile sort, sorts by the radix sort) so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.pic2d_sort_stats(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def step(self, n=1):
        self._check(self._L.pic2d_step(self._ctx, int(n)))

    def step_index(self):
        s = ctypes.c_int64()
        self._check(self._L.pic2d_step_index(self._ctx, ctypes.byref(s)))
        return s.value

    def diagnose(self):
        d = Pic2dDiag()
        self._check(self._L.pic2d_diagnose(self._ctx, ctypes.byref(d)))
        return d.to_dict()

    def probe(self, k, m):
        out = np.empty(8)
        self._check(self._L.pic2d_probe(self._ctx, int(k), int(m), _lib.dptr(out)))
        return out

    def timing(self, on=True):
        """HIP-event timing per class: True every class, "passes" the
        particle passes only (fewer packets in the stream), False off."""
        self._check(self._L.pic2d_timing_enable(self._ctx, _timing_mode(on)))

    def timing_reset(self):
        self._check(self._L.pic2d_timing_reset(self._ctx))

    def timing_get(self, which):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.pic2d_timing_get(self._ctx, _lib.TIMING[which],
                                             ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
