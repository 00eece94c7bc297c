"""Warm-start surrogate: ctypes binding of the U-Net in libdsmc.so
(include/dsmc_unet.h) and the mapping between DSMC per-cell fields and the
network's normalised targets.

Inputs of the network: the solid mask of a 64^3 (in general edge^3) periodic
voxel grid and two scalars per geometry, a Knudsen feature (log10 Kn) and a
body-force feature (log10 Mach of the forcing).  Outputs per collision cell
(edge/4)^3: n_rel - 1, u / u_scale (3 components), T / T_gas - 1, where
n_rel is the cell's particle count over the uniform-seeding expectation
(particles per empty voxel times the cell's empty voxels) and u_scale the
forcing's target drift (Mach times the speed of sound).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from master_amd import config as C
from master_amd import dsmc

P = ctypes.c_void_p


def _lib():
    lib = dsmc.lib()
    if not getattr(lib, "_unet_ready", False):
        for name in ("dsmc_unet_create", "dsmc_unet_get_params", "dsmc_unet_set_params", "dsmc_unet_forward",
                     "dsmc_unet_loss_grad", "dsmc_unet_set_data", "dsmc_unet_train"):
            getattr(lib, name).restype = ctypes.c_int
        lib.dsmc_unet_num_params.restype = ctypes.c_int64
        lib.dsmc_unet_num_params.argtypes = [P]
        lib.dsmc_unet_last_error.restype = ctypes.c_char_p
        lib.dsmc_unet_last_error.argtypes = [P]
        lib.dsmc_unet_destroy.restype = None
        lib.dsmc_unet_destroy.argtypes = [P]
        lib._unet_ready = True
    return lib


def _ptr(a):
    return a.ctypes.data_as(P)


class Net:
    """The surrogate network on the current HIP device."""

    def __init__(self, edge=64, max_batch=4, seed=0):
        self.edge, self.max_batch = edge, max_batch
        h = P()
        st = _lib().dsmc_unet_create(ctypes.c_int(edge), ctypes.c_int(max_batch), ctypes.c_uint64(seed),
                                     ctypes.byref(h))
        if st != 0:
            raise dsmc.DsmcError(f"dsmc_unet_create failed ({st})")
        self.h = h
        self.nparams = int(_lib().dsmc_unet_num_params(h))
        self.cells = (edge // 4) ** 3

    def _call(self, name, *args):
        st = getattr(_lib(), name)(self.h, *args)
        if st != 0:
            msg = _lib().dsmc_unet_last_error(self.h)
            raise dsmc.DsmcError(f"{name}: status {st}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            _lib().dsmc_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self._call("dsmc_unet_get_params", _ptr(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self._call("dsmc_unet_set_params", _ptr(p))

    def _inputs(self, masks, scalars):
        m = np.ascontiguousarray(masks, dtype=np.uint8).reshape(-1, self.edge ** 3)
        s = np.ascontiguousarray(scalars, dtype=np.float32).reshape(-1, 2)
        assert len(m) == len(s) and 1 <= len(m) <= self.max_batch
        return m, s

    def forward(self, masks, scalars):
        m, s = self._inputs(masks, scalars)
        out = np.zeros((len(m), self.cells, 5), np.float32)
        self._call("dsmc_unet_forward", ctypes.c_int(len(m)), _ptr(m), _ptr(s), _ptr(out))
        return out

    def loss_grad(self, masks, scalars, targets, weights):
        m, s = self._inputs(masks, scalars)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(len(m), self.cells, 5)
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(len(m), self.cells)
        out = np.zeros((len(m), self.cells, 5), np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_double()
        self._call("dsmc_unet_loss_grad", ctypes.c_int(len(m)), _ptr(m), _ptr(s), _ptr(t), _ptr(w),
                   ctypes.byref(loss), _ptr(out), _ptr(g))
        return loss.value, out, g

    def set_data(self, masks, scalars, targets, weights):
        n = len(masks)
        m = np.ascontiguousarray(masks, dtype=np.uint8).reshape(n, self.edge ** 3)
        s = np.ascontiguousarray(scalars, dtype=np.float32).reshape(n, 2)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(n, self.cells, 5)
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(n, self.cells)
        self._call("dsmc_unet_set_data", ctypes.c_int(n), _ptr(m), _ptr(s), _ptr(t), _ptr(w))

    def train(self, epochs, batch, lr, seed=0):
        losses = np.zeros(max(epochs, 1), np.float32)
        self._call("dsmc_unet_train", ctypes.c_int(epochs), ctypes.c_int(batch), ctypes.c_float(lr),
                   ctypes.c_uint64(seed), _ptr(losses))
        return losses[:epochs]


# ------------------------------------------------------------------ fields
def features(kn, mach):
    """The two scalar inputs: log10 Kn and log10 Mach of the forcing."""
    return np.array([math.log10(kn), math.log10(mach)], np.float32)


def u_scale(ini, mach):
    T, m, _ = C.gas_scales(ini)
    return mach * math.sqrt(5.0 / 3.0 * T / m)


def cell_geometry(types, cells_per_axis):
    """Empty voxels per cell (global cell order) for a cubic world."""
    n = int(round(len(types) ** (1.0 / 3.0)))
    b = n // cells_per_axis
    t = np.asarray(types).reshape(n, n, n)
    c = cells_per_axis
    return (t == 0).reshape(c, b, c, b, c, b).sum(axis=(1, 3, 5)).reshape(-1)


def targets_from_moments(mom, nsamples, empty, n_total, ini, mach):
    """Per-cell normalised targets and loss weights from summed moments
    (ncells, 7) = (count, sum v, sum v^2) over `nsamples` steps."""
    T, m, _ = C.gas_scales(ini)
    us = u_scale(ini, mach)
    cnt = mom[:, 0]
    per_voxel = n_total / max(empty.sum(), 1)
    expect = per_voxel * empty * nsamples
    w = empty / empty.max()
    out = np.zeros((len(cnt), 5))
    ok = (cnt > 0) & (empty > 0)
    out[ok, 0] = cnt[ok] / expect[ok] - 1.0
    u = np.zeros((len(cnt), 3))
    u[ok] = mom[ok, 1:4] / cnt[ok, None]
    out[:, 1:4] = u / us
    var = np.zeros(len(cnt))
    var[ok] = (mom[ok, 4:7] / cnt[ok, None] - u[ok] ** 2).sum(1) / 3.0
    out[ok, 4] = m * var[ok] / T - 1.0
    return out, np.where(ok, w, 0.0)


def fields_from_outputs(out, ini, mach):
    """Network outputs (ncells, 5) -> dsmc_initialize_fields input."""
    T, m, _ = C.gas_scales(ini)
    us = u_scale(ini, mach)
    f = np.zeros((len(out), 5))
    f[:, 0] = np.maximum(1.0 + out[:, 0], 0.05)
    f[:, 1:4] = out[:, 1:4] * us
    f[:, 4] = T * np.clip(1.0 + out[:, 4], 0.5, 1.5)
    return f
