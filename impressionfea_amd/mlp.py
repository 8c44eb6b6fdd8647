"""Learned warm start: a per-node MLP that predicts each stroke increment's nodal velocity field.

The network (11 -> 128 -> 128 -> 128 -> 3, GELU; its products on BF16 MFMA with FP32 accumulation) runs and trains
through the hand-written HIP kernels behind include/ifea_mlp.h. Inputs per node, all dimensionless: position / billet size (3),
the last two converged nodal velocities / ram speed (6), die-contact flag (1), distance to the nearest die / billet
size (1). Output: the change of the nodal velocity over the coming increment / ram speed. The prediction plugs into
the reference's LoadVelocityGuess slot (Forging3DSolver.cs:854-869); the solver judges it (see Forging3DSolver.step)
and the previous-increment start stays the fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .errors import raise_for

N_IN, WIDTH, DEPTH, N_OUT = 11, 128, 3, 3


def _lib():
    L = N.lib()
    if not getattr(L, "_mlp_bound", False):
        vp, fp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int)
        L.ifea_mlp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int]
        L.ifea_mlp_num_params.restype = C.c_longlong
        L.ifea_mlp_num_params.argtypes = [vp]
        L.ifea_mlp_set_params.argtypes = [vp, fp]
        L.ifea_mlp_get_params.argtypes = [vp, fp]
        L.ifea_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
        L.ifea_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, fp, fp]
        L.ifea_mlp_set_dataset.argtypes = [vp, C.c_int, fp, fp]
        L.ifea_mlp_train_epoch.argtypes = [vp, ip, C.c_int, C.c_float, fp]
        L.ifea_mlp_destroy.argtypes = [vp]
        L.ifea_mlp_destroy.restype = None
        L._mlp_bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def layer_shapes(n_in=N_IN, width=WIDTH, depth=DEPTH, n_out=N_OUT):
    fin = [n_in] + [width] * depth
    fout = [width] * depth + [n_out]
    return list(zip(fin, fout))


def init_params(seed=0, n_in=N_IN, width=WIDTH, depth=DEPTH, n_out=N_OUT):
    """He-style initialisation; the last layer starts small so an untrained network predicts ~no change."""
    rng = np.random.default_rng(seed)
    parts = []
    shapes = layer_shapes(n_in, width, depth, n_out)
    for l, (fi, fo) in enumerate(shapes):
        scale = np.sqrt(2.0 / fi) * (0.1 if l == len(shapes) - 1 else 1.0)
        parts.append((rng.standard_normal((fi, fo)) * scale).ravel())
        parts.append(np.zeros(fo))
    return np.concatenate(parts).astype(np.float32)


class MlpDevice:
    """The network's parameters and kernels on the GPU."""

    def __init__(self, params=None, n_in=N_IN, width=WIDTH, depth=DEPTH, n_out=N_OUT):
        L = _lib()
        h = C.c_void_p()
        rc = L.ifea_mlp_create(C.byref(h), n_in, width, depth, n_out)
        if rc:
            raise_for(rc, "ifea_mlp_create", N.last_error())
        self._h = h
        self.n_in, self.n_out = n_in, n_out
        self.np = int(L.ifea_mlp_num_params(h))
        self.set_params(init_params(0, n_in, width, depth, n_out) if params is None else params)

    def _check(self, rc, what):
        if rc:
            raise_for(rc, what, N.last_error())

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.np
        self._check(_lib().ifea_mlp_set_params(self._h, N.ptr(p, C.c_float)), "set_params")

    def get_params(self):
        p = np.zeros(self.np, np.float32)
        self._check(_lib().ifea_mlp_get_params(self._h, N.ptr(p, C.c_float)), "get_params")
        return p

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._check(_lib().ifea_mlp_forward(self._h, x.shape[0], N.ptr(x, C.c_float), N.ptr(y, C.c_float)), "forward")
        return y

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.np, np.float32)
        self._check(_lib().ifea_mlp_loss_grad(self._h, x.shape[0], N.ptr(x, C.c_float), N.ptr(y, C.c_float),
                                              N.ptr(loss, C.c_float), N.ptr(g, C.c_float)), "loss_grad")
        return float(loss[0]), g

    def set_dataset(self, x, y):
        self._x, self._y = _f32(x), _f32(y)
        self.ndata = self._x.shape[0]
        self._check(_lib().ifea_mlp_set_dataset(self._h, self.ndata, N.ptr(self._x, C.c_float),
                                                N.ptr(self._y, C.c_float)), "set_dataset")

    def train_epoch(self, perm, batch=4096, lr=1e-3):
        p = N.i32(perm)
        loss = np.zeros(1, np.float32)
        self._check(_lib().ifea_mlp_train_epoch(self._h, N.ptr(p, C.c_int), int(batch), float(lr),
                                                N.ptr(loss, C.c_float)), "train_epoch")
        return float(loss[0])

    def close(self):
        if getattr(self, "_h", None):
            _lib().ifea_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def node_features(solver):
    """Per-node inputs at the start of an increment (before PrepareStep), from the solver's host state: coordinates,
    the last two converged nodal velocities, the contact flag and the distance to the nearer die surface (11 values,
    lengths scaled by the billet size and velocities by the ram speed)."""
    cfg = solver.cfg
    L = cfg.size
    x = solver.x
    vg = solver.vg
    vp = solver.vg_prev if solver.n_converged >= 2 else solver.vg
    contact = (solver.contact != 0).astype(np.float64)
    kept = getattr(solver, "clearance", None)
    if kept is not None and kept[0] is x:
        cb, ct = kept[1], kept[2]          # computed by the geometry update that produced x
    else:
        cb = cfg.bottom.clearance(x, 0.0)
        ct = cfg.top.clearance(x, solver.gap)
    dist = np.minimum(np.abs(cb), np.abs(ct))
    f = np.concatenate([x / L, vg / cfg.ram, vp / cfg.ram, contact[:, None], (dist / L)[:, None]], axis=1)
    return f.astype(np.float32)


def trend_start(solver):
    """The trend start 2 v_n - v_{n-1} (ExtrapolateStart, Forging3DSolver.cs:856-859) once two converged increments
    exist (before that, the previous increment's field)."""
    if solver.n_converged < 2:
        return solver.vg.copy()
    return 2.0 * solver.vg - solver.vg_prev


def target(base, vg_after, ram, scale):
    """Training target of one increment: the converged velocity's deviation from the start `base` (the trend start),
    in units of scale * ram."""
    return ((vg_after - base) / (ram * scale)).astype(np.float32)


class LearnedWarmStart:
    """Warm-start callback for Forging3DSolver: v_guess = (2 v_n - v_{n-1}) + scale * ram * net(features), i.e. the
    network predicts the correction to the trend start."""

    def __init__(self, params, scale=1.0):
        self.params = np.asarray(params, np.float32)
        self.scale = float(scale)
        self.dev = MlpDevice(self.params)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], float(z["scale"]) if "scale" in z.files else 1.0)

    def save(self, path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.dev.get_params(), scale=np.float64(self.scale))

    def __call__(self, solver, trend=None):
        dv = self.dev.forward(node_features(solver)).astype(np.float64)
        return (trend_start(solver) if trend is None else trend) + self.scale * solver.cfg.ram * dv
