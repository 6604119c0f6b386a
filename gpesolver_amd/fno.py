"""Fourier-neural-operator warm start for imaginary time (ctypes binding of
the gpe_fno_* entry points of include/gpehip.h) plus the feature / target
conventions shared by training and inference.

Input channels per grid point (cin = 6):
  0  V / V0                      potential in units of V0 = 1e-12 hartree
  1  (eps_dd - 1.4) / 0.2
  2  (log10 N - 4.35) / 0.5
  3  (omega_z[Hz] - 120) / 20
  4  potential type (0 HARMONIC, 1 MEXICAN)
  5  dd / 1500 nm
Target: the ground-state amplitude |psi| / max|psi| (real, >= 0).  The warm
start is psi0 = max(prediction, 0), which gpe_load_buffer normalises.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import hip
from . import params as P


class FnoConfig(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("nx", "ny", "nz", "cin", "width", "layers", "hidden", "modes", "max_batch")]


CIN = 6
V0 = 1e-12


def _bind():
    L = hip.load()
    vp, ip, fp, dp = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.gpe_fno_create.argtypes = [C.POINTER(FnoConfig), C.c_ulonglong, C.POINTER(vp)]
    L.gpe_fno_create.restype = ip
    L.gpe_fno_destroy.argtypes = [vp]
    L.gpe_fno_destroy.restype = None
    L.gpe_fno_last_error.argtypes = [vp]
    L.gpe_fno_last_error.restype = C.c_char_p
    L.gpe_fno_num_params.argtypes = [vp]
    L.gpe_fno_num_params.restype = C.c_long
    for n in ("gpe_fno_get_params", "gpe_fno_set_params", "gpe_fno_get_grads"):
        getattr(L, n).argtypes = [vp, fp]
        getattr(L, n).restype = ip
    L.gpe_fno_forward.argtypes = [vp, fp, ip, fp]
    L.gpe_fno_forward.restype = ip
    L.gpe_fno_loss_grad.argtypes = [vp, fp, fp, ip, dp]
    L.gpe_fno_loss_grad.restype = ip
    L.gpe_fno_adam_step.argtypes = [vp, C.c_float]
    L.gpe_fno_adam_step.restype = ip
    L.gpe_fno_input_dev.argtypes = [vp]
    L.gpe_fno_input_dev.restype = vp
    L.gpe_fno_output_dev.argtypes = [vp]
    L.gpe_fno_output_dev.restype = vp
    L.gpe_fno_forward_dev.argtypes = [vp, ip]
    L.gpe_fno_forward_dev.restype = ip
    L.gpe_fno_set_data.argtypes = [vp, fp, fp, C.c_long]
    L.gpe_fno_set_data.restype = ip
    L.gpe_fno_loss_grad_idx.argtypes = [vp, C.POINTER(C.c_int), ip, dp]
    L.gpe_fno_loss_grad_idx.restype = ip
    return L


def _fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Fno:
    def __init__(self, nx=64, ny=64, nz=16, width=32, layers=4, hidden=64, modes=16, max_batch=8, seed=1, cin=CIN):
        self.L = _bind()
        self.cfg = FnoConfig(nx, ny, nz, cin, width, layers, hidden, modes, max_batch)
        f = C.c_void_p()
        rc = self.L.gpe_fno_create(C.byref(self.cfg), seed, C.byref(f))
        if rc != 0:
            msg = self.L.gpe_fno_last_error(f).decode() if f else "gpe_fno_create failed"
            if f:
                self.L.gpe_fno_destroy(f)
            raise hip.GpeError(msg)
        self.f = f
        self.P = nx * ny * nz

    def close(self):
        if getattr(self, "f", None):
            self.L.gpe_fno_destroy(self.f)
            self.f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise hip.GpeError(self.L.gpe_fno_last_error(self.f).decode())

    @property
    def num_params(self) -> int:
        return self.L.gpe_fno_num_params(self.f)

    def get_params(self) -> np.ndarray:
        w = np.empty(self.num_params, dtype=np.float32)
        self._check(self.L.gpe_fno_get_params(self.f, _fp(w)))
        return w

    def set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.num_params
        self._check(self.L.gpe_fno_set_params(self.f, _fp(w)))

    def get_grads(self) -> np.ndarray:
        g = np.empty(self.num_params, dtype=np.float32)
        self._check(self.L.gpe_fno_get_grads(self.f, _fp(g)))
        return g

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        B = x.shape[0]
        y = np.empty((B, self.P), dtype=np.float32)
        self._check(self.L.gpe_fno_forward(self.f, _fp(x), B, _fp(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double()
        self._check(self.L.gpe_fno_loss_grad(self.f, _fp(x), _fp(t), x.shape[0], C.byref(loss)))
        return loss.value

    def input_ptr(self, slot: int = 0) -> int:
        """Device address of batch slot `slot` of the network input, [P][cin]."""
        return self.L.gpe_fno_input_dev(self.f) + slot * self.P * self.cfg.cin * 4

    def output_ptr(self, slot: int = 0) -> int:
        """Device address of the output of batch slot `slot`, [P]."""
        return self.L.gpe_fno_output_dev(self.f) + slot * self.P * 4

    def forward_device(self, batch: int):
        """Forward pass of the first `batch` inputs already on the device;
        the output stays there (output_ptr)."""
        self._check(self.L.gpe_fno_forward_dev(self.f, batch))

    def set_data(self, x: np.ndarray, t: np.ndarray):
        """Keep the training set [n, P, cin] / [n, P] on the device."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert x.shape[0] == t.shape[0]
        self._check(self.L.gpe_fno_set_data(self.f, _fp(x), _fp(t), x.shape[0]))

    def loss_grad_idx(self, idx) -> float:
        """loss_grad on samples idx of the device-resident training set."""
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        self._check(self.L.gpe_fno_loss_grad_idx(self.f, ii.ctypes.data_as(C.POINTER(C.c_int)), ii.size,
                                                 C.byref(loss)))
        return loss.value

    def adam_step(self, lr: float):
        self._check(self.L.gpe_fno_adam_step(self.f, lr))


def feature_consts(p: P.PhysicalParameters, pot: str) -> np.ndarray:
    """Input channels 1 .. CIN-1 of an instance (constant over the grid)."""
    return np.array([(p.edd - 1.4) / 0.2, (math.log10(p.n_atoms) - 4.35) / 0.5,
                     (P.freq_au_to_Hz(p.omega_z) - 120.0) / 20.0, 1.0 if pot.startswith("MEXICAN") else 0.0,
                     P.len_au_to_nm(p.dd) / 1500.0], dtype=np.float32)


def features(p: P.PhysicalParameters, V: np.ndarray, pot: str) -> np.ndarray:
    """Per-point input channels [P, CIN] for one instance (see module doc)."""
    n = V.size
    x = np.empty((n, CIN), dtype=np.float32)
    x[:, 0] = (V.reshape(-1) / V0).astype(np.float32)
    x[:, 1:] = feature_consts(p, pot)
    return x


def target(psi: np.ndarray) -> np.ndarray:
    a = np.abs(psi).reshape(-1)
    return (a / a.max()).astype(np.float32)


# Potentials even in x, y and z about the grid origin (index n/2 on each
# axis: x_i = (i - n/2) d, src/context/context.cpp:17-33)
EVEN_POTENTIALS = ("HARMONIC", "MEXICAN", "MEXICAN_FREE", "CYLINDRICAL", "FREE")


def mirror_average(a: np.ndarray, axes=(0, 1, 2)) -> np.ndarray:
    """Average of a with its mirror images x_i -> -x_i (index i -> n - i)
    along each of `axes`, over the pairs that are both interior (2 <= i <=
    n - 2).  Indices 0, 1 and n - 1 are kept: the image of the interior plane
    1 is the boundary plane n - 1, which imaginary time holds at zero, so the
    grid itself is not symmetric there."""
    a = np.array(a, dtype=np.float64, copy=True)
    for ax in axes:
        b = np.moveaxis(a, ax, 0)
        n = b.shape[0]
        img = b.copy()
        img[2:n - 1] = b[2:n - 1][::-1]
        a = np.moveaxis(0.5 * (b + img), 0, ax)
    return a


def warm_state(pred: np.ndarray, shape, symmetric: bool = False) -> np.ndarray:
    """Initial state from a prediction: interior amplitude, boundary zero
    (like the reference's generated initial states).  symmetric: the
    prediction is first averaged over the mirror images of an even potential,
    the symmetry imaginary time keeps (the plain MULTIPLE_GAUSS start has it
    too); a slightly asymmetric warm state would seed the unstable left / right
    mode of a double well, which imaginary time then takes thousands of steps
    to settle."""
    a = pred.reshape(shape).astype(np.float64)
    if symmetric:
        a = mirror_average(a)
    a = np.maximum(a, 0.0)
    psi = np.zeros(shape, dtype=np.complex128)
    psi[1:-1, 1:-1, 1:-1] = a[1:-1, 1:-1, 1:-1]
    return psi


def time_training(steps: int = 10, batch: int = 8, seed: int = 0) -> dict:
    """Wall time of the config-4 network at 64x64x16 (width 32, 4 layers, 16
    modes, hidden 64) on random inputs: training steps from host arrays,
    forward passes, and training steps from a device-resident set."""
    import time

    grid = (64, 64, 16)
    npts = grid[0] * grid[1] * grid[2]
    rng = np.random.default_rng(seed)
    m = Fno(*grid, width=32, layers=4, hidden=64, modes=16, max_batch=batch, seed=1)
    x = rng.standard_normal((batch, npts, CIN)).astype(np.float32)
    t = rng.random((batch, npts)).astype(np.float32)
    out = {}
    try:
        for _ in range(3):
            m.loss_grad(x, t)
            m.adam_step(1e-3)
        t0 = time.perf_counter()
        for _ in range(steps):
            m.loss_grad(x, t)
            m.adam_step(1e-3)
        out["train_step_host_ms"] = (time.perf_counter() - t0) / steps * 1e3
        t0 = time.perf_counter()
        for _ in range(2 * steps):
            m.forward(x)
        out["forward_ms"] = (time.perf_counter() - t0) / (2 * steps) * 1e3
        xs = rng.standard_normal((8 * batch, npts, CIN)).astype(np.float32)
        ts = rng.random((8 * batch, npts)).astype(np.float32)
        m.set_data(xs, ts)
        for _ in range(3):
            m.loss_grad_idx(np.arange(batch))
            m.adam_step(1e-3)
        t0 = time.perf_counter()
        for s in range(steps):
            m.loss_grad_idx(np.arange(batch * (s % 8), batch * (s % 8) + batch))
            m.adam_step(1e-3)
        out["train_step_device_ms"] = (time.perf_counter() - t0) / steps * 1e3
    finally:
        m.close()
    return out


if __name__ == "__main__":
    import json
    import sys

    print(json.dumps(time_training(int(sys.argv[1]) if len(sys.argv) > 1 else 10)), flush=True)
