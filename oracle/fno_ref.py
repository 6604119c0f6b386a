"""FP64 numpy restatement of the FNO warm-start network (the checker for the
gpe_fno_* kernels).  TEST INFRASTRUCTURE ONLY: imported by tests/ as the
oracle, never by the product.

The network has no counterpart in the reference; it is restated here from
its definition (gpesolver_amd/csrc/gpe_fno.hip header) using numpy's FFTs
instead of the DFT matrices the GPU uses, with the same flat parameter layout:
  W_lift [cin][W], b_lift [W]
  per layer: R_re [nm][W][W], R_im [nm][W][W], W_l [W][W], b_l [W]
  W_q1 [W][H], b_q1 [H], W_q2 [H], b_q2 [1]
with modes kx, ky in {0..m/2-1} U {n-m/2..n-1} and kz in {0..mz-1},
mz = min(modes, nz/2+1), nm = mx*my*mz ordered (kx, ky, kz).
Gradients are taken by central finite differences in FP64 (tests pick
entries), which pins them by the definition of the derivative.
"""
from __future__ import annotations

import numpy as np


def gelu(x):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x**3)))


class FnoRef:
    def __init__(self, nx, ny, nz, cin, width, layers, hidden, modes):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cin, self.W, self.L, self.H = cin, width, layers, hidden
        self.mx, self.my = min(modes, nx), min(modes, ny)
        self.mz = min(modes, nz // 2 + 1)
        self.nm = self.mx * self.my * self.mz
        self.kx = list(range(self.mx // 2)) + list(range(nx - self.mx // 2, nx))
        self.ky = list(range(self.my // 2)) + list(range(ny - self.my // 2, ny))
        W, H = width, hidden
        sizes = [("lift_w", cin * W), ("lift_b", W)]
        for l in range(layers):
            sizes += [(f"r{l}_re", self.nm * W * W), (f"r{l}_im", self.nm * W * W), (f"w{l}", W * W), (f"b{l}", W)]
        sizes += [("q1_w", W * H), ("q1_b", H), ("q2_w", H), ("q2_b", 1)]
        self.off = {}
        o = 0
        for k, n in sizes:
            self.off[k] = (o, n)
            o += n
        self.nparam = o

    def view(self, w, key, shape):
        o, n = self.off[key]
        return np.asarray(w[o:o + n], dtype=np.float64).reshape(shape)

    def spectral(self, h, Rre, Rim):
        nx, ny, nz, mz = self.nx, self.ny, self.nz, self.mz
        Hk = np.fft.rfft(h, axis=3)[:, :, :, :mz, :]
        Hk = np.fft.fft(Hk, axis=2)[:, :, self.ky, :, :]
        Hk = np.fft.fft(Hk, axis=1)[:, self.kx, :, :, :]
        R = (Rre + 1j * Rim).reshape(self.mx, self.my, mz, self.W, self.W)
        Mk = np.einsum("bxyzi,xyzio->bxyzo", Hk, R)
        full = np.zeros((h.shape[0], nx, ny, nz // 2 + 1, self.W), dtype=np.complex128)
        for a, kx in enumerate(self.kx):
            for b, ky in enumerate(self.ky):
                full[:, kx, ky, :mz, :] = Mk[:, a, b, :, :]
        return np.fft.irfftn(full, s=(nx, ny, nz), axes=(1, 2, 3))

    def forward(self, w, x):
        """x [B, P, cin] -> y [B, P]."""
        B = x.shape[0]
        nx, ny, nz, W, H = self.nx, self.ny, self.nz, self.W, self.H
        xg = np.asarray(x, dtype=np.float64).reshape(B, nx, ny, nz, self.cin)
        h = xg @ self.view(w, "lift_w", (self.cin, W)) + self.view(w, "lift_b", (W,))
        for l in range(self.L):
            s = self.spectral(h, self.view(w, f"r{l}_re", (self.nm, W, W)), self.view(w, f"r{l}_im", (self.nm, W, W)))
            z = h @ self.view(w, f"w{l}", (W, W)) + self.view(w, f"b{l}", (W,)) + s
            h = gelu(z) if l < self.L - 1 else z
        q = gelu(h @ self.view(w, "q1_w", (W, H)) + self.view(w, "q1_b", (H,)))
        y = q @ self.view(w, "q2_w", (H, 1)) + self.view(w, "q2_b", (1,))
        return y.reshape(B, -1)

    def loss(self, w, x, t):
        y = self.forward(w, x)
        return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))

    def fd_grad(self, w, x, t, idx, eps=1e-4):
        """Central finite differences of the loss at parameter indices idx."""
        w = np.asarray(w, dtype=np.float64).copy()
        g = np.zeros(len(idx))
        for n, i in enumerate(idx):
            h = eps * max(1.0, abs(w[i]))
            w0 = w[i]
            w[i] = w0 + h
            lp = self.loss(w, x, t)
            w[i] = w0 - h
            lm = self.loss(w, x, t)
            w[i] = w0
            g[n] = (lp - lm) / (2 * h)
        return g
