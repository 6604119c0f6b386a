"""GPU FNO kernels against the FP64 numpy restatement (oracle/fno_ref.py):
forward output, and gradients of the loss against central finite differences
of the restatement.  The GPU rounds GEMM operands to BF16 (FP32 accumulation),
so agreement is at the BF16 level (a few 1e-3 relative)."""
import numpy as np
import pytest

from gpesolver_amd import fno
from oracle.fno_ref import FnoRef

pytestmark = pytest.mark.gpu

CFG = dict(nx=16, ny=16, nz=8, width=8, layers=2, hidden=16, modes=8, max_batch=4)


def make(seed=5):
    m = fno.Fno(**CFG, seed=seed, cin=3)
    ref = FnoRef(CFG["nx"], CFG["ny"], CFG["nz"], 3, CFG["width"], CFG["layers"], CFG["hidden"], CFG["modes"])
    assert m.num_params == ref.nparam
    rng = np.random.default_rng(seed)
    P = CFG["nx"] * CFG["ny"] * CFG["nz"]
    x = rng.standard_normal((4, P, 3)).astype(np.float32)
    xx, yy, zz = np.meshgrid(np.linspace(-1, 1, 16), np.linspace(-1, 1, 16), np.linspace(-1, 1, 8), indexing="ij")
    t = np.exp(-(xx**2 + yy**2 + zz**2) * 3).reshape(1, -1).repeat(4, 0).astype(np.float32)
    return m, ref, x, t


def test_forward_matches_numpy():
    m, ref, x, t = make()
    w = m.get_params()
    y = m.forward(x)
    yr = ref.forward(w, x)
    assert np.linalg.norm(y - yr) < 2e-2 * np.linalg.norm(yr)


def test_gradients_match_finite_differences():
    m, ref, x, t = make(7)
    w = m.get_params().astype(np.float64)
    loss = m.loss_grad(x, t)
    assert abs(loss - ref.loss(w, x, t)) < 2e-2 * ref.loss(w, x, t)
    g = m.get_grads()
    rng = np.random.default_rng(0)
    idx = []
    for key, (o, n) in ref.off.items():
        # the largest-gradient entries of each parameter group plus random ones
        blk = np.abs(g[o:o + n])
        idx += list(o + np.argsort(blk)[-3:]) + list(o + rng.integers(0, n, 2))
    idx = sorted(set(int(i) for i in idx))
    gfd = ref.fd_grad(w, x, t, idx, eps=1e-3)
    gg = g[idx].astype(np.float64)
    cos = gg @ gfd / (np.linalg.norm(gg) * np.linalg.norm(gfd))
    assert cos > 0.995, cos
    big = np.abs(gfd) > 0.05 * np.abs(gfd).max()
    rel = np.abs(gg[big] - gfd[big]) / np.abs(gfd[big])
    assert np.median(rel) < 0.03 and rel.max() < 0.25, (np.median(rel), rel.max())


def test_adam_reduces_loss():
    m, ref, x, t = make(9)
    l0 = m.loss_grad(x, t)
    for _ in range(30):
        m.loss_grad(x, t)
        m.adam_step(3e-3)
    l1 = m.loss_grad(x, t)
    assert l1 < 0.5 * l0, (l0, l1)


def test_config4_size_forward_loss_and_gradients():
    """The warm-start network at its stated size (64x64x16, width 32, 4
    spectral layers, 16 modes per axis capped at nz/2+1 = 9 along z, six
    input channels): forward pass, loss and the largest-gradient entry of
    every parameter group against the FP64 restatement (central finite
    differences).  BF16 operands, so agreement is at the BF16 level."""
    cfg = dict(nx=64, ny=64, nz=16, width=32, layers=4, hidden=64, modes=16, max_batch=1)
    m = fno.Fno(**cfg, seed=21, cin=fno.CIN)
    ref = FnoRef(64, 64, 16, fno.CIN, 32, 4, 64, 16)
    assert m.num_params == ref.nparam == 18880993
    rng = np.random.default_rng(21)
    P = 64 * 64 * 16
    x = (0.5 * rng.standard_normal((1, P, fno.CIN))).astype(np.float32)
    xx, yy, zz = np.meshgrid(np.linspace(-1, 1, 64), np.linspace(-1, 1, 64), np.linspace(-1, 1, 16), indexing="ij")
    t = np.exp(-(xx**2 + yy**2 + zz**2) * 3).reshape(1, -1).astype(np.float32)
    w = m.get_params().astype(np.float64)
    y = m.forward(x)
    yr = ref.forward(w, x)
    fwd = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    loss = m.loss_grad(x, t)
    lr_ = ref.loss(w, x, t)
    print(f"\nconfig-4 FNO: forward rel L2 {fwd:.2e}, loss {loss:.6e} vs {lr_:.6e}", flush=True)
    assert fwd < 3e-2
    assert abs(loss - lr_) < 3e-2 * lr_
    g = m.get_grads()
    idx = sorted({int(o + np.argmax(np.abs(g[o:o + n]))) for (o, n) in ref.off.values()})
    gfd = ref.fd_grad(w, x, t, idx, eps=1e-3)
    gg = g[idx].astype(np.float64)
    cos = gg @ gfd / (np.linalg.norm(gg) * np.linalg.norm(gfd))
    big = np.abs(gfd) > 0.05 * np.abs(gfd).max()
    rel = np.abs(gg[big] - gfd[big]) / np.abs(gfd[big])
    print(f"config-4 FNO gradients ({len(idx)} entries): cosine {cos:.5f}, median rel {np.median(rel):.2e}, "
          f"max rel {rel.max():.2e}", flush=True)
    assert cos > 0.99
    assert np.median(rel) < 0.05 and rel.max() < 0.3
    m.close()


def test_device_training_set_matches_host_batches():
    """loss_grad_idx on the device-resident training set equals loss_grad on
    the same samples passed from the host, loss and gradients bit for bit
    (the batch is the same inputs, copied on the device)."""
    m, ref, x, t = make(9)
    t = t * np.linspace(0.5, 1.5, 4, dtype=np.float32)[:, None]
    m.set_data(x, t)
    idx = np.array([3, 1, 2, 0], dtype=np.int32)  # 4 x 2048 points: the 8192-point granule of the loss
    l_dev = m.loss_grad_idx(idx)
    g_dev = m.get_grads()
    l_host = m.loss_grad(x[idx], t[idx])
    g_host = m.get_grads()
    assert l_dev == l_host
    np.testing.assert_array_equal(g_dev, g_host)
    m.close()
