"""gpe_calc_energy_noalloc (the allocation-free variant of gpe_calc_energy,
include/gpehip.h) against gpe_calc_energy, through the C ABI: two calls in a
row, the second on a smaller grid, in both precisions, one of them from a
pending real-time state.  Both entry points add the same FP64 per-plane
partials in the same order; only the final formulas run on the device
instead of the host, so the values agree to 1e-13 relative."""
import ctypes as C

import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

KEYS = ("e_kin", "e_pot", "e_int", "e_ext", "e_bmf", "e_total", "mu", "residual")


def _hiprt():
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return L


def _case(shape, prec, rt_steps):
    nx, ny, nz = shape
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5)
    V = P.potential("HARMONIC", p)
    psi = P.add_phase_noise(P.wavefunction("GAUSS", p, sigma_cells=3.0), 1e-2, 7 + nx)
    dev = hip.Device(p.to_gpe(), prec)
    dev.load_pote(V)
    dev.load_buffer(psi)
    dev.calc_fi3d()
    dev.it_steps(5)
    if rt_steps:
        dev.rt_steps(rt_steps)  # leaves a pending half phase: the noalloc call applies it
    return dev


def _noalloc(dev, hrt, out_dev):
    print(f"[no_alloc] shape={dev.shape} prec={dev.precision}", flush=True)
    assert hrt.hipMemset(out_dev, 0xFF, 64) == 0
    dev.energy_noalloc(out_dev.value)
    dev.synchronize()
    got = np.zeros(8)
    assert hrt.hipMemcpy(got.ctypes.data_as(C.c_void_p), out_dev, 64, 2) == 0  # device to host
    return got


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
def test_noalloc_energy_matches_calc_energy(prec):
    hrt = _hiprt()
    out_dev = C.c_void_p()
    assert hrt.hipMalloc(C.byref(out_dev), 64) == 0
    try:
        for shape, rt in (((32, 24, 12), 3), ((16, 16, 8), 0)):
            dev = _case(shape, prec, rt)
            got = _noalloc(dev, hrt, out_dev)
            ref = dev.energy()
            want = np.array([ref[k] for k in KEYS])
            scale = np.maximum(np.abs(want), abs(ref["e_total"]) * 1e-3)
            err = np.abs(got - want) / scale
            print(f"[no_alloc] max rel diff {err.max():.3e}", flush=True)
            assert np.all(np.isfinite(got))
            assert err.max() < 1e-13, dict(zip(KEYS, err))
            dev.close()
    finally:
        hrt.hipFree(out_dev)
