"""The warm state's mirror averaging (gpesolver_amd.fno.mirror_average):
it fixes what the plain MULTIPLE_GAUSS start and the even potentials already
satisfy exactly (src/initializer/initializer.cpp:162-219,
src/parameters/potentials.hpp:76-157 on x_i = (i - n/2) d), and it is a
projection."""
import numpy as np

from gpesolver_amd import fno as F
from gpesolver_amd import params as P


def _p():
    return P.PhysicalParameters.from_physical(24, 20, 12, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5, dd_nm=600.0)


def test_plain_start_and_even_potentials_are_fixed_points():
    p = _p()
    psi = np.abs(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2, sigma_cells=3.0))
    assert np.array_equal(F.mirror_average(psi), psi)
    for pot in ("HARMONIC", "MEXICAN"):
        V = P.potential(pot, p)
        assert np.abs(F.mirror_average(V) - V).max() <= 1e-15 * np.abs(V).max()


def test_projection_and_interior_pairs():
    rng = np.random.default_rng(0)
    a = rng.random((9, 8, 6))
    s = F.mirror_average(a)
    np.testing.assert_allclose(F.mirror_average(s), s, rtol=0, atol=1e-15)
    # pairs (i, n - i) with 2 <= i <= n - 2 are equal; planes 0, 1, n - 1 kept
    for ax, n in enumerate(a.shape):
        b, sb = np.moveaxis(a, ax, 0), np.moveaxis(s, ax, 0)
        for i in range(2, n - 1):
            np.testing.assert_allclose(sb[i], sb[n - i], rtol=0, atol=1e-15)
    assert np.array_equal(s[0], F.mirror_average(a)[0])
    # the asymmetric potential is not symmetrised by the warm start
    assert "MEXICAN_ASYMETRIC" not in F.EVEN_POTENTIALS
