"""Pins the oracle's Poisson (dipolar) solve, IT step, RT split step and
energies against what the mathematics fixes: closed forms, invariants, dense
eigen-solves on tiny grids, and numpy's FFT.  None of these checks retypes the
oracle's formula: each is chosen so that a dropped term, a wrong sign, a wrong
k-lattice or a transposed axis fails it.

Reference functions restated by the oracle:
  Poisson   src/solver/fft_solvers/cpu/fftw_poisson_solver.cpp:12-96
  IT step   src/solver/cpu_solver/cpu_solver.cpp:50-97, solver.cpp:99-105
  norm      src/solver/cpu_solver/cpu_solver.cpp:19-48
  RT step   src/solver/cpu_solver/cpu_solver.cpp:99-131,
            src/solver/fft_solvers/cpu/fftw_split_solver.cpp:11-86, solver.cpp:107-112
  energies  src/solver/cpu_solver/cpu_solver.cpp:137-194
"""
import numpy as np
import pytest
import scipy.linalg


def small_params(oracle, nx=8, ny=8, nz=6, **kw):
    base = dict(dx_nm=120.0, dy_nm=120.0, dz_nm=600.0, n_atoms=2e4, edd=1.5)
    base.update(kw)
    return oracle.make_params(nx, ny, nz, base.pop("dx_nm"), base.pop("dy_nm"), base.pop("dz_nm"), **base)


def phi_of(oracle, psi, d_nm, n_atoms=1.0):
    nx, ny, nz = psi.shape
    p = oracle.make_params(nx, ny, nz, *d_nm, n_atoms=n_atoms)
    o = oracle.Oracle(p)
    o.psi[...] = psi
    o.calc_fi3d()
    return o.fi3d.copy()


# ------------------------------------------------------------------ Vdip table
def test_vdip_invariants(oracle):
    p = small_params(oracle, 6, 8, 4)
    o = oracle.Oracle(p)
    V = o.vdip()
    assert V.shape == (12, 16, 5)
    assert np.all(V >= 0) and np.all(V <= 1)
    assert V[0, 0, 0] == 0.0
    # kx = ky = 0 line, kz != 0: Vdip = 1
    np.testing.assert_array_equal(V[0, 0, 1:], 1.0)
    # kz = 0 plane: Vdip = 0
    np.testing.assert_array_equal(V[:, :, 0], 0.0)
    # even in kx and ky
    np.testing.assert_array_equal(V[1:, :, :], V[:0:-1, :, :])
    np.testing.assert_array_equal(V[:, 1:, :], V[:, :0:-1, :])


# ------------------------------------------------------------------ Poisson
@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (8, 6, 4)])
def test_poisson_axis_sum_identity(oracle, shape):
    """kx^2/k^2 + ky^2/k^2 + kz^2/k^2 = 1 for k != 0.  Computing the oracle's
    Phi (which applies kz^2/k^2) on axis-permuted copies gives the x and y
    multipliers; the three must add up to rho - mean(rho_pad) on the corner.
    A wrong k-lattice, padding, normalisation or axis fails this."""
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.0, 1.0, shape)
    psi = np.sqrt(rho).astype(np.complex128)
    d = (120.0, 150.0, 600.0)
    phz = phi_of(oracle, psi, d)
    phx = phi_of(oracle, np.ascontiguousarray(psi.transpose(2, 1, 0)), (d[2], d[1], d[0])).transpose(2, 1, 0)
    phy = phi_of(oracle, np.ascontiguousarray(psi.transpose(0, 2, 1)), (d[0], d[2], d[1])).transpose(0, 2, 1)
    np_pad = 8 * rho.size
    expect = rho - rho.sum() / np_pad
    np.testing.assert_allclose(phx + phy + phz, expect, atol=1e-12)


def test_poisson_vs_numpy_rfft(oracle):
    """Same padded convolution with numpy's r2c/c2r (a library FFT)."""
    shape = (8, 6, 4)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    d = (120.0, 120.0, 600.0)
    n_atoms = 2e4
    got = phi_of(oracle, psi, d, n_atoms)
    L = oracle.lib()
    dx, dy, dz = (L.orc_len_nm_to_au(v) for v in d)
    Nx, Ny, Nz = (2 * s for s in shape)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, dx)
    ky = 2 * np.pi * np.fft.fftfreq(Ny, dy)
    kz = 2 * np.pi * np.fft.rfftfreq(Nz, dz)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX**2 + KY**2 + KZ**2
    mult = np.where(k2 > 1e-30, KZ**2 / np.where(k2 > 0, k2, 1), 0.0)
    pad = np.zeros((Nx, Ny, Nz))
    pad[: shape[0], : shape[1], : shape[2]] = n_atoms * np.abs(psi) ** 2
    phi = np.fft.irfftn(np.fft.rfftn(pad) * mult, s=(Nx, Ny, Nz), axes=(0, 1, 2))[: shape[0], : shape[1], : shape[2]]
    np.testing.assert_allclose(got, phi, rtol=1e-12, atol=1e-12 * np.abs(phi).max())


def test_poisson_direct_sum_tiny(oracle):
    """Brute force O(Np^2) real-space sum on a 2x2x2 grid (padded 4x4x4)."""
    shape = (2, 2, 2)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.1, 1.0, shape)
    d = (100.0, 200.0, 300.0)
    got = phi_of(oracle, np.sqrt(rho).astype(np.complex128), d)
    L = oracle.lib()
    dd = [L.orc_len_nm_to_au(v) for v in d]
    Np = (4, 4, 4)
    # multiplier on the full padded lattice
    ks = [2 * np.pi * np.array([(i if i <= n // 2 else i - n) for i in range(n)]) / (n * h) for n, h in zip(Np, dd)]
    out = np.zeros(shape)
    for a in np.ndindex(shape):
        acc = 0.0
        for kk in np.ndindex(Np):
            k = [ks[ax][kk[ax]] for ax in range(3)]
            k2 = sum(v * v for v in k)
            mult = k[2] ** 2 / k2 if k2 > 1e-30 else 0.0
            for b in np.ndindex(shape):
                ph = sum(2 * np.pi * kk[ax] * (a[ax] - b[ax]) / Np[ax] for ax in range(3))
                acc += mult * rho[b] * np.cos(ph)
        out[a] = acc / np.prod(Np)
    np.testing.assert_allclose(got, out, atol=1e-13)


# ------------------------------------------------------------------ IT step
def test_it_linear_sine_mode_closed_form(oracle):
    """With V = g = gamma = cdd = 0, sin(pi p i/(n-1)) products are eigenvectors
    of the Dirichlet FD Laplacian: psi' = psi (1 - dtau*lambda) on the interior,
    lambda = sum_ax (2 - 2cos(pi p_ax/(n_ax-1))) / (2 m d_ax^2)."""
    p = small_params(oracle, 10, 9, 7)
    o = oracle.Oracle(p)
    P = o.params
    P.cdd = 0.0
    P.ggp11 = 0.0
    P.gamma = 0.0
    modes = (2, 1, 3)
    n = (P.nx, P.ny, P.nz)
    d = (P.dx, P.dy, P.dz)
    axes = [np.sin(np.pi * m * np.arange(N) / (N - 1)) for m, N in zip(modes, n)]
    psi = np.einsum("i,j,k->ijk", *axes).astype(np.complex128) * (1 + 0.5j)
    o.psi[...] = psi
    o.psii[...] = 0
    o.imag_linear()
    o.imag_nonlinear()
    lam = sum((2 - 2 * np.cos(np.pi * m / (N - 1))) / (2 * P.m * h * h) for m, N, h in zip(modes, n, d))
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(o.psi[inner], psi[inner] * (1 - P.imag_time_dt * lam), rtol=1e-12, atol=1e-14)
    # the boundary is taken from cpsii (zero here): cpu_solver.cpp:96
    assert np.all(o.psi[0] == 0) and np.all(o.psi[-1] == 0)


def dense_fd_hamiltonian(P, V):
    nx, ny, nz = V.shape
    ix = [np.arange(1, nx - 1), np.arange(1, ny - 1), np.arange(1, nz - 1)]
    sizes = [len(a) for a in ix]
    Hs = []
    for n_in, h in zip(sizes, (P.dx, P.dy, P.dz)):
        T = (np.diag(np.full(n_in, 2.0)) - np.diag(np.ones(n_in - 1), 1) - np.diag(np.ones(n_in - 1), -1))
        Hs.append(T / (2 * P.m * h * h))
    I = [np.eye(s) for s in sizes]
    H = (np.kron(np.kron(Hs[0], I[1]), I[2]) + np.kron(np.kron(I[0], Hs[1]), I[2]) + np.kron(np.kron(I[0], I[1]), Hs[2]))
    H += np.diag(V[1:-1, 1:-1, 1:-1].ravel())
    return H


def test_it_converges_to_dense_ground_state(oracle):
    """Linear harmonic case on a tiny grid: repeated IT steps converge to the
    lowest eigenvector of the dense Dirichlet FD Hamiltonian (scipy eigh), and
    the oracle's mu and residual agree with it."""
    p = small_params(oracle, 8, 8, 6, n_atoms=1.0, dx_nm=400.0, dy_nm=400.0, dz_nm=800.0)
    o = oracle.Oracle(p)
    P = o.params
    P.cdd = 0.0
    P.ggp11 = 0.0
    P.gamma = 0.0
    V = oracle.potential("HARMONIC", P)
    o.load_pote(V)
    psi0 = oracle.wavefunction("GAUSS", P)
    o.load_buffer(psi0)
    H = dense_fd_hamiltonian(P, V)
    w, vecs = scipy.linalg.eigh(H)
    gap = w[1] - w[0]
    nsteps = int(np.ceil(40.0 / (gap * P.imag_time_dt)))
    assert nsteps < 20000
    o.it_steps(nsteps)
    e = o.energy()
    g = vecs[:, 0].reshape(6, 6, 4)
    got = o.psi[1:-1, 1:-1, 1:-1].real
    got = got / np.linalg.norm(got)
    g = g * np.sign(np.sum(g * got))
    np.testing.assert_allclose(got, g, atol=1e-8)
    assert abs(e["mu"] - w[0]) < 1e-9 * abs(w[0])
    assert e["residual"] < 1e-6 * abs(w[0])
    # With no interactions E/N equals mu (E_kin + E_pot, n_atoms = 1).
    assert abs(e["e_total"] - e["mu"]) < 1e-10 * abs(w[0])


def test_norm_and_normalize(oracle):
    p = small_params(oracle, 6, 5, 4)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal((6, 5, 4)) + 1j * rng.standard_normal((6, 5, 4))
    o.load_buffer(psi)
    assert abs(np.sum(np.abs(o.psi) ** 2) * o.params.dxdydz - 1.0) < 1e-14
    # cpsii keeps the un-normalised copy (solver.cpp:118-125)
    np.testing.assert_array_equal(o.psii, psi)


def test_energy_mu_relation(oracle):
    """mu = [E_kin + E_pot + 2 E_int + 2 E_dd + 5/2 E_bmf] / N for any state:
    the energies (cpu_solver.cpp:137-194) and <psi|H|psi> are computed by two
    different loops in the oracle."""
    p = small_params(oracle, 12, 10, 8)
    o = oracle.Oracle(p)
    P = o.params
    o.load_pote(oracle.potential("MEXICAN", P))
    rng = np.random.default_rng(5)
    psi = oracle.wavefunction("MULTIPLE_GAUSS", P) * np.exp(1j * rng.uniform(-0.3, 0.3, (12, 10, 8)))
    o.load_buffer(psi)
    o.calc_fi3d()
    e = o.energy()
    mu2 = (e["e_kin"] + e["e_pot"] + 2 * e["e_int"] + 2 * e["e_ext"] + 2.5 * e["e_bmf"]) / P.n_atoms
    assert abs(e["mu"] - mu2) < 1e-12 * abs(mu2)


def test_kinetic_energy_sine_mode(oracle):
    p = small_params(oracle, 10, 9, 7, n_atoms=3.0)
    o = oracle.Oracle(p)
    P = o.params
    modes = (1, 2, 1)
    n = (P.nx, P.ny, P.nz)
    d = (P.dx, P.dy, P.dz)
    axes = [np.sin(np.pi * m * np.arange(N) / (N - 1)) for m, N in zip(modes, n)]
    psi = np.einsum("i,j,k->ijk", *axes).astype(np.complex128)
    o.psi[...] = psi
    e = o.energy()
    lam = sum((2 - 2 * np.cos(np.pi * m / (N - 1))) / (h * h) for m, N, h in zip(modes, n, d))
    expect = P.dxdydz / (2 * P.m) * P.n_atoms * lam * np.sum(np.abs(psi[1:-1, 1:-1, 1:-1]) ** 2)
    assert abs(e["e_kin"] - expect) < 1e-12 * expect


# ------------------------------------------------------------------ RT step
def test_kinetic_step_plane_wave(oracle):
    """A lattice plane wave is an eigenstate of the kinetic split step:
    psi <- psi * exp(-i dt k^2 / 2m) exactly (fftw_split_solver.cpp:11-86)."""
    p = small_params(oracle, 8, 6, 4)
    o = oracle.Oracle(p)
    P = o.params
    n = (P.nx, P.ny, P.nz)
    d = (P.dx, P.dy, P.dz)
    q = (3, -2, 1)
    grids = np.meshgrid(*[np.arange(N) for N in n], indexing="ij")
    phase = sum(2 * np.pi * qq * g / N for qq, g, N in zip(q, grids, n))
    psi = np.exp(1j * phase)
    o.psi[...] = psi
    o.kinetic_step()
    k2 = sum((2 * np.pi * qq / (N * h)) ** 2 for qq, N, h in zip(q, n, d))
    np.testing.assert_allclose(o.psi, psi * np.exp(-1j * P.real_time_dt * k2 / (2 * P.m)), atol=1e-13)


def test_rt_norm_conservation_and_reversibility(oracle):
    p = small_params(oracle, 12, 10, 8)
    o = oracle.Oracle(p)
    P = o.params
    o.load_pote(oracle.potential("MEXICAN", P))
    o.load_buffer(oracle.wavefunction("MULTIPLE_GAUSS", P))
    o.calc_fi3d()
    n0 = np.sum(np.abs(o.psi) ** 2)
    start = o.psi.copy()
    for _ in range(20):
        o.rt_step()
    assert abs(np.sum(np.abs(o.psi) ** 2) / n0 - 1) < 1e-12
    # Potential half step alone: unitary pointwise phase, interior only.
    before = o.psi.copy()
    o.potential_half_step()
    np.testing.assert_allclose(np.abs(o.psi), np.abs(before), rtol=1e-13)
    np.testing.assert_array_equal(o.psi[0], before[0])
    np.testing.assert_array_equal(o.psi[:, :, -1], before[:, :, -1])
    assert not np.allclose(o.psi[1:-1, 1:-1, 1:-1], before[1:-1, 1:-1, 1:-1], rtol=0, atol=0)
    # Kinetic step is undone by the same step with dt -> -dt.
    k = o.psi.copy()
    o.kinetic_step()
    P.real_time_dt = -P.real_time_dt
    o.refresh()
    o.kinetic_step()
    np.testing.assert_allclose(o.psi, k, atol=1e-13 * np.abs(k).max())
    assert start.shape == o.psi.shape


def test_rt_strang_second_order(oracle):
    """Strang splitting: halving dt over a fixed interval cuts the error against
    a fine-dt solution by ~4x (order 2)."""
    p = small_params(oracle, 12, 10, 8, n_atoms=2e4)
    ref_psi = None
    errs = []
    for nsteps in (256, 16, 32):
        P0 = small_params(oracle, 12, 10, 8, n_atoms=2e4)
        P0.real_time_dt = 2.0e10 / nsteps
        o = oracle.Oracle(P0)
        o.load_pote(oracle.potential("HARMONIC", P0))
        o.load_buffer(oracle.wavefunction("GAUSS", P0) * np.exp(0.3j * np.arange(12)[:, None, None]))
        o.calc_fi3d()
        o.rt_steps(0, nsteps)
        if ref_psi is None:
            ref_psi = o.psi.copy()
        else:
            errs.append(np.linalg.norm(o.psi - ref_psi))
    ratio = errs[0] / errs[1]
    assert 3.0 < ratio < 5.5, ratio
    assert p.nx == 12


def test_edd_ramp_update(oracle):
    """update_edd (parameters.cpp:53-59): edd linear in iter, g and gamma follow."""
    p = oracle.make_params(8, 8, 4, 120, 120, 600, edd=None, edd_start=1.5, edd_stop=1.3, iter_real=200)
    L = oracle.lib()
    import ctypes as C
    L.orc_update_edd(C.byref(p), 50)
    assert abs(p.edd - 1.45) < 1e-15
    a = p.add / p.edd
    assert abs(p.ggp11 - 4 * np.pi * a / p.m) < 1e-15 * p.ggp11
    g = 128 * np.sqrt(np.pi) * a**2.5 / 3 / p.m * (1 + 1.5 * p.edd**2)
    assert abs(p.gamma - g) < 1e-14 * g
