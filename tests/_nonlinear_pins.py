"""Checks that fix the oracle's nonlinear terms by what the mathematics
requires, with no retyped formula:

  * the IT operator is the functional derivative of the energy
    (cpu_solver.cpp:50-97 against cpu_solver.cpp:137-194);
  * the dipolar interaction averages to zero over the three orientations of
    the dipole axis, up to its k = 0 term (the -cdd/3 folded into the contact
    term, cpu_solver.cpp:88,117,180);
  * the mean-field terms depend on (N, psi) only through the density
    n = N |psi|^2, so (N s^2, psi / s) gives the same potential and energy
    (the N and N^1.5 factors, cpu_solver.cpp:88-90,117-118,185-189);
  * the RT potential phase is -dt/2 times the IT operator's potential
    (cpu_solver.cpp:99-131), and the whole RT step has the generator
    i d psi/dt = (spectral kinetic + interior potential) psi
    (solver.cpp:107-112, fftw_split_solver.cpp:44-86);
  * the RT step conserves the energy to O(dt^2).

Every function takes the ctypes handle of an oracle build, so the same checks
run on the oracle and on copies of it with planted mistakes
(tests/test_oracle_mutations.py).  Each raises AssertionError on failure.
Parameters are O(1) numbers (not atomic units) so that every term of the
operator carries weight in the comparison.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import pyoracle

INNER = (slice(1, -1),) * 3


def unit_params(L, shape, n_atoms=3.7, m=0.8, d=(0.9, 1.1, 1.3), g=0.9, cdd=0.6, gamma=0.35,
                dtau=0.01, dt=0.05):
    """O(1) parameters.  orc_init_parameters derives w_15 (and dV) from
    n_atoms; the couplings are then set to O(1) values."""
    p = pyoracle.OrcParams()
    p.nx, p.ny, p.nz = shape
    p.dx, p.dy, p.dz = d
    p.m = m
    p.n_atoms = n_atoms
    p.omega_x = p.omega_y = p.omega_z = 0.1
    p.dd = 1.0
    p.edd = 1.5
    p.const_edd = 1
    p.imag_time_dt = dtau
    p.real_time_dt = dt
    L.orc_init_parameters(C.byref(p))
    p.ggp11, p.cdd, p.gamma = g, cdd, gamma
    return p


def it_H(L, p, psi, V):
    """(psi - psi')/dtau of one explicit IT step (linear + nonlinear part,
    before the norm): H psi on the interior, 0 on the boundary."""
    o = pyoracle.Oracle(p, L)
    o.load_pote(V)
    o.psi[...] = psi
    o.psii[...] = psi
    o.calc_fi3d()
    o.imag_linear()
    o.imag_nonlinear()
    return (psi - o.psi) / p.imag_time_dt


def energy(L, p, psi, V):
    o = pyoracle.Oracle(p, L)
    o.load_pote(V)
    o.psi[...] = psi
    o.calc_fi3d()
    return o.energy()


def half_step(L, p, psi, V):
    o = pyoracle.Oracle(p, L)
    o.load_pote(V)
    o.psi[...] = psi
    o.calc_fi3d()
    o.potential_half_step()
    return o.psi.copy()


def rt_run(L, p, psi, V, nsteps, every=1):
    """RT steps from psi; returns the states after every `every` steps."""
    o = pyoracle.Oracle(p, L)
    o.load_pote(V)
    o.psi[...] = psi
    o.calc_fi3d()
    out = [psi.copy()]
    for t in range(nsteps):
        o.rt_step()
        if (t + 1) % every == 0:
            out.append(o.psi.copy())
    return out


def random_state(shape, seed, boundary_zero=True, amp=0.8):
    rng = np.random.default_rng(seed)
    psi = amp * (rng.uniform(0.3, 1.0, shape) * np.exp(1j * rng.uniform(-np.pi, np.pi, shape)))
    if boundary_zero:
        mask = np.zeros(shape, bool)
        mask[INNER] = True
        psi = np.where(mask, psi, 0)
    V = rng.uniform(0.0, 1.5, shape)
    return psi.astype(np.complex128), V


def spectral_kinetic(p, psi):
    """-(1/2m) Laplacian on the periodic grid, by numpy's FFT."""
    ks = [2 * np.pi * np.fft.fftfreq(n, h) for n, h in zip(psi.shape, (p.dx, p.dy, p.dz))]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    return np.fft.ifftn((KX**2 + KY**2 + KZ**2) / (2 * p.m) * np.fft.fftn(psi))


# --------------------------------------------------------------------- pins
def pin_it_step_is_energy_gradient(L, shape=(7, 6, 5), seed=1, h=3e-5, rtol=1e-7):
    """(psi - psi')/dtau = (1/(dV N)) dE/d(conj psi) at every interior point,
    with Phi recomputed from psi for each perturbed energy (central
    differences; dE/d(conj psi) = (dE/dRe + i dE/dIm)/2).  psi is zero on the
    boundary, where the energy sums (interior only) and Phi (whole grid) see
    the same points."""
    p = unit_params(L, shape)
    psi, V = random_state(shape, seed)
    H = it_H(L, p, psi, V)
    o = pyoracle.Oracle(p, L)
    o.load_pote(V)

    def E(x):
        o.psi[...] = x
        o.calc_fi3d()
        return o.energy()["e_total"]

    grad = np.zeros(shape, np.complex128)
    for idx in np.ndindex(*(n - 2 for n in shape)):
        pt = tuple(i + 1 for i in idx)
        d = []
        for step in (h, 1j * h):
            x = psi.copy()
            x[pt] += step
            ep = E(x)
            x[pt] -= 2 * step
            em = E(x)
            d.append((ep - em) / (2 * h))
        grad[pt] = 0.5 * (d[0] + 1j * d[1])
    grad /= p.dxdydz * p.n_atoms
    err = np.abs(grad[INNER] - H[INNER]).max() / np.abs(H[INNER]).max()
    assert err < rtol, f"IT step vs energy gradient: {err:.3e}"
    return err


def _permuted(perm, shape, d):
    return tuple(shape[a] for a in perm), tuple(d[a] for a in perm)


def pin_isotropic_dipole_average(L, shape=(7, 6, 5), seed=2, rtol=1e-11):
    """The dipole axis is the last array axis.  Solving the same problem with
    the arrays (and spacings) permuted so that the dipole lies along x, y and
    z in turn, the three dipolar terms add up to cdd (n - <n>_pad) - cdd n:
    kx^2/k^2 + ky^2/k^2 + kz^2/k^2 = 1 for k != 0, and the -1/3 of each
    orientation removes the local part.  So
        sum_axes H_axis psi - 3 H_nodip psi = -cdd <n>_pad psi,
    <n>_pad = sum N|psi|^2 / (8 nx ny nz), for the IT operator, the RT phase
    angle (-dt/2 times it) and the energy (-cdd/2 N <n>_pad dV sum |psi|^2)."""
    d = (0.9, 1.1, 1.3)
    psi, V = random_state(shape, seed, boundary_zero=False)
    perms = [(0, 1, 2), (2, 1, 0), (0, 2, 1)]
    Hs, phs, Es = [], [], []
    for perm in perms:
        shp, dd = _permuted(perm, shape, d)
        p = unit_params(L, shp, d=dd)
        x = np.ascontiguousarray(psi.transpose(perm))
        v = np.ascontiguousarray(V.transpose(perm))
        inv = np.argsort(perm)
        Hs.append(it_H(L, p, x, v).transpose(inv))
        phs.append(np.angle(half_step(L, p, x, v) / x).transpose(inv))
        Es.append(energy(L, p, x, v)["e_total"])
    p0 = unit_params(L, shape, d=d, cdd=0.0)
    H0 = it_H(L, p0, psi, V)
    ph0 = np.angle(half_step(L, p0, psi, V) / psi)
    E0 = energy(L, p0, psi, V)["e_total"]
    p = unit_params(L, shape, d=d)
    nbar = p.n_atoms * np.sum(np.abs(psi) ** 2) / (8 * psi.size)
    lhs = sum(Hs) - 3 * H0
    rhs = -p.cdd * nbar * psi
    err_it = np.abs(lhs[INNER] - rhs[INNER]).max() / np.abs(H0[INNER]).max()
    assert err_it < rtol, f"IT isotropic dipole average: {err_it:.3e}"
    lhs_ph = sum(phs) - 3 * ph0
    rhs_ph = 0.5 * p.real_time_dt * p.cdd * nbar
    err_ph = np.abs(lhs_ph[INNER] - rhs_ph).max() / np.abs(ph0[INNER]).max()
    assert err_ph < rtol, f"RT isotropic dipole average: {err_ph:.3e}"
    w = np.abs(psi[INNER]) ** 2
    rhs_E = -0.5 * p.cdd * p.n_atoms * nbar * p.dxdydz * w.sum()
    err_E = abs(sum(Es) - 3 * E0 - rhs_E) / abs(E0)
    assert err_E < rtol, f"energy isotropic dipole average: {err_E:.3e}"
    return max(err_it, err_ph, err_E)


def pin_density_scaling(L, shape=(7, 6, 5), seed=3, s=1.7, rtol=1e-12):
    """(N, psi) -> (N s^2, psi/s) keeps n = N|psi|^2, so the potential
    (IT: H psi -> H psi / s; RT: the same phase) and the energy are unchanged."""
    psi, V = random_state(shape, seed)
    p1 = unit_params(L, shape, n_atoms=3.7)
    p2 = unit_params(L, shape, n_atoms=3.7 * s * s)
    H1 = it_H(L, p1, psi, V)
    H2 = it_H(L, p2, psi / s, V)
    err_it = np.abs(H2 - H1 / s).max() / np.abs(H1).max()
    assert err_it < rtol, f"IT density scaling: {err_it:.3e}"
    ph1 = np.angle(half_step(L, p1, psi, V)[INNER] / psi[INNER])
    ph2 = np.angle(half_step(L, p2, psi / s, V)[INNER] / (psi[INNER] / s))
    err_ph = np.abs(ph2 - ph1).max() / np.abs(ph1).max()
    assert err_ph < rtol, f"RT density scaling: {err_ph:.3e}"
    E1 = energy(L, p1, psi, V)
    E2 = energy(L, p2, psi / s, V)
    err_E = max(abs(E2[k] - E1[k]) / abs(E1["e_total"]) for k in ("e_kin", "e_pot", "e_int", "e_ext", "e_bmf"))
    assert err_E < rtol, f"energy density scaling: {err_E:.3e}"
    return max(err_it, err_ph, err_E)


def pin_rt_phase_is_it_potential(L, shape=(7, 6, 5), seed=4, rtol=1e-12):
    """With the kinetic term switched off (m -> infinity) the IT operator is
    diagonal: H psi = Veff psi, Veff real.  The RT half step multiplies the
    interior by exp(-i dt/2 Veff) and leaves the boundary alone."""
    psi, V = random_state(shape, seed, boundary_zero=False)
    p_inf = unit_params(L, shape, m=1e30)
    Hd = it_H(L, p_inf, psi, V)
    veff = Hd[INNER] / psi[INNER]
    assert np.abs(veff.imag).max() < 1e-12 * np.abs(veff).max(), "IT potential part is not real"
    veff = veff.real
    p = unit_params(L, shape, dt=1.0 / np.abs(veff).max())
    got = half_step(L, p, psi, V)
    want = psi.copy()
    want[INNER] = psi[INNER] * np.exp(-0.5j * p.real_time_dt * veff)
    err = np.abs(got - want).max() / np.abs(psi).max()
    assert err < rtol, f"RT phase vs IT potential: {err:.3e}"
    return err


def pin_rt_generator(L, shape=(8, 6, 6), seed=5, rtol=2e-6):
    """G(dt) = i (psi(dt) - psi)/dt of one whole RT step tends to
    (K_spectral + Veff on the interior) psi; 2 G(dt/2) - G(dt) removes the
    O(dt) term.  Veff comes from the IT operator with the kinetic term off."""
    psi, V = random_state(shape, seed, boundary_zero=False)
    p_inf = unit_params(L, shape, m=1e30)
    veff_psi = it_H(L, p_inf, psi, V)
    p = unit_params(L, shape)
    want = spectral_kinetic(p, psi) + veff_psi
    scale = np.abs(want).max()
    dt = 1e-3 / scale
    G = []
    for h in (dt, dt / 2):
        q = unit_params(L, shape, dt=h)
        out = rt_run(L, q, psi, V, 1)[-1]
        G.append(1j * (out - psi) / h)
    R = 2 * G[1] - G[0]
    err = np.abs(R - want).max() / scale
    assert err < rtol, f"RT generator: {err:.3e}"
    return err


def pin_rt_energy_conservation(L, shape=(14, 12, 10), T=2.0):
    """E with the spectral kinetic energy (the RT step's) in place of the FD
    one is conserved by the step.  The state is a moving Gaussian in a trap,
    far from the boundary; the largest deviation over [0, T] is compared at
    dt and dt/2.
      * cdd = 0: the step is a symmetric (Strang) splitting with exact
        sub-flows, so the error is O(dt^2): halving dt divides it by 4.
      * cdd != 0: the first half step of every step uses the Phi of the
        previous step (solver.cpp:107-112 computes Phi once, after the first
        half step, and both half steps read it), which makes the splitting
        asymmetric in Phi: the error is O(dt), halving dt divides it by 2."""
    ratios = []
    for cdd, band, tol in ((0.0, (3.6, 4.4), 1e-4), (0.6, (1.8, 2.2), 2e-3)):
        p = unit_params(L, shape, n_atoms=3.7, cdd=cdd)
        grids = np.meshgrid(*[(np.arange(n) - (n - 1) / 2) * h for n, h in zip(shape, (p.dx, p.dy, p.dz))],
                            indexing="ij")
        r2 = sum((g / (1.3 * h)) ** 2 for g, h in zip(grids, (p.dx, p.dy, p.dz)))
        psi = (0.9 * np.exp(-0.5 * r2) * np.exp(0.4j * grids[0] / p.dx)).astype(np.complex128)
        V = 0.4 * sum(g * g for g in grids)

        def E_spec(x):
            e = energy(L, p, x, V)
            kin = p.dxdydz * p.n_atoms * np.real(np.sum(np.conj(x) * spectral_kinetic(p, x)))
            return e["e_total"] - e["e_kin"] + kin

        E0 = E_spec(psi)
        devs = []
        for n in (40, 80):
            q = unit_params(L, shape, n_atoms=3.7, cdd=cdd, dt=T / n)
            states = rt_run(L, q, psi, V, n, every=n // 40)
            devs.append(max(abs(E_spec(x) - E0) for x in states[1:]))
        ratio = devs[0] / devs[1]
        assert devs[1] < tol * abs(E0), f"RT energy drift (cdd={cdd}) {devs[1] / abs(E0):.3e}"
        assert band[0] < ratio < band[1], f"RT energy error order (cdd={cdd}): ratio {ratio:.3f}"
        ratios.append(ratio)
    return ratios


ALL_PINS = [
    pin_it_step_is_energy_gradient,
    pin_isotropic_dipole_average,
    pin_density_scaling,
    pin_rt_phase_is_it_potential,
    pin_rt_generator,
    pin_rt_energy_conservation,
]
