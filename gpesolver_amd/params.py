"""Host-side parameters, units, potentials and initial states.

Mirrors the reference's configuration layer so that a caller can build the
solver's inputs without the reference:
  units          src/utils/units/units.cpp:3-41
  parameters     src/parameters/parameters.cpp:37-59, parameters.hpp:70-72
  grid coords    src/context/context.cpp:17-33
  potentials     src/parameters/potentials.hpp:76-157
  initial states src/initializer/initializer.cpp:46-219
Everything here is one-time host set-up (vectorised numpy); the hot path runs
in libgpehip.so.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import hip


# ---------------------------------------------------------------- units
def len_nm_to_au(nm: float) -> float:
    return nm * 18.897261260649092


def len_au_to_nm(au: float) -> float:
    return au * 0.05291772105


def mass_Da_to_au(da: float) -> float:
    return da / 0.000548579909


def mass_au_to_Da(au: float) -> float:
    return au * 0.000548579909


def freq_Hz_to_au(hz: float) -> float:
    return hz * (4.1356e-12 / 27211.6)


def freq_au_to_Hz(au: float) -> float:
    return au / (4.1356e-12 / 27211.6)


def ene_au_to_meV(au: float) -> float:
    return au * 27211.407953


POTENTIALS = ("HARMONIC", "MEXICAN", "MEXICAN_FREE", "CYLINDRICAL", "FREE", "MEXICAN_ASYMETRIC",
              "MEXICAN_ASYMETRIC_FREE")
INIT_STRATEGIES = ("COS", "GAUSS", "MULTIPLE_GAUSS", "SETUP_GAUSS", "BINARY_FILE", "TEXT_FILE")
CALC_STRATEGIES = ("IT", "RT", "FS")


@dataclasses.dataclass
class PhysicalParameters:
    """PhysicalParameters (src/parameters/parameters.hpp:10-123), atomic units."""

    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    m: float
    n_atoms: float
    omega_x: float
    omega_y: float
    omega_z: float
    dd: float = len_nm_to_au(1500.0)
    edd: float = 1.5
    const_edd: bool = True
    edd_start: float = 0.0
    edd_stop: float = 0.0
    iter_imag: int = 0
    iter_real: int = 0
    imag_time_dt: float = 1.25e11
    real_time_dt: float = 1.00e10
    pote_key: str = "HARMONIC"
    calc_strategy: str = "FS"
    init_strategy: str = "GAUSS"
    n_gauss_max: int = 1
    bec_droplets: tuple = (1, 1, 1)
    load_filename: str = ""
    fftw_n_threads: int = 4
    # derived (init_parameters)
    aa: float = 0.0
    b: float = 0.0
    add: float = 131.0
    cdd: float = 0.0
    ggp11: float = 0.0
    gamma: float = 0.0
    w_15: float = 0.0

    @classmethod
    def from_physical(cls, nx, ny, nz, dx_nm, dy_nm, dz_nm, m_da=163.929, n_atoms=2e4, edd=1.5,
                      omega_hz=(60.0, 60.0, 120.0), dd_nm=1500.0, edd_start=None, edd_stop=None, **kw):
        """Build from the user units of gpe_params.json (file_manager.cpp:436-548)."""
        p = cls(nx=nx, ny=ny, nz=nz, dx=len_nm_to_au(dx_nm), dy=len_nm_to_au(dy_nm), dz=len_nm_to_au(dz_nm),
                m=mass_Da_to_au(m_da), n_atoms=float(n_atoms), omega_x=freq_Hz_to_au(omega_hz[0]),
                omega_y=freq_Hz_to_au(omega_hz[1]), omega_z=freq_Hz_to_au(omega_hz[2]), dd=len_nm_to_au(dd_nm),
                **kw)
        if edd_start is not None:
            p.const_edd = False
            p.edd_start, p.edd_stop = float(edd_start), float(edd_stop)
            p.edd = float(edd if edd is not None else edd_start)
        else:
            p.edd = float(edd)
        p.init_parameters()
        return p

    def get_dxdydz(self) -> float:
        return self.dx * self.dy * self.dz

    def init_parameters(self):
        """src/parameters/parameters.cpp:37-51."""
        self.aa = self.m * self.omega_x**2 / 4.0 / self.dd**2
        self.b = 0.5 * self.m * self.omega_x**2
        self.add = 131.0
        self.cdd = 12.0 * math.pi * self.add / self.m
        a = self.add / self.edd
        self.ggp11 = 4.0 * math.pi * a / self.m
        self.gamma = 128.0 * math.sqrt(math.pi) * a**2.5 / 3.0 / self.m * (1.0 + 1.5 * self.edd**2)
        self.w_15 = self.n_atoms**1.5

    def coupling_at(self, iteration: int) -> tuple[float, float, float]:
        """update_edd (src/parameters/parameters.cpp:53-59) without mutating:
        (edd, ggp11, gamma) for RT step `iteration`."""
        edd = self.edd_start + (self.edd_stop - self.edd_start) * iteration / self.iter_real
        a = self.add / edd
        g = 4 * math.pi * a / self.m
        gm = 128.0 * math.sqrt(math.pi) * a**2.5 / 3.0 / self.m * (1.0 + 1.5 * edd**2)
        return edd, g, gm

    def update_edd(self, iteration: int):
        self.edd, self.ggp11, self.gamma = self.coupling_at(iteration)

    def to_gpe(self) -> hip.GpeParams:
        g = hip.GpeParams()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.dx, g.dy, g.dz = self.dx, self.dy, self.dz
        g.m, g.n_atoms = self.m, self.n_atoms
        g.cdd, g.ggp11, g.gamma, g.w_15 = self.cdd, self.ggp11, self.gamma, self.w_15
        g.imag_time_dt, g.real_time_dt = self.imag_time_dt, self.real_time_dt
        return g


# ---------------------------------------------------------------- grid
def coords(n: int, d: float) -> np.ndarray:
    """x_i = (i - floor(n/2)) d  (src/context/context.cpp:17-33)."""
    return (np.arange(n) - (n // 2)) * d


def potential(key: str, p: PhysicalParameters) -> np.ndarray:
    """V sampled on every grid point (initializer.cpp:229-248, potentials.hpp:76-157).
    The reference samples through a std::function<double(int, int, int)>
    (initializer.hpp:37), so the coordinates are truncated toward zero to
    whole atomic units before the potential sees them."""
    x = np.trunc(coords(p.nx, p.dx))[:, None, None]
    y = np.trunc(coords(p.ny, p.dy))[None, :, None]
    z = np.trunc(coords(p.nz, p.dz))[None, None, :]
    vy = 0.5 * p.m * y**2 * p.omega_y**2
    vz = 0.5 * p.m * z**2 * p.omega_z**2
    slope = (0.25 / (27211.4 * 11.6 * 1e9)) / p.dd
    if key == "HARMONIC":
        V = 0.5 * p.m * x**2 * p.omega_x**2 + vy + vz
    elif key == "MEXICAN":
        V = -p.b * x**2 + p.aa * x**4 + vy + vz
    elif key == "MEXICAN_FREE":
        V = p.aa * x**4 + vy + vz
    elif key == "CYLINDRICAL":
        r2 = x**2 + y**2
        V = 0.5 * p.m * r2 * ((p.omega_y + p.omega_x) / 2.0) ** 2 + vz
    elif key == "FREE":
        V = vz + 0 * x + 0 * y
    elif key == "MEXICAN_ASYMETRIC":
        V = -p.b * x**2 + p.aa * x**4 - slope * x + vy + vz
    elif key == "MEXICAN_ASYMETRIC_FREE":
        V = p.aa * x**4 - slope * x + vy + vz
    else:
        raise ValueError(f"Not know potential option: {key}")
    return np.ascontiguousarray(np.broadcast_to(V, (p.nx, p.ny, p.nz)), dtype=np.float64)


def _gauss_sum(p, centers, sig, use_z_center):
    x = coords(p.nx, p.dx)[:, None, None]
    y = coords(p.ny, p.dy)[None, :, None]
    z = coords(p.nz, p.dz)[None, None, :]
    val = np.zeros((p.nx, p.ny, p.nz))
    for (cx, cy, cz) in centers:
        zc = (z - cz) if use_z_center else z
        val = val + (np.exp(-0.5 * (x - cx) ** 2 / sig[0] ** 2) * np.exp(-0.5 * (y - cy) ** 2 / sig[1] ** 2)
                     * np.exp(-0.5 * zc**2 / sig[2] ** 2))
    return val


def wavefunction(strategy: str, p: PhysicalParameters, n_gauss_max: int | None = None, drops=None,
                 sigma_cells: float = 0.0) -> np.ndarray:
    """Initial state on the interior points, boundary zero (initializer.cpp:46-219).
    sigma_cells > 0 sets the GAUSS / MULTIPLE_GAUSS width to that many cells
    per axis (the reference uses L/10 and L/20)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    Lx, Ly, Lz = nx * p.dx, ny * p.dy, nz * p.dz
    if strategy == "COS":
        x = coords(nx, p.dx)[:, None, None] - p.dd
        y = coords(ny, p.dy)[None, :, None]
        z = coords(nz, p.dz)[None, None, :]
        rrr = (nx // 2) * p.dx
        val = np.cos(2 * np.pi * x / rrr) * np.cos(2 * np.pi * y / rrr) * np.cos(1 * np.pi * z / rrr)
    elif strategy == "GAUSS":
        sig = (Lx / 10.0, Ly / 10.0, Lz / 10.0)
        if sigma_cells > 0:
            sig = (sigma_cells * p.dx, sigma_cells * p.dy, sigma_cells * p.dz)
        val = _gauss_sum(p, [(0.0, 0.0, 0.0)], sig, False)
    elif strategy == "MULTIPLE_GAUSS":
        nmax = n_gauss_max if n_gauss_max is not None else p.n_gauss_max
        if nmax <= 0:
            raise ValueError("Number of Gaussian maxima must be positive.")
        sig = (Lx / 20.0, Ly / 20.0, Lz / 20.0)
        if sigma_cells > 0:
            sig = (sigma_cells * p.dx, sigma_cells * p.dy, sigma_cells * p.dz)
        centers = []
        for idx in range(nmax):
            cx = p.dd if idx % 2 else -p.dd
            if nmax % 2 == 1:
                cy = (idx - nmax // 2) * (ny * p.dy) / 2.0 / (nmax + 1.0)
            else:
                cy = coords(ny, p.dy)[0] + (idx // 2 + 1) * (ny * p.dy) / (nmax // 2 + 1)
            centers.append((cx, cy, 0.0))
        val = _gauss_sum(p, centers, sig, False)
    elif strategy == "SETUP_GAUSS":
        bx, by, bz = drops if drops is not None else p.bec_droplets
        centers = [(-Lx / 2 + Lx / (bx + 1) * (a + 1), -Ly / 2 + Ly / (by + 1) * (b + 1), -Lz / 2 + Lz / (bz + 1) * (c + 1))
                   for a in range(bx) for b in range(by) for c in range(bz)]
        val = _gauss_sum(p, centers, (Lx / 20.0, Ly / 20.0, Lz / 20.0), True)
    else:
        raise ValueError(f"unsupported init strategy {strategy}")
    psi = np.zeros((nx, ny, nz), dtype=np.complex128)
    psi[1:-1, 1:-1, 1:-1] = val[1:-1, 1:-1, 1:-1]
    return psi


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from SplitMix64 (public-domain generator)."""
    mask = (1 << 64) - 1
    out = np.empty(n, dtype=np.float64)
    state = np.uint64(seed & mask)
    # vectorised: state_k = seed + k * golden
    golden = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        s = state + golden * np.arange(1, n + 1, dtype=np.uint64)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out


def add_phase_noise(psi: np.ndarray, amplitude: float, seed: int) -> np.ndarray:
    """psi * exp(i theta), theta ~ U[-amplitude, amplitude] from SplitMix64(seed),
    one draw per grid point in layout order."""
    u = splitmix64(seed, psi.size).reshape(psi.shape)
    return psi * np.exp(1j * amplitude * (2.0 * u - 1.0))
