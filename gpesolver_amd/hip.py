"""ctypes binding of libgpehip.so (the C ABI declared in include/gpehip.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, it never falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpehip.so")

GPE_FP32 = 32
GPE_FP64 = 64

# Every symbol include/gpehip.h declares.
EXPORTS = (
    "gpe_create", "gpe_destroy", "gpe_last_error", "gpe_version",
    "gpe_load_buffer", "gpe_load_pote", "gpe_export_psi", "gpe_export_fi3d", "gpe_import_fi3d",
    "gpe_set_coupling", "gpe_init_potential", "gpe_init_wavefunction", "gpe_export_pote", "gpe_calc_fi3d", "gpe_calc_norm", "gpe_normalize", "gpe_imag_iter_step",
    "gpe_real_potential_half_step", "gpe_real_kinetic_step", "gpe_calc_energy", "gpe_calc_energy_noalloc",
    "gpe_snapshot_begin", "gpe_snapshot_end",
    "gpe_imag_time_iters", "gpe_real_time_iters", "gpe_imag_time_until", "gpe_synchronize", "gpe_set_timing",
    "gpe_get_timing", "gpe_autotune", "gpe_set_tiles", "gpe_get_tiles",
    "gpe_batch_create", "gpe_batch_destroy", "gpe_batch_instance", "gpe_batch_last_error", "gpe_batch_it_until",
    "gpe_batch_set_fallback", "gpe_batch_fallback_taken", "gpe_batch_autotune",
    "gpe_fno_create", "gpe_fno_destroy", "gpe_fno_last_error", "gpe_fno_num_params", "gpe_fno_get_params",
    "gpe_fno_set_params", "gpe_fno_get_grads", "gpe_fno_forward", "gpe_fno_loss_grad", "gpe_fno_adam_step",
    "gpe_fno_set_data", "gpe_fno_loss_grad_idx", "gpe_fno_input_dev", "gpe_fno_output_dev", "gpe_fno_forward_dev",
    "gpe_warm_features", "gpe_load_amplitude",
    "gpe_gemm_test", "gpe_gemm_rows_test", "gpe_gemm_wgrad_test",
)
# Every symbol include/gpesolver_host.h declares.
HOST_EXPORTS = ("gpes_validate_params_json", "gpes_run", "gpes_write_files")


class GpeParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("m", C.c_double), ("n_atoms", C.c_double),
        ("cdd", C.c_double), ("ggp11", C.c_double), ("gamma", C.c_double), ("w_15", C.c_double),
        ("imag_time_dt", C.c_double), ("real_time_dt", C.c_double),
    ]


class GpeEnergies(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("e_kin", "e_pot", "e_int", "e_ext", "e_bmf", "e_total", "mu", "residual")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class GpeTrap(C.Structure):
    _fields_ = [("omega_x", C.c_double), ("omega_y", C.c_double), ("omega_z", C.c_double), ("dd", C.c_double)]


# potential keys and initial states (include/gpehip.h, GPE_POT_* / GPE_INIT_*)
POT_KEYS = {"HARMONIC": 0, "MEXICAN": 1, "MEXICAN_FREE": 2, "CYLINDRICAL": 3, "FREE": 4, "MEXICAN_ASYMETRIC": 5,
            "MEXICAN_ASYMETRIC_FREE": 6}
INIT_KEYS = {"COS": 0, "GAUSS": 1, "MULTIPLE_GAUSS": 2, "SETUP_GAUSS": 3}


class GpeError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load libgpehip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GpeError(f"{path} is missing: run `python -m gpesolver_amd.build` (or __graft_entry__.build())")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.gpe_create.argtypes = [C.POINTER(GpeParams), ip, C.POINTER(vp)]
    L.gpe_create.restype = ip
    L.gpe_destroy.argtypes = [vp]
    L.gpe_destroy.restype = None
    L.gpe_last_error.argtypes = [vp]
    L.gpe_last_error.restype = C.c_char_p
    L.gpe_version.restype = C.c_char_p
    L.gpe_warm_features.argtypes = [vp, vp, ip, C.c_double, C.POINTER(C.c_float)]
    L.gpe_warm_features.restype = ip
    L.gpe_load_amplitude.argtypes = [vp, vp, ip, dp]
    L.gpe_load_amplitude.restype = ip
    L.gpe_init_potential.argtypes = [vp, ip, C.POINTER(GpeTrap)]
    L.gpe_init_potential.restype = ip
    L.gpe_init_wavefunction.argtypes = [vp, ip, C.POINTER(GpeTrap), ip, ip, ip, ip, C.c_double, C.c_double,
                                        C.c_ulonglong]
    L.gpe_init_wavefunction.restype = ip
    for name in ("gpe_load_buffer", "gpe_load_pote", "gpe_export_psi", "gpe_export_fi3d", "gpe_import_fi3d",
                 "gpe_export_pote"):
        getattr(L, name).argtypes = [vp, dp]
        getattr(L, name).restype = ip
    L.gpe_set_coupling.argtypes = [vp, C.c_double, C.c_double]
    L.gpe_set_coupling.restype = ip
    for name in ("gpe_calc_fi3d", "gpe_normalize", "gpe_imag_iter_step", "gpe_real_potential_half_step",
                 "gpe_real_kinetic_step", "gpe_synchronize"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = ip
    L.gpe_calc_norm.argtypes = [vp, dp]
    L.gpe_calc_norm.restype = ip
    L.gpe_calc_energy.argtypes = [vp, C.POINTER(GpeEnergies)]
    L.gpe_calc_energy.restype = ip
    L.gpe_calc_energy_noalloc.argtypes = [vp, vp]
    L.gpe_calc_energy_noalloc.restype = ip
    L.gpe_snapshot_begin.argtypes = [vp]
    L.gpe_snapshot_begin.restype = ip
    L.gpe_snapshot_end.argtypes = [vp, dp, C.POINTER(GpeEnergies)]
    L.gpe_snapshot_end.restype = ip
    L.gpe_imag_time_iters.argtypes = [vp, C.c_long]
    L.gpe_imag_time_iters.restype = ip
    L.gpe_imag_time_until.argtypes = [vp, C.c_long, C.c_long, C.c_double, C.POINTER(C.c_long), dp]
    L.gpe_imag_time_until.restype = ip
    L.gpe_real_time_iters.argtypes = [vp, C.c_long, dp, dp]
    L.gpe_real_time_iters.restype = ip
    L.gpe_autotune.argtypes = [vp, C.c_long, C.POINTER(C.c_int), dp, dp]
    L.gpe_autotune.restype = ip
    L.gpe_set_tiles.argtypes = [vp, ip, ip, ip, ip, ip, ip]
    L.gpe_set_tiles.restype = ip
    L.gpe_get_tiles.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.gpe_get_tiles.restype = ip
    L.gpe_batch_autotune.argtypes = [vp, C.c_long, C.POINTER(C.c_int)]
    L.gpe_batch_autotune.restype = ip
    L.gpe_set_timing.argtypes = [vp, ip]
    L.gpe_set_timing.restype = ip
    L.gpe_get_timing.argtypes = [vp, dp, dp, dp]
    L.gpe_get_timing.restype = ip
    L.gpe_batch_create.argtypes = [C.POINTER(GpeParams), ip, ip, C.POINTER(vp)]
    L.gpe_batch_create.restype = ip
    L.gpe_batch_destroy.argtypes = [vp]
    L.gpe_batch_destroy.restype = None
    L.gpe_batch_instance.argtypes = [vp, ip]
    L.gpe_batch_instance.restype = vp
    L.gpe_batch_last_error.argtypes = [vp]
    L.gpe_batch_last_error.restype = C.c_char_p
    L.gpe_batch_it_until.argtypes = [vp, C.c_long, C.c_long, C.c_double, C.POINTER(C.c_long), dp]
    L.gpe_batch_it_until.restype = ip
    L.gpe_batch_set_fallback.argtypes = [vp, ip, dp]
    L.gpe_batch_set_fallback.restype = ip
    L.gpe_batch_fallback_taken.argtypes = [vp, ip]
    L.gpe_batch_fallback_taken.restype = ip
    L.gpes_validate_params_json.argtypes = [C.c_char_p, C.c_char_p, ip]
    L.gpes_validate_params_json.restype = ip
    L.gpes_run.argtypes = [C.c_char_p, C.c_double, dp, C.POINTER(C.c_long), C.c_char_p, ip]
    L.gpes_run.restype = ip
    L.gpes_write_files.argtypes = [C.c_char_p, C.c_char_p, dp, C.c_char_p, dp, dp, ip, C.c_char_p, ip]
    L.gpes_write_files.restype = ip
    _lib = L
    return L


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Device:
    """One solver instance on the GPU (a gpe_handle)."""

    def __init__(self, params: GpeParams, precision: int = GPE_FP64, _borrowed=None):
        self.L = load()
        self.params = params
        self.precision = precision
        self.shape = (params.nx, params.ny, params.nz)
        self._owned = _borrowed is None
        if _borrowed is not None:
            self.h = C.c_void_p(_borrowed)
            return
        h = C.c_void_p()
        rc = self.L.gpe_create(C.byref(params), precision, C.byref(h))
        if rc != 0:
            raise GpeError(self.L.gpe_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self.L.gpe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise GpeError(self.L.gpe_last_error(self.h).decode())

    # data movement
    def load_buffer(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        assert a.shape == self.shape
        self._check(self.L.gpe_load_buffer(self.h, _dp(a.view(np.float64))))

    def load_pote(self, V: np.ndarray):
        a = np.ascontiguousarray(V, dtype=np.float64)
        assert a.shape == self.shape
        self._check(self.L.gpe_load_pote(self.h, _dp(a)))

    def export_psi(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.complex128)
        self._check(self.L.gpe_export_psi(self.h, _dp(out.view(np.float64))))
        return out

    def export_fi3d(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.gpe_export_fi3d(self.h, _dp(out)))
        return out

    def import_fi3d(self, phi: np.ndarray):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        self._check(self.L.gpe_import_fi3d(self.h, _dp(a)))

    def set_coupling(self, ggp11: float, gamma: float):
        self._check(self.L.gpe_set_coupling(self.h, ggp11, gamma))

    # hooks
    def calc_fi3d(self):
        self._check(self.L.gpe_calc_fi3d(self.h))

    def calc_norm(self) -> float:
        x = C.c_double()
        self._check(self.L.gpe_calc_norm(self.h, C.byref(x)))
        return x.value

    def normalize(self):
        self._check(self.L.gpe_normalize(self.h))

    def imag_iter_step(self):
        self._check(self.L.gpe_imag_iter_step(self.h))

    def potential_half_step(self):
        self._check(self.L.gpe_real_potential_half_step(self.h))

    def kinetic_step(self):
        self._check(self.L.gpe_real_kinetic_step(self.h))

    # on-device initial state and potential (gpe_init_potential / gpe_init_wavefunction)
    @staticmethod
    def trap_of(p) -> GpeTrap:
        return GpeTrap(p.omega_x, p.omega_y, p.omega_z, p.dd)

    def init_potential(self, key: str, p):
        """V of potential `key` for the parameters p (PhysicalParameters), on the device."""
        self._check(self.L.gpe_init_potential(self.h, POT_KEYS[key], C.byref(self.trap_of(p))))

    def init_wavefunction(self, strategy: str, p, n_gauss_max: int = 2, drops=(3, 3, 2), sigma_cells: float = 0.0,
                          phase_noise: float = 0.0, seed: int = 0):
        """psi_0 of `strategy` on the device, then load_buffer's norm and normalize."""
        self._check(self.L.gpe_init_wavefunction(self.h, INIT_KEYS[strategy], C.byref(self.trap_of(p)), n_gauss_max,
                                                 drops[0], drops[1], drops[2], sigma_cells, phase_noise, seed))

    def warm_features(self, x_dev: int, cin: int, v_scale: float, consts):
        """Network input of this instance, [P][cin] at device address x_dev
        (gpe_warm_features)."""
        c = (C.c_float * 8)(*[float(v) for v in consts])
        self._check(self.L.gpe_warm_features(self.h, C.c_void_p(x_dev), cin, v_scale, c))

    def load_amplitude(self, amp_dev: int, mirror: bool) -> float:
        """psi0 from the predicted amplitude at device address amp_dev, loaded
        and normalised like load_buffer (gpe_load_amplitude); returns the norm
        before normalising."""
        n = C.c_double()
        self._check(self.L.gpe_load_amplitude(self.h, C.c_void_p(amp_dev), 1 if mirror else 0, C.byref(n)))
        return n.value

    def export_pote(self) -> np.ndarray:
        V = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.gpe_export_pote(self.h, _dp(V)))
        return V

    def energy(self) -> dict:
        e = GpeEnergies()
        self._check(self.L.gpe_calc_energy(self.h, C.byref(e)))
        return e.as_dict()

    def snapshot_begin(self):
        """Enqueue a copy of the state and its energies (gpe_snapshot_begin)."""
        self._check(self.L.gpe_snapshot_begin(self.h))

    def snapshot_end(self):
        """(psi, energies) of the snapshot in flight (gpe_snapshot_end)."""
        psi = np.empty(self.shape, dtype=np.complex128)
        e = GpeEnergies()
        self._check(self.L.gpe_snapshot_end(self.h, _dp(psi.view(np.float64)), C.byref(e)))
        return psi, e.as_dict()

    def energy_noalloc(self, out_dev: int):
        """gpe_calc_energy_noalloc: the eight energies to device memory at
        out_dev (8 doubles), enqueued on the handle's stream."""
        self._check(self.L.gpe_calc_energy_noalloc(self.h, C.c_void_p(out_dev)))

    # loops
    def it_steps(self, n: int):
        self._check(self.L.gpe_imag_time_iters(self.h, n))

    def it_until(self, max_steps: int, check_every: int = 100, tol: float = 1e-6) -> tuple[int, float]:
        """Imaginary time until |dE/N| < tol |E/N| between checks (gpe_imag_time_until):
        (steps run, last E/N)."""
        st, e = C.c_long(), C.c_double()
        self._check(self.L.gpe_imag_time_until(self.h, max_steps, check_every, tol, C.byref(st), C.byref(e)))
        return st.value, e.value

    def rt_steps(self, n: int, ggp11=None, gamma=None):
        if ggp11 is None:
            self._check(self.L.gpe_real_time_iters(self.h, n, None, None))
        else:
            g = np.ascontiguousarray(ggp11, dtype=np.float64)
            gm = np.ascontiguousarray(gamma, dtype=np.float64)
            assert g.shape == (n,) and gm.shape == (n,)
            self._check(self.L.gpe_real_time_iters(self.h, n, _dp(g), _dp(gm)))

    def synchronize(self):
        self._check(self.L.gpe_synchronize(self.h))

    def autotune(self, steps: int = 200) -> dict:
        """Pick the fastest launch-shape variant for this grid (gpe_autotune)."""
        cfg = C.c_int()
        rt, it = C.c_double(), C.c_double()
        self._check(self.L.gpe_autotune(self.h, steps, C.byref(cfg), C.byref(rt), C.byref(it)))
        return {"split_plane_pass": bool(cfg.value & 1), "register_x_pass": bool(cfg.value & 2),
                "graph_replay": bool(cfg.value & 4), "rt_step_us": rt.value, "it_step_us": it.value,
                "tiles": self.tiles()["cur"]}

    TILE_KEYS = ("cp", "ck", "pen_c", "pen_r", "pen_yreg", "pen_zreg")

    def tiles(self) -> dict:
        """Launch tiles and their bounds (gpe_get_tiles)."""
        cur, bound = (C.c_int * 6)(), (C.c_int * 6)()
        self._check(self.L.gpe_get_tiles(self.h, cur, bound))
        return {"cur": dict(zip(self.TILE_KEYS, cur[:])), "bound": dict(zip(self.TILE_KEYS, bound[:]))}

    def set_tiles(self, cp: int = 0, ck: int = 0, pen_c: int = 0, pen_r: int = 0, pen_yreg: int = -1,
                  pen_zreg: int = -1):
        """Force launch tiles (gpe_set_tiles; 0, or a flag -1, keeps one)."""
        self._check(self.L.gpe_set_tiles(self.h, cp, ck, pen_c, pen_r, pen_yreg, pen_zreg))

    def set_timing(self, on: bool):
        self._check(self.L.gpe_set_timing(self.h, 1 if on else 0))

    def timing(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.gpe_get_timing(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"plane_us": a.value, "line_us": b.value, "total_ms": c.value}


def validate_params_json(text: str) -> str | None:
    """FileManager::load_all_v1 + check_params on a JSON text: None if valid,
    else the error message (the reference throws std::runtime_error)."""
    L = load()
    buf = C.create_string_buffer(512)
    rc = L.gpes_validate_params_json(text.encode(), buf, 512)
    return None if rc == 0 else buf.value.decode()


def run_directory(workdir: str, it_tol: float = 0.0) -> dict:
    """Run the host front end (gpe_params.json in workdir) on the GPU."""
    L = load()
    e = C.c_double()
    n = C.c_long()
    buf = C.create_string_buffer(512)
    rc = L.gpes_run(workdir.encode(), it_tol, C.byref(e), C.byref(n), buf, 512)
    if rc != 0:
        raise GpeError(buf.value.decode())
    return {"energy_per_atom": e.value, "it_steps": n.value}


class Batch:
    """Many instances stepped together (gpe_batch): grids may differ."""

    def __init__(self, params_list, precision: int = GPE_FP32):
        self.L = load()
        arr = (GpeParams * len(params_list))(*params_list)
        b = C.c_void_p()
        rc = self.L.gpe_batch_create(arr, len(params_list), precision, C.byref(b))
        if rc != 0:
            raise GpeError(self.L.gpe_last_error(None).decode())
        self.b = b
        self.instances = [Device(p, precision, _borrowed=self.L.gpe_batch_instance(b, i))
                          for i, p in enumerate(params_list)]

    def close(self):
        if getattr(self, "b", None):
            for d in self.instances:
                d.h = None
            self.L.gpe_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fallback(self, i: int, psi):
        """State instance i restarts from if its own turns non-finite or its
        residual grows (gpe_batch_set_fallback); None clears it."""
        if psi is None:
            rc = self.L.gpe_batch_set_fallback(self.b, i, None)
        else:
            a = np.ascontiguousarray(psi, dtype=np.complex128)
            assert a.shape == self.instances[i].shape
            rc = self.L.gpe_batch_set_fallback(self.b, i, _dp(a.view(np.float64)))
        if rc != 0:
            raise GpeError(self.L.gpe_batch_last_error(self.b).decode())

    def autotune(self, steps: int = 100) -> list:
        """Per instance: register or LDS x pass and its cp tile, chosen per x
        length by timing whole batched steps (gpe_batch_autotune)."""
        n = len(self.instances)
        cfg = (C.c_int * n)()
        if self.L.gpe_batch_autotune(self.b, steps, cfg) != 0:
            raise GpeError(self.L.gpe_batch_last_error(self.b).decode())
        return [{"register_x_pass": bool(c & 2), "cp": c >> 8} for c in cfg[:]]

    def fallbacks_taken(self) -> np.ndarray:
        return np.array([bool(self.L.gpe_batch_fallback_taken(self.b, i)) for i in range(len(self.instances))])

    def it_until(self, max_steps: int, check_every: int = 100, tol: float = 0.0):
        n = len(self.instances)
        steps = (C.c_long * n)()
        e = np.zeros(n)
        rc = self.L.gpe_batch_it_until(self.b, max_steps, check_every, tol, steps, _dp(e))
        if rc != 0:
            raise GpeError(self.L.gpe_batch_last_error(self.b).decode())
        return np.array(steps[:], dtype=np.int64), e


def write_files(workdir: str, name: str, psi: np.ndarray, pote_name: str, V: np.ndarray, energies: np.ndarray):
    """The reference's on-disk formats (gpes_write_files) for the grid of
    workdir/gpe_params.json; energies is [n, 5] (kin, pot, int, ext, bmf)."""
    L = load()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    V = np.ascontiguousarray(V, dtype=np.float64)
    en = np.ascontiguousarray(energies, dtype=np.float64).reshape(-1, 5)
    err = C.create_string_buffer(512)
    rc = L.gpes_write_files(workdir.encode(), name.encode(), _dp(psi.view(np.float64)), pote_name.encode(), _dp(V),
                            _dp(en), en.shape[0], err, 512)
    if rc != 0:
        raise GpeError(err.value.decode())
