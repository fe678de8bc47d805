"""Python binding of libgsk.so, the MI355X phase-field step library (C ABI in include/gsk.h).

The library is the only compute path: there is no CPU fallback. Importing this
module loads ``libgsk.so`` from the package directory and raises if it is absent.

Mirrors the reference plugin surface for the phase-field step
(General-Simulator ``src/dynamics.h:9-25``): a simulation handle is created with the
parameters ``Dynamics_mart::Initialize`` / ``Dynamics_precipitate::Initialize`` read
(``src/dynamics_mart.cu:14-173``, ``src/dynamics_precipitate.cu:17-110``), the
Khachaturyan tensor is built once (``src/gtensorb.cu:55-108``), and ``step`` advances
the fields like ``Calculate`` inside ``GS::Run`` (``src/gs.cu:201-209``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GSK_LIBRARY names another build of the same library (for example one compiled with other tile shapes)
LIB_PATH = os.environ.get("GSK_LIBRARY") or os.path.join(os.path.dirname(_HERE), "libgsk.so")

GSK_OK = 0
PHYS_MART, PHYS_PREC = 0, 1
PATH_FUSED32, PATH_PLAIN64 = 0, 1
RULE_SPECTRAL, RULE_REFERENCE = 0, 1
K_EXACT, K_REFERENCE = 0, 1
F32, F64 = 0, 1
RED_ENERGY, RED_MASS, RED_MEAN_ETA, RED_SUM_ETA2, RED_MART_FORCES = 1, 2, 3, 4, 5


class GskError(RuntimeError):
    """Raised when a gsk_* call returns a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"gsk error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("nvar", ctypes.c_int), ("physics", ctypes.c_int), ("path", ctypes.c_int),
        ("kconv", ctypes.c_int), ("device", ctypes.c_int), ("dx", ctypes.c_double), ("seed", ctypes.c_ulonglong),
        ("L", ctypes.c_double), ("kappa", ctypes.c_double), ("dt", ctypes.c_double), ("noise", ctypes.c_double),
        ("A", ctypes.c_double), ("B", ctypes.c_double), ("C", ctypes.c_double), ("w_el", ctypes.c_double),
        ("a", ctypes.c_double * 8), ("c1", ctypes.c_double), ("c2", ctypes.c_double), ("M", ctypes.c_double),
        ("kappa_c", ctypes.c_double), ("xi", ctypes.c_double), ("S", ctypes.c_double),
        ("rule", ctypes.c_int), ("chem_form", ctypes.c_int), ("w_grad", ctypes.c_double), ("w_chem", ctypes.c_double),
        ("w_dis", ctypes.c_double), ("temperature", ctypes.c_double), ("T0", ctypes.c_double),
        ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
    ]


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Load libgsk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libgsk.so not found at {LIB_PATH}; build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong
    L.gsk_last_error.restype = ctypes.c_char_p
    L.gsk_version.restype = ctypes.c_char_p
    L.gsk_config_default.argtypes = [ctypes.POINTER(Config), I]
    L.gsk_config_reference.argtypes = [ctypes.POINTER(Config), I]
    L.gsk_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    L.gsk_destroy.argtypes = [P]
    L.gsk_build_btensor.argtypes = [P, P, P, I]
    L.gsk_set_field.argtypes = [P, ctypes.c_char_p, P, I]
    L.gsk_get_field.argtypes = [P, ctypes.c_char_p, P, I]
    L.gsk_init_uniform.argtypes = [P, ctypes.c_char_p, D, D, U]
    L.gsk_step.argtypes = [P, I]
    L.gsk_set_step.argtypes = [P, ctypes.c_longlong]
    L.gsk_get_step.argtypes = [P]
    L.gsk_get_step.restype = ctypes.c_longlong
    L.gsk_set_param.argtypes = [P, ctypes.c_char_p, D]
    L.gsk_reduce.argtypes = [P, I, P]
    L.gsk_time_steps.argtypes = [P, I, P]
    L.gsk_kernel_times.argtypes = [P, I, P, I, P, ctypes.c_char_p, I]
    L.gsk_kernel_bytes.argtypes = [P, P, I, P]
    L.gsk_noise_field.argtypes = [U, ctypes.c_uint, ctypes.c_uint, ctypes.c_longlong, P, I]
    L.gsk_btensor_full.argtypes = [I, D, I, P, P, I, P, I, P, I]
    L.gsk_fft_r2c.argtypes = [I, I, P, P, I]
    L.gsk_fft_c2r.argtypes = [I, I, P, P, I]
    L.gsk_fft_compare.argtypes = [I, I, I, P, I]
    L.gsk_philox4x32.argtypes = [P, U, P]
    L.gsk_ensemble_create.argtypes = [P, I, ctypes.POINTER(P)]
    L.gsk_ensemble_destroy.argtypes = [P]
    L.gsk_ensemble_size.argtypes = [P]
    L.gsk_ensemble_member.argtypes = [P, I, ctypes.POINTER(P)]
    L.gsk_ensemble_build_btensor.argtypes = [P, P, P, I]
    L.gsk_ensemble_set_schedule.argtypes = [P, I, I, I]
    L.gsk_ensemble_get_schedule.argtypes = [P, P, P, P]
    L.gsk_ensemble_step.argtypes = [P, I]
    L.gsk_ensemble_autotune.argtypes = [P, I, P, I, P]
    L.gsk_adjoint_forward.argtypes = [P, I, I]
    L.gsk_adjoint_backward.argtypes = [P, P, P, P]
    L.gsk_adjoint_info.argtypes = [P, P]
    L.gsk_adjoint_release.argtypes = [P]
    L.gsk_stress.argtypes = [I, D, I, P, P, I, D, I, P, P, P, I]
    L.gsk_reduce_eta_noalloc.argtypes = [P, I, ctypes.c_longlong, I, P, I]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != GSK_OK:
        raise GskError(rc, lib().gsk_last_error().decode())


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


# ----------------------------------------------------------------------------- helpers
def cubic_modulus(c11: float, c12: float, c44: float) -> np.ndarray:
    """C[i][j][k][l] of a cubic crystal (the reference's cijkl, src/dynamics_mart.cu:80-89)."""
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    if i == j and k == l:
                        C[i, j, k, l] = c11 if i == k else c12
                    elif (i == k and j == l) or (i == l and j == k):
                        C[i, j, k, l] = c44
    return C


def mart_eigenstrains(nvar: int = 3, base: float = -0.04, axis: float = 0.12) -> np.ndarray:
    """Variant p: diag(base, base, base) + axis * e_p e_p (cubic -> tetragonal)."""
    e = np.zeros((nvar, 3, 3))
    for p in range(nvar):
        e[p] = np.eye(3) * base
        e[p, p % 3, p % 3] += axis
    return e


def default_config(physics: int = PHYS_MART, **kw) -> Config:
    c = Config()
    _check(lib().gsk_config_default(ctypes.byref(c), physics))
    for k, v in kw.items():
        if k == "a":
            for i, x in enumerate(v):
                c.a[i] = x
        else:
            setattr(c, k, v)
    return c


def reference_config(physics: int = PHYS_MART, **kw) -> Config:
    """Config of the plugins' own update rule with the reference's Initialize defaults."""
    c = Config()
    _check(lib().gsk_config_reference(ctypes.byref(c), physics))
    for k, v in kw.items():
        if k == "a":
            for i, x in enumerate(v):
                c.a[i] = x
        else:
            setattr(c, k, v)
    return c


class Sim:
    """One phase-field simulation on one GPU (a gsk_sim handle)."""

    def __init__(self, cfg: Config, _borrowed: Optional[ctypes.c_void_p] = None, _owner=None):
        self.cfg = cfg
        self._own = _borrowed is None
        self._owner = _owner  # keeps an ensemble alive while its member is in use
        if _borrowed is None:
            h = ctypes.c_void_p()
            _check(lib().gsk_create(ctypes.byref(cfg), ctypes.byref(h)))
        else:
            h = _borrowed
        self._h = h
        self.n, self.nvar, self.physics = cfg.n, cfg.nvar, cfg.physics
        # grid edges (nx, ny, nz): the box grid's, or n, n, n
        self.dims = (cfg.n, cfg.ny or cfg.n, cfg.nz or cfg.n)

    def close(self) -> None:
        if self._h and self._own:
            lib().gsk_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup
    def build_btensor(self, C: np.ndarray, eps: np.ndarray) -> None:
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1, 9)
        _check(lib().gsk_build_btensor(self._h, _ptr(C), _ptr(eps), eps.shape[0]))

    def _shape(self, name: str):
        d = self.dims
        if name == "eta":
            return (self.nvar,) + d
        if name == "dislocationstress":
            return (6,) + d
        return d

    def set_field(self, name: str, a: np.ndarray) -> None:
        a = np.asarray(a)
        dt = F64 if a.dtype == np.float64 else F32
        a = np.ascontiguousarray(a, dtype=np.float64 if dt == F64 else np.float32)
        if a.size != int(np.prod(self._shape(name))):
            raise ValueError(f"{name}: expected {self._shape(name)}, got {a.shape}")
        _check(lib().gsk_set_field(self._h, name.encode(), _ptr(a), dt))

    def get_field(self, name: str, dtype=np.float32) -> np.ndarray:
        out = np.empty(self._shape(name), dtype=dtype)
        _check(lib().gsk_get_field(self._h, name.encode(), _ptr(out), F64 if dtype == np.float64 else F32))
        return out

    def init_uniform(self, name: str, lo: float, hi: float, seed: int) -> None:
        _check(lib().gsk_init_uniform(self._h, name.encode(), lo, hi, seed))

    def set_param(self, key: str, value: float) -> None:
        _check(lib().gsk_set_param(self._h, key.encode(), value))

    # -- run
    def step(self, nsteps: int = 1) -> None:
        _check(lib().gsk_step(self._h, nsteps))

    @property
    def step_index(self) -> int:
        return int(lib().gsk_get_step(self._h))

    def set_step(self, s: int) -> None:
        _check(lib().gsk_set_step(self._h, s))

    def reduce(self, what: int) -> np.ndarray:
        out = np.zeros(16)
        _check(lib().gsk_reduce(self._h, what, _ptr(out)))
        return out

    def energy(self) -> np.ndarray:
        """[chemical, gradient, elastic, total] in FP64."""
        return self.reduce(RED_ENERGY)[:4].copy()

    def mass(self) -> float:
        """Sum of c: FP64 mean times n^3 plus the k = 0 mode of the FP32 fluctuation state."""
        return float(self.reduce(RED_MASS)[0])

    def mass_real(self) -> float:
        """FP64 sum of the stored FP32 concentration field."""
        return float(self.reduce(RED_MASS)[1])

    def mart_forces(self) -> np.ndarray:
        """Martensite reference rule: means of the gradient, chemical, elastic and dislocation force
        terms of the last step (the reference plugin's Get columns)."""
        return self.reduce(RED_MART_FORCES)[:4].copy()

    def mean_eta(self) -> np.ndarray:
        return self.reduce(RED_MEAN_ETA)[: self.nvar].copy()

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _check(lib().gsk_time_steps(self._h, nsteps, ctypes.byref(ms)))
        return ms.value

    def kernel_times(self, nsteps: int):
        ms = np.zeros(128)
        nk = ctypes.c_int()
        names = ctypes.create_string_buffer(8192)
        _check(lib().gsk_kernel_times(self._h, nsteps, _ptr(ms), 128, ctypes.byref(nk), names, 8192))
        return names.value.decode().split(","), ms[: nk.value].copy()

    def kernel_bytes(self) -> np.ndarray:
        b = np.zeros(128)
        nk = ctypes.c_int()
        _check(lib().gsk_kernel_bytes(self._h, _ptr(b), 128, ctypes.byref(nk)))
        return b[: nk.value].copy()


# ----------------------------------------------------------------------------- free functions
def _adjoint_forward(self, nsteps: int, K: int = 20) -> None:
    """Advance nsteps like step(), recording a checkpoint every K steps for adjoint_backward."""
    _check(lib().gsk_adjoint_forward(self._h, nsteps, K))


def _adjoint_backward(self, dL_detaT: np.ndarray):
    """Reverse sweep over the recorded run: returns (dLoss/deta0, [dLoss/ds_1..s_V, dLoss/dL, dLoss/dkappa])."""
    g = np.ascontiguousarray(dL_detaT, dtype=np.float32)
    if g.size != int(np.prod(self._shape("eta"))):
        raise ValueError("dL_detaT must have the shape of eta")
    ge = np.empty(self._shape("eta"), dtype=np.float32)
    gp = np.zeros(self.nvar + 2)
    _check(lib().gsk_adjoint_backward(self._h, _ptr(g), _ptr(ge), _ptr(gp)))
    return ge, gp


def _adjoint_info(self) -> dict:
    out = np.zeros(6)
    _check(lib().gsk_adjoint_info(self._h, _ptr(out)))
    return {"checkpoint_bytes": out[0], "recompute_bytes": out[1], "work_bytes": out[2], "forward_ms": out[3],
            "backward_ms": out[4], "stored_states": int(out[5])}


def _adjoint_release(self) -> None:
    _check(lib().gsk_adjoint_release(self._h))


Sim.adjoint_forward = _adjoint_forward
Sim.adjoint_backward = _adjoint_backward
Sim.adjoint_info = _adjoint_info
Sim.adjoint_release = _adjoint_release


def noise_field(seed: int, step: int, field: int, nvox: int, device: int = 0) -> np.ndarray:
    out = np.empty(nvox, dtype=np.float32)
    _check(lib().gsk_noise_field(seed, step, field, nvox, _ptr(out), device))
    return out


def btensor_full(n: int, dx: float, kconv: int, C: np.ndarray, eps1: np.ndarray, eps2: Optional[np.ndarray] = None,
                 device: int = 0) -> np.ndarray:
    C = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
    e1 = np.ascontiguousarray(eps1, dtype=np.float64).reshape(-1, 9)
    e2 = e1 if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64).reshape(-1, 9)
    out = np.empty((e1.shape[0], e2.shape[0], n, n, n))
    _check(lib().gsk_btensor_full(n, dx, kconv, _ptr(C), _ptr(e1), e1.shape[0], _ptr(e2), e2.shape[0], _ptr(out), device))
    return out


def fft_r2c(x: np.ndarray, device: int = 0) -> np.ndarray:
    """Unnormalised 3D R2C of real fields x[count, n, n, n] -> complex64 [count, n, n, n//2+1]."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    count, n = x.shape[0], x.shape[1]
    out = np.empty((count, n, n, n // 2 + 1), dtype=np.complex64)
    _check(lib().gsk_fft_r2c(n, count, _ptr(x), _ptr(out), device))
    return out


def fft_c2r(X: np.ndarray, device: int = 0) -> np.ndarray:
    """Unnormalised 3D C2R of half spectra X[count, n, n, n//2+1] -> float32 [count, n, n, n]."""
    X = np.ascontiguousarray(X, dtype=np.complex64)
    count, n = X.shape[0], X.shape[1]
    out = np.empty((count, n, n, n), dtype=np.float32)
    _check(lib().gsk_fft_c2r(n, count, _ptr(X), _ptr(out), device))
    return out


def fft_compare(n: int, count: int = 1, reps: int = 10, device: int = 0) -> dict:
    """Time the library's 3D R2C/C2R against rocFFT on the same data (ms per call of `count` fields)."""
    out = np.zeros(8)
    _check(lib().gsk_fft_compare(n, count, reps, _ptr(out), device))
    return {"hand_r2c_ms": out[0], "rocfft_r2c_ms": out[1], "hand_c2r_ms": out[2], "rocfft_c2r_ms": out[3],
            "max_rel_diff": out[4]}


SCHED_SEQUENTIAL, SCHED_STREAMS, SCHED_GRAPH, SCHED_BATCHED, SCHED_BATCHED_GRAPH = 0, 1, 2, 3, 4


class Ensemble:
    """Many independent simulations on one GPU stepped together (gsk_ensemble).

    ``members[i]`` are ordinary :class:`Sim` objects borrowed from the ensemble.
    """

    def __init__(self, cfgs):
        arr = (Config * len(cfgs))(*cfgs)
        h = ctypes.c_void_p()
        _check(lib().gsk_ensemble_create(arr, len(cfgs), ctypes.byref(h)))
        self._h = h
        self.members = []
        for i, c in enumerate(cfgs):
            m = ctypes.c_void_p()
            _check(lib().gsk_ensemble_member(h, i, ctypes.byref(m)))
            self.members.append(Sim(c, _borrowed=m, _owner=self))

    def close(self) -> None:
        if self._h:
            for m in self.members:
                m._h = ctypes.c_void_p()
            lib().gsk_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self) -> int:
        return len(self.members)

    def build_btensor(self, C: np.ndarray, eps: np.ndarray) -> None:
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1, 9)
        _check(lib().gsk_ensemble_build_btensor(self._h, _ptr(C), _ptr(eps), eps.shape[0]))

    def set_schedule(self, mode: int, branches: int = 0, steps_per_graph: int = 10) -> None:
        _check(lib().gsk_ensemble_set_schedule(self._h, mode, branches, steps_per_graph))

    def schedule(self):
        m, b, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().gsk_ensemble_get_schedule(self._h, ctypes.byref(m), ctypes.byref(b), ctypes.byref(k)))
        return m.value, b.value, k.value

    def step(self, nsteps: int = 1) -> None:
        _check(lib().gsk_ensemble_step(self._h, nsteps))

    def autotune(self, trial_steps: int = 5):
        """Time every schedule candidate, keep the fastest; returns rows (mode, branches, steps_per_graph, ms)."""
        tab = np.zeros((64, 4))
        rows = ctypes.c_int()
        _check(lib().gsk_ensemble_autotune(self._h, trial_steps, _ptr(tab), 64, ctypes.byref(rows)))
        return [tuple(r) for r in tab[: rows.value]]


def philox4x32(ctr, seed: int) -> np.ndarray:
    """One Philox4x32-10 block (the library's generator) for counter ctr[4] and a 64-bit seed."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _check(lib().gsk_philox4x32(_ptr(c), ctypes.c_ulonglong(seed), _ptr(out)))
    return out


def philox_uniform(seed: int, i: int, j: int) -> float:
    """u in [0, 1) from word 0 of the block at counter (i, j, 0, 2): (x >> 8) * 2^-24."""
    return float(philox4x32([i, j, 0, 2], seed)[0] >> 8) * 2.0**-24


def stress(eta: np.ndarray, C: np.ndarray, eps: np.ndarray, xi: float = 4000.0, source: int = 2, dx: float = 1.0,
           kconv: int = K_EXACT, device: int = 0):
    """The stress plugin on the GPU: returns (stress [6,n,n,n], defect [n,n,n]) as float32."""
    e = np.ascontiguousarray(eta, dtype=np.float32)
    V, n = e.shape[0], e.shape[1]
    Cc = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
    ep = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1, 9)
    out = np.empty((6, n, n, n), np.float32)
    d = np.empty((n, n, n), np.float32)
    _check(lib().gsk_stress(n, dx, kconv, _ptr(Cc), _ptr(ep), V, xi, source, _ptr(e), _ptr(out), _ptr(d), device))
    return out, d


def reduce_eta_noalloc(d_eta: int, nvar: int, nvox: int, what: int, device: int = 0) -> np.ndarray:
    """gsk_reduce_eta_noalloc on nvar fields of nvox floats at device address d_eta: the means of
    eta_v (RED_MEAN_ETA) or the sum of eta^2 (RED_SUM_ETA2), allocating no device memory."""
    out = np.zeros(16)
    _check(lib().gsk_reduce_eta_noalloc(ctypes.c_void_p(d_eta), nvar, nvox, what, _ptr(out), device))
    return out[: (nvar if what == RED_MEAN_ETA else 1)].copy()


def version() -> str:
    return lib().gsk_version().decode()
