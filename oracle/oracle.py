"""ctypes binding of the FP64 CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the CPU-baseline leg of bench.py load this
module; the product (libgsk.so and its bindings) never does. See oracle.h for the
reference files each routine restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class MartParams(ctypes.Structure):
    _fields_ = [("A", ctypes.c_double), ("B", ctypes.c_double), ("C", ctypes.c_double),
                ("kappa", ctypes.c_double), ("L", ctypes.c_double), ("dt", ctypes.c_double),
                ("w_el", ctypes.c_double), ("noise", ctypes.c_double), ("dx", ctypes.c_double),
                ("kconv", ctypes.c_int), ("seed", ctypes.c_uint64)]


class MartRefParams(ctypes.Structure):
    _fields_ = [("a1", ctypes.c_double), ("a2", ctypes.c_double), ("a3", ctypes.c_double),
                ("w_grad", ctypes.c_double), ("w_chem", ctypes.c_double), ("w_el", ctypes.c_double),
                ("w_dis", ctypes.c_double), ("w_noise", ctypes.c_double), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("kconv", ctypes.c_int), ("chem_form", ctypes.c_int),
                ("seed", ctypes.c_uint64)]


class PrecParams(ctypes.Structure):
    _fields_ = [("A", ctypes.c_double * 8), ("c1", ctypes.c_double), ("c2", ctypes.c_double),
                ("M", ctypes.c_double), ("L", ctypes.c_double), ("kappa_c", ctypes.c_double),
                ("kappa_eta", ctypes.c_double), ("xi", ctypes.c_double), ("S", ctypes.c_double), ("dt", ctypes.c_double),
                ("noise", ctypes.c_double), ("dx", ctypes.c_double), ("kconv", ctypes.c_int),
                ("seed", ctypes.c_uint64)]


class Grid(ctypes.Structure):
    """A box grid nx x ny x nz with spacings dx, dy, dz (or_grid)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


def grid(nx: int, ny: int, nz: int, dx: float = 1.0, dy: float = 1.0, dz: float = 1.0) -> Grid:
    return Grid(nx, ny, nz, dx, dy, dz)


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, I, D, U64, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64
        L.or_set_threads.argtypes = [I]
        L.or_get_threads.restype = I
        L.or_fft1d.argtypes = [I, P, I]
        L.or_fft3d.argtypes = [I, P, I]
        L.or_philox4x32_10.argtypes = [P, P, P]
        L.or_normal4.argtypes = [P, P]
        L.or_log_f.argtypes = [ctypes.c_float]
        L.or_log_f.restype = ctypes.c_float
        L.or_sqrt_f.argtypes = [ctypes.c_float]
        L.or_sqrt_f.restype = ctypes.c_float
        L.or_sincos_f.argtypes = [ctypes.c_float, P, P]
        L.or_noise_field.argtypes = [U64, ctypes.c_uint32, ctypes.c_uint32, I64, P]
        L.or_uniform_field.argtypes = [U64, ctypes.c_uint32, I64, D, D, P]
        L.or_kcomp.argtypes = [I, I, D, I]
        L.or_kcomp.restype = D
        L.or_btensor.argtypes = [I, D, I, P, P, I, P, I, P]
        L.or_mart_run.argtypes = [I, I, ctypes.POINTER(MartParams), P, P, I, I]
        L.or_mart_energy.argtypes = [I, I, ctypes.POINTER(MartParams), P, P, P]
        L.or_mart_force.argtypes = [I, I, ctypes.POINTER(MartParams), P, P, P]
        L.or_mart_adjoint.argtypes = [I, I, ctypes.POINTER(MartParams), P, P, I, I, P, P, P]
        L.or_mart_ref_run.argtypes = [I, I, ctypes.POINTER(MartRefParams), P, P, P, P, I, I]
        L.or_stress.argtypes = [I, I, P, ctypes.c_double, I, P, P, P]
        L.or_prec_ref_run.argtypes = [I, I, ctypes.POINTER(PrecParams), P, P, P, I, I]
        L.or_prec_run.argtypes = [I, I, ctypes.POINTER(PrecParams), P, P, P, I, I]
        L.or_prec_energy.argtypes = [I, I, ctypes.POINTER(PrecParams), P, P, P, P]
        L.or_prec_local.argtypes = [I, ctypes.POINTER(PrecParams), D, P, P, P, P]
        G = ctypes.POINTER(Grid)
        L.or_fft3d_box.argtypes = [G, P, I]
        L.or_btensor_box.argtypes = [G, I, P, P, I, P, I, P]
        L.or_mart_run_box.argtypes = [G, I, ctypes.POINTER(MartParams), P, P, I, I]
        L.or_mart_energy_box.argtypes = [G, I, ctypes.POINTER(MartParams), P, P, P]
        L.or_mart_ref_run_box.argtypes = [G, I, ctypes.POINTER(MartRefParams), P, P, P, P, I, I]
        L.or_prec_run_box.argtypes = [G, I, ctypes.POINTER(PrecParams), P, P, P, I, I]
        L.or_prec_energy_box.argtypes = [G, I, ctypes.POINTER(PrecParams), P, P, P, P]
        L.or_prec_ref_run_box.argtypes = [G, I, ctypes.POINTER(PrecParams), P, P, P, I, I]
        L.or_sum.argtypes = [P, I64]
        L.or_sum.restype = D
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


def threads() -> int:
    return lib().or_get_threads()


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft1d(a.size, _p(a), sign)
    return a


def fft3d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft3d(a.shape[0], _p(a), sign)
    return a


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def normal4(x) -> np.ndarray:
    a = np.asarray(x, dtype=np.uint32)
    o = np.zeros(4, dtype=np.float32)
    lib().or_normal4(_p(a), _p(o))
    return o


def log_f(u: float) -> float:
    return lib().or_log_f(u)


def sqrt_f(x: float) -> float:
    return lib().or_sqrt_f(x)


def sincos_f(t: float):
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().or_sincos_f(t, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def noise_field(seed: int, step: int, field: int, nvox: int) -> np.ndarray:
    o = np.empty(nvox, dtype=np.float32)
    lib().or_noise_field(seed, step, field, nvox, _p(o))
    return o


def uniform_field(seed: int, field: int, nvox: int, lo: float, hi: float) -> np.ndarray:
    o = np.empty(nvox, dtype=np.float32)
    lib().or_uniform_field(seed, field, nvox, lo, hi, _p(o))
    return o


def kcomp(m: int, n: int, dx: float = 1.0, kconv: int = 0) -> float:
    return lib().or_kcomp(m, n, dx, kconv)


def btensor(n: int, C: np.ndarray, eps1: np.ndarray, eps2=None, dx: float = 1.0, kconv: int = 0) -> np.ndarray:
    C = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
    e1 = np.ascontiguousarray(eps1, dtype=np.float64).reshape(-1, 9)
    e2 = e1 if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64).reshape(-1, 9)
    B = np.empty((e1.shape[0], e2.shape[0], n, n, n))
    lib().or_btensor(n, dx, kconv, _p(C), _p(e1), e1.shape[0], _p(e2), e2.shape[0], _p(B))
    return B


def mart_params(**kw) -> MartParams:
    p = MartParams(A=0.2, B=3.0, C=3.0, kappa=1.0, L=1.0, dt=0.01, w_el=1.0, noise=0.0, dx=1.0, kconv=0, seed=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def prec_params(**kw) -> PrecParams:
    p = PrecParams()
    A = kw.pop("A", [0.0, 54.0, -17.0, 7.0, 2.5, 0.2, 0.2, 0.2])
    for i in range(8):
        p.A[i] = A[i]
    defaults = dict(c1=0.1, c2=0.44, M=0.4, L=0.4, kappa_c=300.0, kappa_eta=300.0, xi=400.0, S=0.0, dt=0.1, noise=0.0,
                    dx=1.0, kconv=0, seed=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def mart_run(eta: np.ndarray, B: np.ndarray, p: MartParams, nsteps: int, step0: int = 0) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    V, n = e.shape[0], e.shape[1]
    Bc = np.ascontiguousarray(B, dtype=np.float64)
    lib().or_mart_run(n, V, ctypes.byref(p), _p(Bc), _p(e), step0, nsteps)
    return e


def mart_energy(eta: np.ndarray, B: np.ndarray, p: MartParams) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64)
    out = np.zeros(4)
    lib().or_mart_energy(e.shape[1], e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B)), _p(e), _p(out))
    return out


def mart_force(eta: np.ndarray, B: np.ndarray, p: MartParams) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64)
    F = np.empty_like(e)
    lib().or_mart_force(e.shape[1], e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B)), _p(e), _p(F))
    return F


def mart_adjoint(eta0: np.ndarray, B: np.ndarray, p: MartParams, nsteps: int, gT: np.ndarray, step0: int = 0):
    """Store-all FP64 adjoint: returns (dLoss/deta0, [dLoss/ds_1..s_V, dLoss/dL, dLoss/dkappa])."""
    e = np.ascontiguousarray(eta0, dtype=np.float64)
    V, n = e.shape[0], e.shape[1]
    g = np.ascontiguousarray(gT, dtype=np.float64)
    ge = np.empty_like(e)
    gp = np.zeros(V + 2)
    lib().or_mart_adjoint(n, V, ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)), _p(e), step0, nsteps,
                          _p(g), _p(ge), _p(gp))
    return ge, gp


def mart_ref_params(**kw) -> MartRefParams:
    # reference defaults (src/dynamics_mart.cu:20-39 HEAD side, LPC[2..3] :41)
    p = MartRefParams(a1=0.0, a2=32.05, a3=37.5, w_grad=2.5, w_chem=1.0, w_el=1e5, w_dis=1.0, w_noise=1e-4, dt=0.01,
                      dx=1.0, kconv=1, chem_form=0, seed=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def mart_ref_run(eta, B, p: MartRefParams, nsteps: int, step0: int = 0, defect=None, K=None) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    V, n = e.shape[0], e.shape[1]
    d = None if defect is None else np.ascontiguousarray(defect, dtype=np.float64)
    k = None if K is None else np.ascontiguousarray(K, dtype=np.float64)
    lib().or_mart_ref_run(n, V, ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)), _p(e),
                          None if d is None else _p(d), None if k is None else _p(k), step0, nsteps)
    return e


def prec_ref_run(c, eta, B, p: PrecParams, nsteps: int, step0: int = 0):
    cc = np.ascontiguousarray(c, dtype=np.float64).copy()
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    V, n = e.shape[0], e.shape[1]
    lib().or_prec_ref_run(n, V, ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)), _p(cc), _p(e),
                          step0, nsteps)
    return cc, e


BASE_STRAINS = np.zeros((6, 3, 3))
for _b, (_i, _j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
    BASE_STRAINS[_b, _i, _j] = 1.0  # single unit entries (src/dynamics_stress.cu:70-75)


def stress(eta, C, eps, xi=4000.0, source=2, dx=1.0, kconv=0):
    """Stress plugin: returns (stress [6,n,n,n], defect [n,n,n])."""
    e = np.ascontiguousarray(eta, dtype=np.float64)
    V, n = e.shape[0], e.shape[1]
    B = np.ascontiguousarray(btensor(n, C, BASE_STRAINS, eps, dx=dx, kconv=kconv))
    out = np.empty((6, n, n, n))
    d = np.empty((n, n, n))
    lib().or_stress(n, V, _p(B), xi, source, _p(e), _p(out), _p(d))
    return out, d


def prec_run(c: np.ndarray, eta: np.ndarray, Bc: np.ndarray, p: PrecParams, nsteps: int, step0: int = 0):
    cc = np.ascontiguousarray(c, dtype=np.float64).copy()
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    V, n = e.shape[0], e.shape[1]
    lib().or_prec_run(n, V, ctypes.byref(p), _p(np.ascontiguousarray(Bc, dtype=np.float64)), _p(cc), _p(e), step0,
                      nsteps)
    return cc, e


def prec_energy(c: np.ndarray, eta: np.ndarray, Bc: np.ndarray, p: PrecParams) -> np.ndarray:
    cc = np.ascontiguousarray(c, dtype=np.float64)
    e = np.ascontiguousarray(eta, dtype=np.float64)
    out = np.zeros(4)
    lib().or_prec_energy(e.shape[1], e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(Bc, dtype=np.float64)),
                         _p(cc), _p(e), _p(out))
    return out


def prec_local(c: float, eta: np.ndarray, p: PrecParams):
    e = np.ascontiguousarray(eta, dtype=np.float64)
    mu = ctypes.c_double()
    F = np.zeros(e.size)
    f = ctypes.c_double()
    lib().or_prec_local(e.size, ctypes.byref(p), c, _p(e), ctypes.byref(mu), _p(F), ctypes.byref(f))
    return mu.value, F, f.value


# ----------------------------------------------------------------------------- box grids
# Fields on a box grid g have shape (..., g.nx, g.ny, g.nz); the params' dx is not used (the
# grid carries dx, dy, dz).
def _dims(g: Grid):
    return (g.nx, g.ny, g.nz)


def fft3d_box(x: np.ndarray, g: Grid, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    assert a.shape == _dims(g)
    lib().or_fft3d_box(ctypes.byref(g), _p(a), sign)
    return a


def btensor_box(g: Grid, C: np.ndarray, eps1: np.ndarray, eps2=None, kconv: int = 0) -> np.ndarray:
    C = np.ascontiguousarray(C, dtype=np.float64).reshape(81)
    e1 = np.ascontiguousarray(eps1, dtype=np.float64).reshape(-1, 9)
    e2 = e1 if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64).reshape(-1, 9)
    B = np.empty((e1.shape[0], e2.shape[0]) + _dims(g))
    lib().or_btensor_box(ctypes.byref(g), kconv, _p(C), _p(e1), e1.shape[0], _p(e2), e2.shape[0], _p(B))
    return B


def mart_run_box(eta, B, p: MartParams, g: Grid, nsteps: int, step0: int = 0) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    assert e.shape[1:] == _dims(g)
    lib().or_mart_run_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)),
                          _p(e), step0, nsteps)
    return e


def mart_energy_box(eta, B, p: MartParams, g: Grid) -> np.ndarray:
    e = np.ascontiguousarray(eta, dtype=np.float64)
    out = np.zeros(4)
    lib().or_mart_energy_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)),
                             _p(e), _p(out))
    return out


def mart_ref_run_box(eta, B, p: MartRefParams, g: Grid, nsteps: int, step0: int = 0, defect=None, K=None):
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    d = None if defect is None else np.ascontiguousarray(defect, dtype=np.float64)
    k = None if K is None else np.ascontiguousarray(K, dtype=np.float64)
    lib().or_mart_ref_run_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)),
                              _p(e), None if d is None else _p(d), None if k is None else _p(k), step0, nsteps)
    return e


def prec_run_box(c, eta, Bc, p: PrecParams, g: Grid, nsteps: int, step0: int = 0):
    cc = np.ascontiguousarray(c, dtype=np.float64).copy()
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    lib().or_prec_run_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(Bc, dtype=np.float64)),
                          _p(cc), _p(e), step0, nsteps)
    return cc, e


def prec_energy_box(c, eta, Bc, p: PrecParams, g: Grid) -> np.ndarray:
    cc = np.ascontiguousarray(c, dtype=np.float64)
    e = np.ascontiguousarray(eta, dtype=np.float64)
    out = np.zeros(4)
    lib().or_prec_energy_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(Bc, dtype=np.float64)),
                             _p(cc), _p(e), _p(out))
    return out


def prec_ref_run_box(c, eta, B, p: PrecParams, g: Grid, nsteps: int, step0: int = 0):
    cc = np.ascontiguousarray(c, dtype=np.float64).copy()
    e = np.ascontiguousarray(eta, dtype=np.float64).copy()
    lib().or_prec_ref_run_box(ctypes.byref(g), e.shape[0], ctypes.byref(p), _p(np.ascontiguousarray(B, dtype=np.float64)),
                              _p(cc), _p(e), step0, nsteps)
    return cc, e
