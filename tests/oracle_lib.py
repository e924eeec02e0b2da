"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared
with; nothing in the product package imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

EOS_CONFORMAL, EOS_WB, EOS_TABLE = 0, 1, 2
# field ids (oracle/kt_oracle.h)
TTT, TTX, TTY, TTN = 0, 1, 2, 3
PITT, PITX, PITY, PITN, PIXX, PIXY, PIXN, PIYY, PIYN, PINN, PI = range(4, 15)
E, P, UT, UX, UY, UN = 15, 16, 17, 18, 19, 20
UTP, UXP, UYP, UNP = 21, 22, 23, 24
REG = 25
# validity diagnostics (HydrodynamicValidity.cu), oracle field order
(V_KN_TAUPI, V_KN_TAUBULK, V_RE_PI, V_RE2_PI, V_RE_BULK, V_RE2_BULK, V_TAUPI, V_DXUX, V_DYUY,
 V_THETA, V_REG) = range(11)
V_N = 11


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("etabar", C.c_double),
        ("bulk", C.c_int),
        ("eos", C.c_int),
        ("zeta_tpeak", C.c_double),
        ("T_reg", C.c_double),
        ("tab_n", C.c_int), ("tab_per_octave", C.c_int), ("tab_emin_exp", C.c_int),
        ("tab_p", C.POINTER(C.c_double)), ("tab_cs2", C.POINTER(C.c_double)),
        ("tab_T", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.kto_create.restype = C.c_void_p
        L.kto_create.argtypes = [C.POINTER(Params)]
        L.kto_destroy.argtypes = [C.c_void_p]
        L.kto_set_primitives.argtypes = [C.c_void_p, C.c_double] + [dp] * 8
        L.kto_step.argtypes = [C.c_void_p, C.c_double]
        L.kto_steps.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
        L.kto_euler_only.argtypes = [C.c_void_p, C.c_double]
        L.kto_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp]
        L.kto_set_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.kto_ghosted_size.restype = C.c_long
        L.kto_ghosted_size.argtypes = [C.c_void_p]
        L.kto_minmod.restype = C.c_double
        L.kto_minmod.argtypes = [C.c_double, C.c_double]
        L.kto_approx_deriv.restype = C.c_double
        L.kto_approx_deriv.argtypes = [C.c_double] * 3
        L.kto_eos.argtypes = [C.POINTER(Params), C.c_double, dp, dp, dp]
        L.kto_eos_p_batch.argtypes = [C.POINTER(Params), C.c_long, dp, dp]
        L.kto_e_of_T.restype = C.c_double
        L.kto_e_of_T.argtypes = [C.POINTER(Params), C.c_double]
        L.kto_zeta_over_s.restype = C.c_double
        L.kto_zeta_over_s.argtypes = [C.POINTER(Params), C.c_double]
        L.kto_inferred.argtypes = [C.POINTER(Params), C.c_double, dp, C.c_double, dp]
        L.kto_flux.argtypes = [C.POINTER(Params), C.c_int, dp, C.c_double, C.c_double,
                               C.c_int, dp]
        L.kto_last_newton_iters.restype = C.c_int
        L.kto_regulate_factor.restype = C.c_double
        L.kto_regulate_factor.argtypes = [C.c_double, dp, C.c_double, C.c_double, dp, C.c_double]
        L.kto_source_cell.argtypes = [C.POINTER(Params), C.c_double, dp, C.c_double,
                                      C.c_double, dp, dp, dp, dp, dp]
        L.kto_validity_cell.argtypes = [C.POINTER(Params), C.c_double, C.c_double, C.c_double,
                                        dp, dp, dp, dp, C.c_double, dp]
        L.kto_validity.argtypes = [C.c_void_p, C.c_double, C.c_int, dp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(nx, ny, nz, dx, dy, dz, dt, etabar, bulk, eos, zeta_tpeak=1.01355,
                T_reg=0.0, table=None):
    """table: gpuvh EosTable-like object with p, cs2, T arrays and grid info."""
    P = Params()
    P.nx, P.ny, P.nz = nx, ny, nz
    P.dx, P.dy, P.dz, P.dt = dx, dy, dz, dt
    P.etabar = etabar
    P.bulk = int(bulk)
    P.eos = eos
    P.zeta_tpeak = zeta_tpeak
    P.T_reg = T_reg
    keep = []
    if table is not None:
        P.tab_n = table.n
        P.tab_per_octave = table.per_octave
        P.tab_emin_exp = table.emin_exp
        for name in ("p", "cs2", "T"):
            arr = np.ascontiguousarray(getattr(table, name), dtype=np.float64)
            keep.append(arr)
            setattr(P, "tab_" + name, _ptr(arr))
    P._keep = keep  # keep table arrays alive with the struct
    return P


class Oracle:
    """One oracle context (the reference's module globals, made an object)."""

    def __init__(self, params):
        self.P = params
        self.h = lib().kto_create(C.byref(params))
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz

    def close(self):
        if self.h:
            lib().kto_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_primitives(self, t0, e, p, ut, ux, uy, un, pi10=None, Pi=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel(order="F")
                for a in (e, p, ut, ux, uy, un)]
        pi = None
        if pi10 is not None:
            pi = np.ascontiguousarray(
                np.stack([np.asarray(a, dtype=np.float64).ravel(order="F") for a in pi10]))
        Pi_a = None if Pi is None else np.ascontiguousarray(
            np.asarray(Pi, dtype=np.float64).ravel(order="F"))
        lib().kto_set_primitives(self.h, t0, *[_ptr(a) for a in arrs], _ptr(pi), _ptr(Pi_a))

    def step(self, t):
        lib().kto_step(self.h, t)

    def steps(self, t0, n_begin, nsteps):
        lib().kto_steps(self.h, t0, n_begin, nsteps)

    def euler_only(self, t):
        lib().kto_euler_only(self.h, t)

    def get(self, field, ghosted=False, from_stage=False):
        """Field as a (nx, ny, nz) array (Fortran order: i fastest)."""
        if ghosted:
            n = lib().kto_ghosted_size(self.h)
            out = np.empty(n, dtype=np.float64)
            lib().kto_get(self.h, field, 1, int(from_stage), _ptr(out))
            return out.reshape((self.nx + 4, self.ny + 4, self.nz + 4), order="F")
        out = np.empty(self.nx * self.ny * self.nz, dtype=np.float64)
        lib().kto_get(self.h, field, 0, int(from_stage), _ptr(out))
        return out.reshape((self.nx, self.ny, self.nz), order="F")

    def validity(self, t, prev_q):
        """The V_N validity arrays as (V_N, nx, ny, nz) (interior cells);
        prev_q: pi and Pi from q^n of the last step, as the reference."""
        n = lib().kto_ghosted_size(self.h)
        out = np.zeros(V_N * n, dtype=np.float64)
        lib().kto_validity(self.h, t, int(prev_q), _ptr(out))
        g = out.reshape((V_N, self.nz + 4, self.ny + 4, self.nx + 4))
        return np.ascontiguousarray(g[:, 2:-2, 2:-2, 2:-2].transpose(0, 3, 2, 1))

    def set_field(self, field, ghosted_array):
        a = np.ascontiguousarray(np.asarray(ghosted_array, dtype=np.float64).ravel(order="F"))
        lib().kto_set_field(self.h, field, _ptr(a))


def eos(params, e):
    p, cs2, T = C.c_double(), C.c_double(), C.c_double()
    lib().kto_eos(C.byref(params), e, C.byref(p), C.byref(cs2), C.byref(T))
    return p.value, cs2.value, T.value


def eos_p_batch(params, e):
    e = np.ascontiguousarray(np.asarray(e, dtype=np.float64))
    out = np.empty_like(e)
    lib().kto_eos_p_batch(C.byref(params), e.size, _ptr(e), _ptr(out))
    return out


def inferred(params, t, q, e_prev):
    q = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
    out = np.empty(6, dtype=np.float64)
    lib().kto_inferred(C.byref(params), t, _ptr(q), e_prev, _ptr(out))
    return out


def source_cell(params, t, Q, e, p, u, up, du, dp):
    """loadSourceTerms2 of one cell; du is (3, 4) = d_i u^m, dp (3,)."""
    arr = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (Q, u, up, du, dp)]
    S = np.zeros(15, dtype=np.float64)
    lib().kto_source_cell(C.byref(params), t, _ptr(arr[0]), e, p, _ptr(arr[1]), _ptr(arr[2]),
                          _ptr(arr[3]), _ptr(arr[4]), _ptr(S))
    return S


def validity_cell(params, t, e, p, u, up, du, pi10, Pi):
    """checkValidityKernel of one cell; du is (3, 4) = d_i u^m.  Returns the
    first V_REG validity values."""
    arr = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (u, up, du, pi10)]
    out = np.zeros(V_N, dtype=np.float64)
    lib().kto_validity_cell(C.byref(params), t, e, p, _ptr(arr[0]), _ptr(arr[1]), _ptr(arr[2]),
                            _ptr(arr[3]), Pi, _ptr(out))
    return out[:V_REG]


def regulate_factor(t, pi10, e, p, u, e_reg):
    pi = np.ascontiguousarray(np.asarray(pi10, dtype=np.float64))
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    return lib().kto_regulate_factor(t, _ptr(pi), e, p, _ptr(uu), e_reg)


def flux(params, direction, data, t, e_prev, fwd):
    data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
    H = np.empty(15, dtype=np.float64)
    lib().kto_flux(C.byref(params), direction, _ptr(data), t, e_prev, int(fwd), _ptr(H))
    return H


def e_of_T(params, T):
    """equilibriumEnergyDensity(T) of the oracle's EoS."""
    return lib().kto_e_of_T(C.byref(params), T)
