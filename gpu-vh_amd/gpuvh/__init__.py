"""Python host binding of the MI355X gpu-vh KT library (libgpuvh.so).

The library is the product: a C ABI (include/gpuvh.h) over HIP kernels for
gfx950.  This module is a thin ctypes layer mirroring the reference's driver
surface -- LatticeParameters / HydroParameters / InitialConditionParameters,
the conserved-variable field order and twoStepRungeKutta -- so that tests,
benchmarks and embedding code can drive it.  It never falls back to a CPU
implementation: if the shared library is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# GPUVH_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GPUVH_LIB") or os.path.join(PKG_DIR, "lib", "libgpuvh.so")
DATA_DIR = os.path.join(PKG_DIR, "data")
EOS_TABLE_PATH = os.path.join(DATA_DIR, "eos_qcdlike_2000.dat")

HBARC = 0.197326938  # GeV fm (rhic/src/HydroPlugin.cpp:132)

EOS_CONFORMAL, EOS_WB, EOS_TABLE = 0, 1, 2
FP64, MIXED, MIXED_F64_ROOTS = 0, 1, 2
# status codes (include/gpuvh.h GVH_*)
OK, EINVAL, EHIP, EIO, ENOMEM, ENOCONV = 0, -1, -2, -3, -4, -5

# field ids, include/gpuvh.h
(TTT, TTX, TTY, TTN, PITT, PITX, PITY, PITN, PIXX, PIXY, PIXN, PIYY, PIYN, PINN, PI,
 E, P, UT, UX, UY, UN, UTP, UXP, UYP, UNP) = range(25)
# freeze-out: surface row columns and flags (include/gpuvh.h)
FO_COLUMNS = 20
FO_EXACT_TAU, FO_NORMAL_TAU, FO_CURRENT_FIELDS = 1, 2, 4
FO_VARS = ["ux", "uy", "un", "e", "pixx", "pixy", "pixn", "piyy", "piyn", "Pi"]

# validity diagnostics (include/gpuvh.h GVH_V_*): HydrodynamicValidity.cu
(V_KN_TAUPI, V_KN_TAUBULK, V_RE_PI, V_RE2_PI, V_RE_BULK, V_RE2_BULK, V_TAUPI, V_DXUX, V_DYUY,
 V_THETA, V_REGULATIONS) = range(11)
VALIDITY_FIELDS = 11
VALIDITY_CURRENT_PI = 1
VALIDITY_NAMES = ["KnTaupi", "KnTauPi", "Rpi", "R2pi", "RPi", "R2Pi", "taupi", "dxux", "dyuy",
                  "theta", "regulations"]

FIELD_NAMES = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy",
               "pixn", "piyy", "piyn", "pinn", "Pi", "e", "p", "ut", "ux", "uy", "un",
               "utp", "uxp", "uyp", "unp"]


class GvhError(RuntimeError):
    pass


class LatticeParameters(C.Structure):
    """rhic/include/LatticeParameters.h:17-31"""
    _fields_ = [
        ("numLatticePointsX", C.c_int), ("numLatticePointsY", C.c_int),
        ("numLatticePointsRapidity", C.c_int),
        ("numComputationalLatticePointsX", C.c_int), ("numComputationalLatticePointsY", C.c_int),
        ("numComputationalLatticePointsRapidity", C.c_int),
        ("numProperTimePoints", C.c_int),
        ("latticeSpacingX", C.c_double), ("latticeSpacingY", C.c_double),
        ("latticeSpacingRapidity", C.c_double), ("latticeSpacingProperTime", C.c_double),
    ]


class InitialConditionParameters(C.Structure):
    """rhic/include/InitialConditionParameters.h:13-26"""
    _fields_ = [
        ("initialConditionType", C.c_int), ("numberOfNucleonsPerNuclei", C.c_int),
        ("numberOfSourceFiles", C.c_int), ("initialEnergyDensity", C.c_double),
        ("scatteringCrossSectionNN", C.c_double), ("impactParameter", C.c_double),
        ("fractionOfBinaryCollisions", C.c_double), ("rapidityVariance", C.c_double),
        ("rapidityMean", C.c_double), ("sourceType", C.c_int),
    ]


class HydroParameters(C.Structure):
    """rhic/include/HydroParameters.h:15-22"""
    _fields_ = [
        ("initialProperTimePoint", C.c_double), ("shearViscosityToEntropyDensity", C.c_double),
        ("freezeoutTemperatureGeV", C.c_double), ("initializePimunuNavierStokes", C.c_int),
        ("regulationTemperatureGeV", C.c_double),
    ]


class Options(C.Structure):
    _fields_ = [
        ("eos", C.c_int), ("bulk", C.c_int), ("precision", C.c_int),
        ("zeta_tpeak", C.c_double),
        ("tab_n", C.c_int), ("tab_per_octave", C.c_int), ("tab_emin_exp", C.c_int),
        ("tab_p", C.POINTER(C.c_double)), ("tab_cs2", C.POINTER(C.c_double)),
        ("tab_T", C.POINTER(C.c_double)),
        ("stream", C.c_void_p),
    ]


class RunOptions(C.Structure):
    """gvh_run_options: FREQ, FOFREQ (rhic/src/HydroPlugin.cpp:39-40), freeze-out flags."""
    _fields_ = [("freq", C.c_int), ("fofreq", C.c_int), ("fo_flags", C.c_int),
                ("snapshots", C.c_int), ("surface", C.c_int), ("verbose", C.c_int),
                ("validity", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("frozen", C.c_int), ("t_final", C.c_double),
                ("surface_rows", C.c_longlong), ("ms_steps", C.c_double),
                ("ms_freezeout", C.c_double), ("ms_total", C.c_double)]


class BatchOptions(C.Structure):
    _fields_ = [("hbm_bytes", C.c_ulonglong), ("max_events", C.c_int), ("fofreq", C.c_int),
                ("autotune", C.c_int)]


class Event(C.Structure):
    _fields_ = [("lattice", LatticeParameters), ("hydro", HydroParameters),
                ("e", C.POINTER(C.c_double)), ("surface", C.c_int), ("fo_flags", C.c_int),
                ("keep_state", C.c_int), ("newton_budget", C.c_int)]


class EventResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("steps", C.c_int), ("frozen", C.c_int),
                ("t_final", C.c_double), ("surface_rows", C.c_longlong),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("status", C.c_int), ("newton_not_converged", C.c_ulonglong)]


class Diagnostics(C.Structure):
    _fields_ = [("newton_not_converged", C.c_ulonglong), ("e_nan", C.c_ulonglong),
                ("e_clamped", C.c_ulonglong)]


_lib = None


def lib():
    """Load libgpuvh.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GvhError("libgpuvh.so not built (%s); run __graft_entry__.build() or "
                       "make -C gpu-vh_amd" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    L.gvh_last_error.restype = C.c_char_p
    sig = {
        "gvh_read_lattice_parameters": [C.c_char_p, C.POINTER(LatticeParameters)],
        "gvh_read_hydro_parameters": [C.c_char_p, C.POINTER(HydroParameters)],
        "gvh_read_ic_parameters": [C.c_char_p, C.POINTER(InitialConditionParameters)],
        "gvh_create": [C.POINTER(LatticeParameters), C.POINTER(HydroParameters),
                       C.POINTER(Options), C.POINTER(C.c_void_p)],
        "gvh_destroy": [C.c_void_p],
        "gvh_set_initial_state": [C.c_void_p, C.c_double] + [dp] * 8,
        "gvh_set_field": [C.c_void_p, C.c_int, dp],
        "gvh_step": [C.c_void_p, C.c_double],
        "gvh_steps": [C.c_void_p, C.c_double, C.c_int, C.c_int],
        "gvh_step_timed": [C.c_void_p, C.c_double, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "gvh_steps_timed": [C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_float),
                            C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "gvh_synchronize": [C.c_void_p],
        "gvh_get_field": [C.c_void_p, C.c_int, C.c_int, dp],
        "gvh_count_above": [C.c_void_p, C.c_double, C.POINTER(C.c_longlong)],
        "gvh_validity_enable": [C.c_void_p],
        "gvh_check_validity": [C.c_void_p, C.c_double, C.c_int],
        "gvh_get_validity": [C.c_void_p, C.c_int, C.c_int, dp],
        "gvh_get_diagnostics": [C.c_void_p, C.POINTER(Diagnostics)],
        "gvh_energy_density_of_T": [C.c_void_p, C.c_double, dp],
        "gvh_equilibrium_pressure": [C.c_void_p, C.c_longlong, dp, dp],
        "gvh_dims": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "gvh_freezeout_init": [C.c_void_p, C.c_int, C.c_double, C.c_int],
        "gvh_freezeout_record": [C.c_void_p, C.c_int],
        "gvh_freezeout_find": [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_longlong)],
        "gvh_freezeout_rows": [C.c_void_p, dp, C.c_longlong],
        "gvh_freezeout_slot": [C.c_void_p, C.c_int, C.c_int, dp],
        "gvh_cornelius_cells": [C.c_int, C.c_longlong, dp, dp, C.c_double, C.POINTER(C.c_int),
                                dp, dp],
        "gvh_cornelius_cells_noalloc": [C.c_int, C.c_longlong, C.c_void_p, dp, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
        "gvh_run": [C.POINTER(LatticeParameters), C.POINTER(InitialConditionParameters),
                    C.POINTER(HydroParameters), C.c_char_p, C.c_char_p, C.POINTER(Options),
                    C.POINTER(RunOptions), C.POINTER(RunStats)],
        "gvh_run_hydro": [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Options),
                          C.POINTER(RunOptions)],
    }
    L.gvh_run_options_default.argtypes = [C.POINTER(RunOptions)]
    L.gvh_run_options_default.restype = None
    L.gvh_batch_options_default.argtypes = [C.POINTER(BatchOptions)]
    L.gvh_batch_options_default.restype = None
    bsig = {
        "gvh_batch_create": [C.POINTER(Options), C.POINTER(BatchOptions), C.POINTER(C.c_void_p)],
        "gvh_batch_destroy": [C.c_void_p],
        "gvh_batch_submit": [C.c_void_p, C.POINTER(Event), C.POINTER(C.c_longlong)],
        "gvh_batch_step": [C.c_void_p, C.POINTER(C.c_int)],
        "gvh_batch_run": [C.c_void_p],
        "gvh_batch_poll": [C.c_void_p, C.POINTER(EventResult), C.c_int, C.POINTER(C.c_int)],
        "gvh_batch_rows": [C.c_void_p, C.c_longlong, dp, C.c_longlong],
        "gvh_batch_field": [C.c_void_p, C.c_longlong, C.c_int, dp],
        "gvh_batch_release": [C.c_void_p, C.c_longlong],
        "gvh_batch_stats": [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                            C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)],
        "gvh_autotune": [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                         C.POINTER(C.c_float)],
    }
    for name, args in bsig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().gvh_last_error()
        raise GvhError("gpuvh error %d: %s" % (rc, msg.decode() if msg else ""))


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_lattice_parameters(config_dir):
    L = LatticeParameters()
    _check(lib().gvh_read_lattice_parameters(config_dir.encode(), C.byref(L)))
    return L


def read_hydro_parameters(config_dir):
    H = HydroParameters()
    _check(lib().gvh_read_hydro_parameters(config_dir.encode(), C.byref(H)))
    return H


def read_ic_parameters(config_dir):
    I = InitialConditionParameters()
    _check(lib().gvh_read_ic_parameters(config_dir.encode(), C.byref(I)))
    return I


def lattice(nx, ny, nz, dx, dy, dz, dt, nt=0):
    L = LatticeParameters()
    L.numLatticePointsX, L.numLatticePointsY, L.numLatticePointsRapidity = nx, ny, nz
    L.numComputationalLatticePointsX = nx + 4
    L.numComputationalLatticePointsY = ny + 4
    L.numComputationalLatticePointsRapidity = nz + 4
    L.numProperTimePoints = nt
    L.latticeSpacingX, L.latticeSpacingY = dx, dy
    L.latticeSpacingRapidity, L.latticeSpacingProperTime = dz, dt
    return L


def hydro(t0, eta_over_s, T_fo_GeV=0.15, T_reg_GeV=0.0):
    H = HydroParameters()
    H.initialProperTimePoint = t0
    H.shearViscosityToEntropyDensity = eta_over_s
    H.freezeoutTemperatureGeV = T_fo_GeV
    H.initializePimunuNavierStokes = 0
    H.regulationTemperatureGeV = T_reg_GeV
    return H


class EosTable:
    """Tabulated EoS on the pseudo-log grid (see data/make_eos_table.py)."""

    def __init__(self, path=EOS_TABLE_PATH):
        n = per_octave = emin_exp = None
        with open(path) as f:
            for line in f:
                if not line.startswith("#"):
                    break
                tok = line[1:].split()
                if tok and tok[0] == "nodes":
                    n, per_octave, emin_exp = int(tok[1]), int(tok[3]), int(tok[5])
        d = np.loadtxt(path)
        self.e, self.p, self.cs2, self.T = (np.ascontiguousarray(c) for c in d.T)
        self.n = n if n is not None else len(self.e)
        self.per_octave, self.emin_exp = per_octave, emin_exp

    def node_e(self, k):
        k = np.asarray(k)
        return np.ldexp(1.0 + (k % self.per_octave) / self.per_octave,
                        self.emin_exp + k // self.per_octave)


def options(eos=EOS_WB, bulk=True, precision=FP64, zeta_tpeak=1.01355, table=None,
            stream=None):
    o = Options()
    o.eos, o.bulk, o.precision, o.zeta_tpeak = eos, int(bulk), precision, zeta_tpeak
    o._keep = []
    if eos == EOS_TABLE:
        if table is None:
            table = EosTable()
        o.tab_n, o.tab_per_octave, o.tab_emin_exp = table.n, table.per_octave, table.emin_exp
        for name in ("p", "cs2", "T"):
            a = np.ascontiguousarray(getattr(table, name), dtype=np.float64)
            o._keep.append(a)
            setattr(o, "tab_" + name, _ptr(a))
    o.stream = stream
    return o


class Context:
    """One lattice's device state (the reference's module globals as an object)."""

    def __init__(self, lattice_params, hydro_params, opts):
        self.L, self.H, self.opts = lattice_params, hydro_params, opts
        h = C.c_void_p()
        _check(lib().gvh_create(C.byref(lattice_params), C.byref(hydro_params),
                                C.byref(opts), C.byref(h)))
        self.h = h
        self.nx = lattice_params.numLatticePointsX
        self.ny = lattice_params.numLatticePointsY
        self.nz = lattice_params.numLatticePointsRapidity
        self.dt = lattice_params.latticeSpacingProperTime

    def close(self):
        if self.h:
            lib().gvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _flat(a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))

    def set_initial_state(self, t0, e, p, ut, ux, uy, un, pi10=None, Pi=None):
        n = self.nx * self.ny * self.nz
        arrs = [self._flat(a) for a in (e, p, ut, ux, uy, un)]
        pi = None if pi10 is None else np.ascontiguousarray(
            np.stack([self._flat(a) for a in pi10]))
        Pi_a = None if Pi is None else self._flat(Pi)
        # the library reads exactly n values per field (10 n for pi)
        for name, a in zip(("e", "p", "ut", "ux", "uy", "un"), arrs):
            if a.size != n:
                raise ValueError("%s has %d values, the lattice %d" % (name, a.size, n))
        if pi is not None and pi.shape != (10, n):
            raise ValueError("pi10 must hold 10 fields of %d values, got %s" % (n, pi.shape))
        if Pi_a is not None and Pi_a.size != n:
            raise ValueError("Pi has %d values, the lattice %d" % (Pi_a.size, n))
        _check(lib().gvh_set_initial_state(self.h, t0, *[_ptr(a) for a in arrs],
                                           _ptr(pi), _ptr(Pi_a)))

    def step(self, t):
        _check(lib().gvh_step(self.h, t))

    def steps(self, t0, n_begin, nsteps):
        _check(lib().gvh_steps(self.h, t0, n_begin, nsteps))

    def step_timed(self, t):
        a, b = C.c_float(), C.c_float()
        _check(lib().gvh_step_timed(self.h, t, C.byref(a), C.byref(b)))
        return a.value, b.value

    def steps_timed(self, t0, n_begin, nsteps):
        """nsteps steps with HIP events around every stage kernel; returns the
        summed stage-1 and stage-2 kernel ms and the GPU ms of the sequence."""
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        _check(lib().gvh_steps_timed(self.h, t0, n_begin, nsteps, C.byref(a), C.byref(b),
                                     C.byref(c)))
        return a.value, b.value, c.value

    def synchronize(self):
        _check(lib().gvh_synchronize(self.h))

    def get(self, field, ghosted=False):
        """Field as an (nx, ny, nz) array (x fastest) or the ghosted grid."""
        if ghosted:
            shape = (self.nx + 4, self.ny + 4, self.nz + 4)
        else:
            shape = (self.nx, self.ny, self.nz)
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _check(lib().gvh_get_field(self.h, field, int(ghosted), _ptr(out)))
        return out.reshape(shape, order="F")

    def validity_enable(self):
        """Allocate the validity arrays and record the regulation factor of
        every later step (gvh_validity_enable)."""
        _check(lib().gvh_validity_enable(self.h))

    def check_validity(self, t, flags=0):
        """checkValidity at proper time t (gvh_check_validity)."""
        _check(lib().gvh_check_validity(self.h, t, flags))

    def get_validity(self, field, ghosted=False):
        """Validity field V_* as an (nx, ny, nz) array or the ghosted grid."""
        if ghosted:
            shape = (self.nx + 4, self.ny + 4, self.nz + 4)
        else:
            shape = (self.nx, self.ny, self.nz)
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _check(lib().gvh_get_validity(self.h, field, int(ghosted), _ptr(out)))
        return out.reshape(shape, order="F")

    def set_field(self, field, ghosted_array):
        a = self._flat(ghosted_array)
        _check(lib().gvh_set_field(self.h, field, _ptr(a)))

    def count_above(self, e_threshold):
        c = C.c_longlong()
        _check(lib().gvh_count_above(self.h, e_threshold, C.byref(c)))
        return c.value

    def diagnostics(self):
        d = Diagnostics()
        _check(lib().gvh_get_diagnostics(self.h, C.byref(d)))
        return {"newton_not_converged": d.newton_not_converged, "e_nan": d.e_nan,
                "e_clamped": d.e_clamped}

    def equilibrium_pressure(self, e):
        e = np.ascontiguousarray(np.asarray(e, dtype=np.float64))
        p = np.empty_like(e)
        _check(lib().gvh_equilibrium_pressure(self.h, e.size, _ptr(e), _ptr(p)))
        return p

    def energy_density_of_T(self, T):
        e = C.c_double()
        _check(lib().gvh_energy_density_of_T(self.h, T, C.byref(e)))
        return e.value

    # -------------------------------------------------------- freeze-out
    def freezeout_init(self, fofreq, e_fo, flags=0):
        _check(lib().gvh_freezeout_init(self.h, fofreq, e_fo, flags))

    def freezeout_record(self, n):
        _check(lib().gvh_freezeout_record(self.h, n))

    def freezeout_find(self, n, t):
        """Search the window of loop index n; returns the (nrows, 20) rows."""
        cnt = C.c_longlong()
        _check(lib().gvh_freezeout_find(self.h, n, t, C.byref(cnt)))
        rows = np.empty((cnt.value, FO_COLUMNS), dtype=np.float64)
        if cnt.value:
            _check(lib().gvh_freezeout_rows(self.h, _ptr(rows), cnt.value))
        return rows

    def autotune(self, steps=3):
        """Time the kernel arrangements on the current state (restored) and
        keep the fastest; returns (variant, strip, ms per step)."""
        v, st, ms = C.c_int(), C.c_int(), C.c_float()
        _check(lib().gvh_autotune(self.h, steps, C.byref(v), C.byref(st), C.byref(ms)))
        return v.value, st.value, ms.value

    def freezeout_slot(self, logical, var):
        out = np.empty(self.nx * self.ny * self.nz, dtype=np.float64)
        _check(lib().gvh_freezeout_slot(self.h, logical, var, _ptr(out)))
        return out.reshape((self.nx, self.ny, self.nz), order="F")


def cornelius_cells(dim, corners, dx, v0):
    """Surface elements of independent cells on the device.  corners:
    (ncells, 2^dim), last axis fastest.  Returns (counts, centroids, normals)
    with centroids/normals of shape (ncells, 10, dim)."""
    v = np.ascontiguousarray(np.asarray(corners, dtype=np.float64).reshape(-1, 1 << dim))
    n = v.shape[0]
    d = np.ascontiguousarray(np.asarray(dx, dtype=np.float64))
    counts = np.zeros(n, dtype=np.int32)
    cen = np.zeros((n, 10, dim))
    nrm = np.zeros((n, 10, dim))
    _check(lib().gvh_cornelius_cells(dim, n, _ptr(v), _ptr(d), v0,
                                     counts.ctypes.data_as(C.POINTER(C.c_int)), _ptr(cen),
                                     _ptr(nrm)))
    return counts, cen, nrm


def run_options(freq=10, fofreq=10, fo_flags=0, snapshots=True, surface=True, verbose=False,
                validity=False):
    ro = RunOptions()
    lib().gvh_run_options_default(C.byref(ro))
    ro.freq, ro.fofreq, ro.fo_flags = freq, fofreq, fo_flags
    ro.snapshots, ro.surface, ro.verbose = int(snapshots), int(surface), int(verbose)
    ro.validity = int(validity)
    return ro


def run(lattice_params, ic_params, hydro_params, root_dir, output_dir, opts, ro=None):
    """One event through the run loop (rhic/src/HydroPlugin.cpp:107-434)."""
    st = RunStats()
    _check(lib().gvh_run(C.byref(lattice_params), C.byref(ic_params), C.byref(hydro_params),
                         root_dir.encode(), output_dir.encode(), C.byref(opts),
                         C.byref(ro) if ro is not None else None, C.byref(st)))
    return {f: getattr(st, f) for f, _ in RunStats._fields_}


class Batch:
    """Many events advancing together on the device (include/gpuvh.h, batches)."""

    def __init__(self, opts, hbm_bytes=0, max_events=64, fofreq=10, autotune=True):
        bo = BatchOptions()
        lib().gvh_batch_options_default(C.byref(bo))
        bo.hbm_bytes, bo.max_events, bo.fofreq, bo.autotune = hbm_bytes, max_events, fofreq, int(autotune)
        self.opts = opts
        h = C.c_void_p()
        _check(lib().gvh_batch_create(C.byref(opts), C.byref(bo), C.byref(h)))
        self.h = h
        self.shapes = {}

    def close(self):
        if self.h:
            lib().gvh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, lattice_params, hydro_params, e, surface=False, fo_flags=0, keep_state=False,
               newton_budget=0):
        """Queue an event; newton_budget > 0 retires it with status
        ENOCONV after a step in which a root-find reached that many
        iterations (0: the reference's limit)."""
        ev = Event()
        ev.lattice, ev.hydro = lattice_params, hydro_params
        a = Context._flat(e)
        n = (lattice_params.numLatticePointsX * lattice_params.numLatticePointsY
             * lattice_params.numLatticePointsRapidity)
        if a.size != n:
            raise ValueError("e has %d values, the lattice %d" % (a.size, n))
        ev.e = _ptr(a)
        ev.surface, ev.fo_flags, ev.keep_state = int(surface), fo_flags, int(keep_state)
        ev.newton_budget = int(newton_budget)
        i = C.c_longlong()
        _check(lib().gvh_batch_submit(self.h, C.byref(ev), C.byref(i)))
        self.shapes[i.value] = (lattice_params.numLatticePointsX, lattice_params.numLatticePointsY,
                                lattice_params.numLatticePointsRapidity)
        return i.value

    def step(self):
        n = C.c_int()
        _check(lib().gvh_batch_step(self.h, C.byref(n)))
        return n.value

    def run(self):
        _check(lib().gvh_batch_run(self.h))

    def poll(self, max_results=1024):
        out = (EventResult * max_results)()
        n = C.c_int()
        _check(lib().gvh_batch_poll(self.h, out, max_results, C.byref(n)))
        return [{f: getattr(out[k], f) for f, _ in EventResult._fields_} for k in range(n.value)]

    def rows(self, event_id, nrows):
        r = np.empty((nrows, FO_COLUMNS), dtype=np.float64)
        if nrows:
            _check(lib().gvh_batch_rows(self.h, event_id, _ptr(r), nrows))
        return r

    def field(self, event_id, field):
        nx, ny, nz = self.shapes[event_id]
        out = np.empty(nx * ny * nz, dtype=np.float64)
        _check(lib().gvh_batch_field(self.h, event_id, field, _ptr(out)))
        return out.reshape((nx, ny, nz), order="F")

    def release(self, event_id):
        _check(lib().gvh_batch_release(self.h, event_id))

    def stats(self):
        q, r = C.c_longlong(), C.c_longlong()
        u, b = C.c_ulonglong(), C.c_ulonglong()
        _check(lib().gvh_batch_stats(self.h, C.byref(q), C.byref(r), C.byref(u), C.byref(b)))
        return {"queued": q.value, "running": r.value, "hbm_used": u.value, "hbm_budget": b.value}
