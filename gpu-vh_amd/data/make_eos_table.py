"""Generate the synthetic QCD-like equation-of-state table eos_qcdlike_2000.dat.

Model: the squared speed of sound dips from the conformal 1/3 to 0.15 at
T_c = 0.16 GeV,
    c_s^2(T) = 1/3 - (1/3 - 0.15) exp(-(T - T_c)^2 / (2 sigma^2)),  sigma = 0.03 GeV.
Thermodynamics follows from d ln s / d ln T = 1/c_s^2, p(T) = int_0^T s dT',
e = T s - p.  The entropy is normalised so that the high-temperature limit is
the reference's conformal EoS e = 13.8997 T^4 (T in fm^-1,
rhic/include/EquationOfState.cuh:24).

The table holds 2000 nodes in e on a pseudo-logarithmic grid: 50 nodes per
power of two, node k at e_k = 2^(E0 + k // 50) * (1 + (k % 50)/50) with
E0 = -24, i.e. e from 6.0e-8 to 6.5e4 fm^-4.  Within one power of two the
nodes are equally spaced, so a lookup needs only the exponent and mantissa
of e and interpolates linearly in e.  Columns: e [fm^-4], p [fm^-4],
c_s^2, T [fm^-1].

Run:  python make_eos_table.py   (writes the .dat file next to this script)
"""
import os

import numpy as np
from scipy.integrate import cumulative_trapezoid

HBARC = 0.197326938
EOS_FACTOR = 13.8997
N_NODES = 2000
PER_OCTAVE = 50
EMIN_EXP = -24
T_C, SIGMA, CS2_MIN = 0.16, 0.03, 0.15


def cs2_of_T(T_fm):
    T = T_fm * HBARC
    return 1.0 / 3.0 - (1.0 / 3.0 - CS2_MIN) * np.exp(-(T - T_C) ** 2 / (2 * SIGMA ** 2))


def node_e(k):
    k = np.asarray(k)
    return np.ldexp(1.0 + (k % PER_OCTAVE) / PER_OCTAVE, EMIN_EXP + k // PER_OCTAVE)


def thermodynamics(n=400001, lnT_lo=np.log(1e-4), lnT_hi=np.log(200.0)):
    """s, p, e on a fine grid in ln T (T in fm^-1)."""
    lnT = np.linspace(lnT_lo, lnT_hi, n)
    T = np.exp(lnT)
    inv = 1.0 / cs2_of_T(T)
    # ln s(T) = ln s(T_hi) - int_T^{T_hi} dlnT / c_s^2
    cum = cumulative_trapezoid(inv, lnT, initial=0.0)
    s_hi = 4.0 / 3.0 * EOS_FACTOR * T[-1] ** 3
    ln_s = np.log(s_hi) - (cum[-1] - cum)
    s = np.exp(ln_s)
    # p = int_0^T s dT; below the grid s ~ T^3 so p(T_lo) = s T / 4
    p = s[0] * T[0] / 4.0 + cumulative_trapezoid(s * T, lnT, initial=0.0)
    e = T * s - p
    return T, s, p, e


def build():
    T, s, p, e = thermodynamics()
    assert np.all(np.diff(e) > 0), "e(T) must increase"
    ek = node_e(np.arange(N_NODES))
    lnTk = np.interp(np.log(ek), np.log(e), np.log(T))
    # one Newton polish per node on e(T) using de/dT = T ds/dT = s / c_s^2
    Tk = np.exp(lnTk)
    for _ in range(3):
        ln_e_model = np.interp(np.log(Tk), np.log(T), np.log(e))
        # d ln e / d ln T = T s / (c_s^2 e)
        slope = np.interp(np.log(Tk), np.log(T), T * s / (cs2_of_T(T) * e))
        Tk = Tk * np.exp((np.log(ek) - ln_e_model) / slope)
    pk = np.exp(np.interp(np.log(Tk), np.log(T), np.log(p)))
    return ek, pk, cs2_of_T(Tk), Tk


def main():
    ek, pk, ck, Tk = build()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eos_qcdlike_2000.dat")
    with open(path, "w") as f:
        f.write("# synthetic QCD-like EoS: c_s^2 dips to %.2f at T=%.2f GeV (sigma %.2f GeV)\n"
                % (CS2_MIN, T_C, SIGMA))
        f.write("# nodes %d per_octave %d emin_exp %d\n" % (N_NODES, PER_OCTAVE, EMIN_EXP))
        f.write("# columns: e[fm^-4] p[fm^-4] cs2 T[fm^-1]\n")
        for row in zip(ek, pk, ck, Tk):
            f.write("%.17e %.17e %.17e %.17e\n" % row)
    print("wrote", path)


if __name__ == "__main__":
    main()
