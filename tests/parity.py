"""Shared helpers of the parity tests: build the same seeded event for the
HIP library and for the CPU oracle, and compare fields per cell."""
import numpy as np

import gpuvh
import oracle_lib as O
from gpuvh import ic

# fields compared per cell: e, u^mu, pi^{mu nu} (and Pi when evolved)
PI_FIELDS = [gpuvh.PITT, gpuvh.PITX, gpuvh.PITY, gpuvh.PITN, gpuvh.PIXX, gpuvh.PIXY,
             gpuvh.PIXN, gpuvh.PIYY, gpuvh.PIYN, gpuvh.PINN]
U_FIELDS = [gpuvh.UT, gpuvh.UX, gpuvh.UY, gpuvh.UN]


def rel_err(x, ref, floor_frac=1e-6):
    """max_cells |x - ref| / max(|ref|, floor_frac * max_cells |ref|).

    The floor keeps components that cross zero (u^x, pi^{xy}, ...) from
    dividing by ~0; a field that is zero everywhere must match to 1e-300."""
    scale = float(np.max(np.abs(ref)))
    if scale == 0.0:
        return float(np.max(np.abs(x))) / 1e-300 if np.any(x != 0) else 0.0
    den = np.maximum(np.abs(ref), floor_frac * scale)
    return float(np.max(np.abs(x - ref) / den))


class Case:
    """One lattice + physics configuration, runnable on both sides."""

    def __init__(self, nx, ny, nz, dx, dz, dt, t0, etabar, bulk, eos, zeta_tpeak=1.01355,
                 T_reg_GeV=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dz, self.dt, self.t0 = dx, dz, dt, t0
        self.etabar, self.bulk, self.eos = etabar, bulk, eos
        self.zeta_tpeak, self.T_reg_GeV = zeta_tpeak, T_reg_GeV
        self.table = gpuvh.EosTable() if eos == gpuvh.EOS_TABLE else None

    def oracle(self):
        P = O.make_params(self.nx, self.ny, self.nz, self.dx, self.dx, self.dz, self.dt,
                          self.etabar, self.bulk, self.eos, zeta_tpeak=self.zeta_tpeak,
                          T_reg=self.T_reg_GeV / gpuvh.HBARC, table=self.table)
        return O.Oracle(P)

    def context(self, precision=gpuvh.FP64):
        L = gpuvh.lattice(self.nx, self.ny, self.nz, self.dx, self.dx, self.dz, self.dt)
        H = gpuvh.hydro(self.t0, self.etabar, T_reg_GeV=self.T_reg_GeV)
        o = gpuvh.options(self.eos, self.bulk, precision, self.zeta_tpeak, table=self.table)
        return gpuvh.Context(L, H, o)

    def pressure(self, e):
        orc = self.oracle()
        f = np.vectorize(lambda x: O.eos(orc.P, float(x))[0])
        return f(e)


def hotspot_state(case, seed, n_hot=20, amp=8.0, sigma=0.5, pi_noise=0.0, u_noise=0.0):
    """Seeded e on the case lattice (eta plateau profile when nz > 1), u at
    rest unless u_noise, pi zero unless pi_noise (random symmetric values of
    the stated relative size, for exercising the regulation path)."""
    e2 = ic.hotspot_event_2d(case.nx, case.ny, case.dx, case.dx, seed, n_hot=n_hot,
                             hot_sigma=sigma, hot_amp_max=amp, base_peak=10.0, base_sigma=1.0,
                             centre_radius=0.35 * case.nx * case.dx)
    eL = ic.longitudinal_profile(case.nz, case.dz, eta_flat=0.3 * case.nz * case.dz,
                                 eta_variance=0.05) if case.nz > 1 else np.ones(1)
    e = ic.quantise_like_reference(e2[:, :, None] * eL[None, None, :])
    shape = e.shape
    rng = np.random.default_rng(seed)
    ut, ux, uy, un, pi10, Pi = ic.at_rest(shape)
    if u_noise:
        ux = u_noise * rng.standard_normal(shape)
        uy = u_noise * rng.standard_normal(shape)
        un = 0.1 * u_noise * rng.standard_normal(shape) if case.nz > 1 else np.zeros(shape)
        ut = np.sqrt(1 + ux ** 2 + uy ** 2 + (case.t0 * un) ** 2)
    if pi_noise:
        pi10 = [pi_noise * e * rng.standard_normal(shape) for _ in range(10)]
        if case.nz == 1:
            for m in (3, 6, 8):  # tn, xn, yn vanish in 2+1D
                pi10[m] = np.zeros(shape)
    p = case.pressure(e)
    return e, p, ut, ux, uy, un, pi10, Pi


def run_both(case, state, nsteps, precision=gpuvh.FP64):
    """Evolve the same state nsteps on the oracle and on the device."""
    e, p, ut, ux, uy, un, pi10, Pi = state
    orc = case.oracle()
    orc.set_primitives(case.t0, e, p, ut, ux, uy, un, pi10, Pi)
    orc.steps(case.t0, 0, nsteps)
    ctx = case.context(precision)
    ctx.set_initial_state(case.t0, e, p, ut, ux, uy, un, pi10, Pi)
    ctx.steps(case.t0, 0, nsteps)
    ctx.synchronize()
    return orc, ctx


def compare(orc, ctx, fields):
    """{field name: relative error} for the listed fields."""
    out = {}
    for f in fields:
        out[gpuvh.FIELD_NAMES[f]] = rel_err(ctx.get(f), orc.get(f))
    return out
