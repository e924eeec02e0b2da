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


# Floor of the per-cell relative error, as a fraction of the field's tensor
# scale (floor_scales).  A component much smaller than its tensor is a sum of
# O(scale) terms that cancel, so its rounding noise is ~1e-15 x scale whatever
# the operation order; below 1e-4 x scale the 1e-10 relative bar therefore
# becomes an absolute bar of 1e-14 x scale.
FLOOR_FRAC = 1e-4


def rel_err(x, ref, floor_frac=FLOOR_FRAC, scale=None):
    """max_cells |x - ref| / max(|ref|, floor_frac * scale).

    scale defaults to the field's own maximum; a field that is zero
    everywhere must match exactly."""
    if scale is None:
        scale = float(np.max(np.abs(ref)))
    if scale == 0.0:
        return 0.0 if not np.any(x != ref) else float("inf")
    den = np.maximum(np.abs(ref), floor_frac * scale)
    return float(np.max(np.abs(x - ref) / den))


# number of eta indices of each field: in Milne coordinates a component with
# n eta indices carries tau^-n relative to its orthonormal-frame value
ETA_INDICES = {gpuvh.UN: 1, gpuvh.PITN: 1, gpuvh.PIXN: 1, gpuvh.PIYN: 1, gpuvh.PINN: 2}


def floor_scales(ref, tau):
    """Per-field scale for the relative-error floor: components of u^mu and
    of pi^{mu nu} are floored at FLOOR_FRAC of the largest component of the
    same tensor in the orthonormal frame (tau^n_eta x component), converted
    back to the field's own units.  Components that vanish by symmetry (u^eta,
    pi^{x eta}, ... on the eta = 0 plane) are computed as sums of O(1) terms
    that cancel, so their rounding noise is set by the tensor's scale, not by
    their own.  Scalars (e, Pi) use their own maximum."""
    def phys_max(fields):
        return max(float(np.max(np.abs(ref[f]))) * tau ** ETA_INDICES.get(f, 0) for f in fields)
    out = {}
    u_f = [f for f in U_FIELDS if f in ref]
    pi_f = [f for f in PI_FIELDS if f in ref]
    su = phys_max(u_f) if u_f else 0.0
    sp = phys_max(pi_f) if pi_f else 0.0
    for f in ref:
        if f in U_FIELDS:
            out[f] = su / tau ** ETA_INDICES.get(f, 0)
        elif f in PI_FIELDS:
            out[f] = sp / tau ** ETA_INDICES.get(f, 0)
        else:
            out[f] = float(np.max(np.abs(ref[f])))
    return out


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
        return O.eos_p_batch(self.oracle().P, e)


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


def compare(orc, ctx, fields, tau, report=None):
    """{field name: relative error} for the listed fields, with the floors
    of floor_scales.  report (a list) collects the worst cell per field."""
    ref = {f: orc.get(f) for f in fields}
    got = {f: ctx.get(f) for f in fields}
    scales = floor_scales(ref, tau)
    out = {}
    for f in fields:
        name = gpuvh.FIELD_NAMES[f]
        out[name] = rel_err(got[f], ref[f], scale=scales[f])
        if report is not None and scales[f] > 0:
            den = np.maximum(np.abs(ref[f]), FLOOR_FRAC * scales[f])
            r = np.abs(got[f] - ref[f]) / den
            w = np.unravel_index(np.argmax(r), r.shape)
            report.append((name, out[name], w, float(ref[f][w]), float(got[f][w] - ref[f][w]),
                           scales[f]))
    return out
