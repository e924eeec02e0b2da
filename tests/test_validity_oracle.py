"""Pins of the oracle's validity diagnostics (kto_validity_cell and
kto_validity, restating HydrodynamicValidity.cu:19-235) against what the
mathematics fixes:

- Bjorken flow (u = (1,0,0,0)): theta = 1/tau, sigma = diag(0, 1, 1,
  -2/tau^2)/(3 tau), and for pi proportional to sigma every number has a
  closed form;
- a covariant restatement in Milne coordinates (theta, sigma^{mu nu},
  omega^{mu nu}, I2, I3, I4, J^{mu nu} from nabla_mu u^nu with Christoffel
  symbols) for random states, flows with u^eta included.  This kernel keeps
  an older vorticity convention (HydrodynamicValidity.cu:116-121): it equals
  minus the covariant omega plus terms along the acceleration,
  Delta^tau_eta = tau^2 u^eta (Du^tau - tau (u^eta)^2)/2 + (3/2) u^tau tau^2 Du^eta,
  Delta^x_eta = u^x tau^2 Du^eta, Delta^y_eta = u^y tau^2 Du^eta (the
  difference of the reference's expressions from the covariant ones, written
  with the covariant acceleration Du^mu = u^nu nabla_nu u^mu); states on the
  set where these vanish and general states are both drawn;
- the lattice glue: central differences of linear velocity profiles are the
  slopes, and pi, Pi come from q^n of the last step.

CPU only.
"""
import numpy as np
import pytest

import oracle_lib as O
from test_oracle_pins import HBARC, metric, params, pi_components, transverse_traceless, unit_u

F32 = lambda x: float(np.float32(x))  # noqa: E731  reference float literals
DELTA_PIPI, TAU_PIPI, LAMBDA_PIPI_BULK, DELTA_BULK = 1.33333, 1.42857, 1.2, 0.666667


def outputs(P, t, e, p, theta, pipi, ss, JJ, J, Pi):
    """The ten numbers of a cell from the scalars (HydrodynamicValidity.cu:214-227
    read as formulas: Kn = tau_x theta, Re^-1 = |current| / pressure)."""
    peos, cs2, T = O.eos(P, e)
    zetabar = O.lib().kto_zeta_over_s(P, T)
    a = F32(0.333333) - cs2
    tauPiInv = 15 * a * a * (T / zetabar)
    etabar = P.etabar
    out = np.zeros(O.V_REG)
    out[O.V_RE_PI] = np.sqrt(abs(pipi)) / p
    out[O.V_RE2_PI] = T / (2 * etabar * (e + p)) * np.sqrt(abs(JJ / ss))
    out[O.V_RE_BULK] = abs(Pi) / p
    out[O.V_RE2_BULK] = abs(J / zetabar / theta) * T / (e + p)
    out[O.V_KN_TAUPI] = 5 * etabar * theta / T
    out[O.V_KN_TAUBULK] = theta / tauPiInv
    out[O.V_TAUPI] = 5 * etabar / T
    out[O.V_THETA] = theta
    return out


@pytest.mark.parametrize("eos,bulk", [(O.EOS_WB, True), (O.EOS_TABLE, True),
                                      (O.EOS_CONFORMAL, False)])
def test_validity_bjorken_closed_form(eos, bulk):
    """u = (1,0,0,0) at rest in Milne coordinates, pi = a sigma_hat with
    sigma_hat = diag(0, 1, 1, -2/tau^2): theta = 1/tau, sigma = sigma_hat/(3 tau),
    sigma:sigma = 2/(3 tau^2), pi:pi = 6 a^2, pi:sigma = 2a/tau, the vorticity
    vanishes, I4 = pi/(3 tau), so J^{mu nu} = c sigma_hat with
    c = (3 a delta + a tau_pp - lambda Pi)/(3 tau) and J:J = 6 c^2."""
    P = params(eos, bulk=bulk, etabar=0.16, zeta_tpeak=0.16 / HBARC)
    rng = np.random.default_rng(11)
    for _ in range(40):
        t = rng.uniform(0.3, 5.0)
        e = 10 ** rng.uniform(-0.5, 2)
        p = O.eos(P, e)[0]
        a = rng.uniform(-0.3, 0.3) * p
        Pi = rng.uniform(-0.3, 0.0) * p if bulk else 0.0
        u = np.array([1.0, 0.0, 0.0, 0.0])
        pi10 = pi_components(np.diag([0.0, a, a, -2 * a / t**2]))
        got = O.validity_cell(P, t, e, p, u, u, np.zeros((3, 4)), pi10, Pi)
        c = (3 * a * DELTA_PIPI + a * TAU_PIPI - LAMBDA_PIPI_BULK * Pi) / (3 * t)
        cs2 = O.eos(P, e)[1]
        J = -DELTA_BULK * Pi / t + F32(1.6) * (F32(0.333333) - cs2) * (2 * a / t)
        want = outputs(P, t, e, p, 1.0 / t, 6 * a * a, 2.0 / (3 * t * t), 6 * c * c, J, Pi)
        assert got[O.V_DXUX] == 0.0 and got[O.V_DYUY] == 0.0
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def covariant_validity(P, t, e, p, u, up, du, pi, Pi, dt, older_vorticity=False):
    """The validity numbers from the covariant tensors.  nab[mu, nu] =
    nabla_mu u^nu with Gamma^tau_{eta eta} = tau, Gamma^eta_{tau eta} = 1/tau;
    sigma and omega from the transverse gradient; the kernel's vorticity
    is -omega, plus its terms along the acceleration when older_vorticity."""
    g, gi = metric(t)
    Gam = np.zeros((4, 4, 4))
    Gam[0, 3, 3] = t
    Gam[3, 0, 3] = Gam[3, 3, 0] = 1.0 / t
    d_u = np.zeros((4, 4))
    d_u[0] = (u - up) / dt
    d_u[1:] = du
    nab = d_u + np.einsum("nml,l->mn", Gam, u)
    Du = u @ nab
    theta = np.trace(nab)
    Delta = gi - np.outer(u, u)
    nperp = gi @ nab - np.outer(u, Du)
    sigma = 0.5 * (nperp + nperp.T) - Delta * theta / 3.0
    omega_kernel = -0.5 * (nperp - nperp.T)
    om_mixed = omega_kernel @ g
    if older_vorticity:
        # the kernel's extra terms along the acceleration, as omega^mu_nu
        ut, ux, uy, un = u
        a_eta = t * t * Du[3]
        dlt = np.zeros((4, 4))
        dlt[0, 3] = t * t * un * (Du[0] - t * un * un) / 2 + 1.5 * ut * a_eta
        dlt[1, 3] = ux * a_eta
        dlt[2, 3] = uy * a_eta
        dlt[3, 0] = dlt[0, 3] / (t * t)
        dlt[3, 1] = -dlt[1, 3] / (t * t)
        dlt[3, 2] = -dlt[2, 3] / (t * t)
        om_mixed = om_mixed + dlt
    sg_mixed = sigma @ g
    I2 = theta * pi
    I3 = pi @ om_mixed.T + (pi @ om_mixed.T).T
    pisig = np.einsum("ab,ab", pi, g @ sigma @ g)
    I4 = 0.5 * (pi @ sg_mixed.T + (pi @ sg_mixed.T).T) - Delta * pisig / 3.0
    Jmn = DELTA_PIPI * I2 - I3 + TAU_PIPI * I4 - LAMBDA_PIPI_BULK * Pi * sigma
    contract = lambda A, B: np.einsum("ab,ab", g @ A @ g, B)  # noqa: E731
    cs2 = O.eos(P, e)[1]
    J = -DELTA_BULK * Pi * theta + F32(1.6) * (F32(0.333333) - cs2) * pisig
    out = outputs(P, t, e, p, theta, contract(pi, pi), contract(sigma, sigma),
                  contract(Jmn, Jmn), J, Pi)
    out[O.V_DXUX] = du[0][1]
    out[O.V_DYUY] = du[1][2]
    return out


@pytest.mark.parametrize("eos,bulk", [(O.EOS_WB, True), (O.EOS_TABLE, True),
                                      (O.EOS_CONFORMAL, False)])
def test_validity_covariant(eos, bulk):
    """kto_validity_cell == the covariant restatement for random normalised
    u (u^eta included), random spatial gradients, transverse traceless pi,
    and time derivatives fixed by u.d u^tau = tau (u^eta)^2 and
    D u^eta = 0 (where the kernel's vorticity is -omega)."""
    dt = 0.01
    P = params(eos, bulk=bulk, etabar=0.16, zeta_tpeak=0.16 / HBARC, dt=dt)
    rng = np.random.default_rng(12)
    for _ in range(100):
        t = rng.uniform(0.4, 4.0)
        u = unit_u(t, rng.uniform(-0.7, 0.7), rng.uniform(-0.7, 0.7), rng.uniform(-0.4, 0.4) / t)
        ut, ux, uy, un = u
        du = rng.standard_normal((3, 4)) * np.array([1.0, 1.0, 1.0, 1.0 / t])
        dtu = np.zeros(4)
        dtu[1:3] = rng.standard_normal(2)
        # u^mu d_mu u^tau (partial derivatives) = 0 and D u^eta = 0, i.e.
        # u^mu d_mu u^eta = -2 u^tau u^eta / tau
        dtu[0] = -(ux * du[0, 0] + uy * du[1, 0] + un * du[2, 0]) / ut
        dtu[3] = (-2 * ut * un / t - ux * du[0, 3] - uy * du[1, 3] - un * du[2, 3]) / ut
        up = u - dt * dtu
        e = 10 ** rng.uniform(0, 2)
        p = O.eos(P, e)[0]
        Pi = -0.2 * p * rng.uniform() if bulk else 0.0
        pi = transverse_traceless(t, u, rng, 0.2 * e)
        want = covariant_validity(P, t, e, p, u, up, du, pi, Pi, dt)
        got = O.validity_cell(P, t, e, p, u, up, du, pi_components(pi), Pi)
        np.testing.assert_allclose(got, want, rtol=1e-8, atol=0)


@pytest.mark.parametrize("eos,bulk", [(O.EOS_WB, True), (O.EOS_TABLE, True),
                                      (O.EOS_CONFORMAL, False)])
def test_validity_older_vorticity_general_flow(eos, bulk):
    """kto_validity_cell == the covariant restatement with the older
    vorticity's acceleration terms, for random normalised u (u^eta included),
    random spatial gradients and random time derivatives (so Du^tau and
    Du^eta are arbitrary); transverse traceless pi.  The restatement without
    those terms must miss (the states resolve them)."""
    dt = 0.01
    P = params(eos, bulk=bulk, etabar=0.16, zeta_tpeak=0.16 / HBARC, dt=dt)
    rng = np.random.default_rng(13)
    resolved = 0
    for _ in range(100):
        t = rng.uniform(0.4, 4.0)
        u = unit_u(t, rng.uniform(-0.7, 0.7), rng.uniform(-0.7, 0.7), rng.uniform(-0.4, 0.4) / t)
        du = rng.standard_normal((3, 4)) * np.array([1.0, 1.0, 1.0, 1.0 / t])
        dtu = rng.standard_normal(4) * np.array([1.0, 1.0, 1.0, 1.0 / t])
        up = u - dt * dtu
        e = 10 ** rng.uniform(0, 2)
        p = O.eos(P, e)[0]
        Pi = -0.2 * p * rng.uniform() if bulk else 0.0
        pi = transverse_traceless(t, u, rng, 0.2 * e)
        want = covariant_validity(P, t, e, p, u, up, du, pi, Pi, dt, older_vorticity=True)
        got = O.validity_cell(P, t, e, p, u, up, du, pi_components(pi), Pi)
        np.testing.assert_allclose(got, want, rtol=1e-8, atol=0)
        plain = covariant_validity(P, t, e, p, u, up, du, pi, Pi, dt)
        if abs(plain[O.V_RE2_PI] / got[O.V_RE2_PI] - 1) > 1e-6:
            resolved += 1
    assert resolved > 90


def test_validity_grid_differences_and_buffers():
    """On the lattice: linear profiles u^x = a x, u^y = b y, u^eta = c eta
    (u^tau normalised) give dxux = a, dyuy = b and theta = u^tau/tau + a + b + c
    at cells whose stencil stays off the ghost cells; after a step, pi and Pi
    come from the state before the step and e, p, u from the state after it."""
    nx, ny, nz, h = 9, 8, 7, 0.1
    P = O.make_params(nx, ny, nz, h, h, h, 0.01, 0.16, True, O.EOS_WB, zeta_tpeak=0.16 / HBARC)
    orc = O.Oracle(P)
    x = (np.arange(nx) - (nx - 1) / 2) * h
    y = (np.arange(ny) - (ny - 1) / 2) * h
    z = (np.arange(nz) - (nz - 1) / 2) * h
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a, b, c, t0 = 0.3, -0.2, 0.1, 0.8
    ux, uy, un = a * X, b * Y, c * Z
    ut = np.sqrt(1 + ux**2 + uy**2 + t0 * t0 * un**2)
    e = 5.0 + X + 0.5 * Y
    p = O.eos_p_batch(P, e)
    rng = np.random.default_rng(3)
    pi10 = [0.01 * rng.standard_normal((nx, ny, nz)) for _ in range(10)]
    Pi = -0.05 * p
    orc.set_primitives(t0, e, p, ut, ux, uy, un, pi10, Pi)
    v = orc.validity(t0, prev_q=0)
    core = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(v[O.V_DXUX][core], a, rtol=1e-12)
    np.testing.assert_allclose(v[O.V_DYUY][core], b, rtol=1e-12)
    np.testing.assert_allclose(v[O.V_THETA][core], (ut / t0 + a + b + c)[core], rtol=1e-12)
    np.testing.assert_array_equal(v[O.V_REG], 1.0)  # DynamicalVariables.cu:81
    np.testing.assert_allclose(v[O.V_RE_BULK], np.abs(Pi) / p, rtol=1e-15)

    pi_before = [orc.get(f) for f in range(O.PITT, O.PINN + 1)]
    Pi_before = orc.get(O.PI)
    orc.step(t0)
    t1 = t0 + 0.01
    v = orc.validity(t1, prev_q=1)
    p1 = orc.get(O.P)
    gmet = np.array([1.0, -1.0, -1.0, -t1 * t1])
    idx = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    pipi = sum((1 if m == n else 2) * gmet[m] * gmet[n] * pi_before[k] ** 2
               for k, (m, n) in enumerate(idx))
    np.testing.assert_allclose(v[O.V_RE_PI], np.sqrt(np.abs(pipi)) / p1, rtol=1e-12)
    np.testing.assert_allclose(v[O.V_RE_BULK], np.abs(Pi_before) / p1, rtol=1e-15)
    v_now = orc.validity(t1, prev_q=0)
    np.testing.assert_allclose(v_now[O.V_RE_BULK], np.abs(orc.get(O.PI)) / p1, rtol=1e-15)
    # the step's regulation factors are kept (1 where no cell is regulated)
    assert np.all(v[O.V_REG] <= 1.0)
    orc.close()
