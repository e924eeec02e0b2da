"""Pins of the oracle's Newton inversion (energy_from_conserved, restating
EnergyMomentumTensor.cu:19-48) and of its regulation factor
(kto_regulate_factor, RegulateDissipativeCurrents.cu:69-112) by what the
mathematics fixes.

Newton: the update solves f(e) = e + D(e) = 0 with
D = (A - H)/(2 c_t), c_t = p(e0)/e0, A = M0 (1 - c_t) + Pi,
B = M0 (M0 + Pi) - M, H = sqrt|A^2 + 4 c_t B|, i.e. D is the smaller root
of c_t D^2 - A D - B = 0, and the physical solution is the fixed point
e = M0 - M/(M0 + p(e) + Pi).  Equations of state built as tables with
exactly linear or affine p(e) make these checks independent of the
reference's fitted functions:

- p = c e (c_s^2 = c on every node): c_t = c does not depend on e0, so one
  update lands on the root of c e^2 + A e - B = 0 (closed form) and the
  next meets the stop rule |de| <= 1e-3 |e|;
- p = c e + p0 (c_s^2 = c): the reference's derivative
  fp = 1 - d (B + D H - d c_t D M0/e0)/(c_t e0 H), d = c_s^2 - c_t, is
  rewritten with the root identity B + D H = -c_t D^2 as
  fp = 1 + d D (D + d M0/e0)/(e0 H) and iterated independently; the exact
  derivative of f would be 1 + d D (D + M0)/(e0 H), which converges to a
  visibly different stopping iterate, so a Jacobian slip fails the pin.

CPU only.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

import oracle_lib as O
from test_oracle_pins import metric, pi_components, transverse_traceless, unit_u


class LinearTable:
    """Tabulated EoS with p = c e + p0 at every node and c_s^2 = c, on the
    pseudo-log grid of the library's tables (nodes 2^E (1 + j/M))."""

    def __init__(self, c, p0=0.0, per_octave=16, emin_exp=-12, octaves=24):
        self.per_octave, self.emin_exp = per_octave, emin_exp
        self.n = per_octave * octaves + 1
        k = np.arange(self.n)
        e = np.ldexp(1.0 + (k % per_octave) / per_octave, emin_exp + k // per_octave)
        self.p = c * e + p0
        self.cs2 = np.full(self.n, float(c))
        self.T = (e / 13.8997) ** 0.25


def table_params(tab, bulk=True):
    return O.make_params(1, 1, 1, 0.1, 0.1, 0.1, 0.01, 0.2, bulk, O.EOS_TABLE, table=tab)


def conserved_inputs(rng, with_pi=True):
    """Random M0, M, Pi of a physical state (e, u, pi, Pi) and the state's q."""
    t = rng.uniform(0.4, 3.0)
    u = unit_u(t, *rng.uniform(-0.8, 0.8, 2), rng.uniform(-0.4, 0.4) / t)
    e = 10 ** rng.uniform(-0.3, 2.5)
    return t, u, e


def q_of_state(t, e, P_total, u, pi, Pi):
    g, gi = metric(t)
    T = (e + P_total) * np.outer(u, u) - P_total * gi + pi
    q = np.zeros(15)
    q[0:4] = T[0, :]
    q[4:14] = pi_components(pi)
    q[14] = Pi
    return q


def m_values(t, q):
    M0 = q[0] - q[4]
    M = (q[1] - q[5]) ** 2 + (q[2] - q[6]) ** 2 + t * t * (q[3] - q[7]) ** 2
    return M0, M, q[14]


def test_newton_linear_eos_closed_form():
    """p = c e: the oracle returns the closed-form root
    e* = 2B/(A + sqrt(A^2 + 4 c B)) of c e^2 + A e - B = 0 (the at-rest
    energy density of the state it was built from), after exactly two
    iterations when the guess is off by more than the stop tolerance and
    one when it is within it."""
    rng = np.random.default_rng(21)
    for c in (0.1, 1.0 / 3.0, 0.2871):
        P = table_params(LinearTable(c))
        for _ in range(200):
            t, u, e = conserved_inputs(rng)
            Pi = -0.3 * c * e * rng.uniform()
            pi = transverse_traceless(t, u, rng, 0.1 * e)
            q = q_of_state(t, e, c * e + Pi, u, pi, Pi)
            M0, M, Pi_ = m_values(t, q)
            A = M0 * (1 - c) + Pi_
            B = M0 * (M0 + Pi_) - M
            root = 2 * B / (A + np.sqrt(A * A + 4 * c * B))
            assert root == pytest.approx(e, rel=1e-9)  # the state's own e
            far = e * rng.choice([0.5, 0.8, 1.3, 2.0])
            out = O.inferred(P, t, q, far)
            assert out[0] == pytest.approx(root, rel=1e-13)
            assert O.lib().kto_last_newton_iters() == 2
            near = root * (1 + 2e-4 * rng.uniform(-1, 1))
            out = O.inferred(P, t, q, near)
            assert out[0] == pytest.approx(root, rel=1e-13)
            assert O.lib().kto_last_newton_iters() == 1
            # the inferred velocity follows: u^mu of the state
            np.testing.assert_allclose(out[2:6], u, rtol=1e-9, atol=1e-11)


def restated_newton(M0, M, Pi, e0, peos, cs2eos, exact_derivative=False, max_iter=200):
    """Independent iteration: D from the rationalised root -2B/(A + H),
    fp from the root identity (or the exact df/de for the planted-slip
    comparison); stop rule |e - e0| <= 1e-3 |e| returning the new iterate."""
    for j in range(max_iter):
        ct = peos(e0) / e0
        d = cs2eos(e0) - ct
        A = M0 * (1 - ct) + Pi
        B = M0 * (M0 + Pi) - M
        H = np.sqrt(abs(A * A + 4 * ct * B))
        D = -2 * B / (A + H)
        f = e0 + D
        if exact_derivative:
            fp = 1 + d * D * (D + M0) / (e0 * H)
        else:
            fp = 1 + d * D * (D + d * M0 / e0) / (e0 * H)
        e = e0 - f / fp
        if abs(e - e0) <= 1e-3 * abs(e):
            return e, j + 1
        e0 = e
    raise AssertionError("no convergence")


def test_newton_affine_eos_iterates():
    """p = c e + p0 with c_s^2 = c (= dp/de exactly): the oracle's stopping
    iterate and iteration count equal those of the independent iteration
    with the reference's Jacobian in its root-identity form; the result is
    within the linear-convergence bound of the true fixed point
    e = M0 - M/(M0 + p(e) + Pi); and the exact-Jacobian iteration stops
    elsewhere (so the pin resolves the Jacobian)."""
    rng = np.random.default_rng(22)
    distinct = 0
    for c, p0 in ((0.25, -0.3), (1.0 / 3.0, -1.0), (0.18, 0.4)):
        tab = LinearTable(c, p0)
        P = table_params(tab)
        peos = lambda e, c=c, p0=p0: c * e + p0  # noqa: E731
        cs2eos = lambda e, c=c: c  # noqa: E731
        for _ in range(150):
            t, u, e = conserved_inputs(rng)
            e = max(e, 5.0 * abs(p0) / c)  # keep p > 0 in the range used
            Pi = -0.2 * peos(e) * rng.uniform()
            pi = transverse_traceless(t, u, rng, 0.05 * e)
            q = q_of_state(t, e, peos(e) + Pi, u, pi, Pi)
            M0, M, Pi_ = m_values(t, q)
            guess = e * rng.uniform(0.85, 1.15)
            out = O.inferred(P, t, q, guess)
            iters = O.lib().kto_last_newton_iters()
            want, n_want = restated_newton(M0, M, Pi_, guess, peos, cs2eos)
            assert iters == n_want
            assert out[0] == pytest.approx(want, rel=1e-12)
            # fixed point of the physical equation
            g = lambda x: x - (M0 - M / (M0 + peos(x) + Pi_))  # noqa: E731
            star = brentq(g, 0.5 * e, 2.0 * e, xtol=1e-15 * e)
            assert star == pytest.approx(e, rel=1e-9)
            assert abs(out[0] - star) <= 2e-3 * star
            exact, _ = restated_newton(M0, M, Pi_, guess, peos, cs2eos, exact_derivative=True)
            if abs(exact - out[0]) > 1e-9 * star:
                distinct += 1
    assert distinct > 300  # the Jacobian form is resolved in most cases


# ------------------------------------------------------------ regulation
def regulation_factor_covariant(t, pi, e, p, u):
    """rho = max(sqrt|pi:pi| / sqrt(e^2 + 3 p^2),
    g_{mu nu} pi^{mu nu} / (xi0 sqrt|pi:pi|),
    (pi^{mu nu} g_{nu l} u^l) / (xi0 sqrt|pi:pi|) for mu = tau, x, y, eta),
    xi0 = 0.1, rho_max = 1, signed (not absolute) ratios; factor
    tanh(rho)/rho (1 for |rho| < 1e-7).  Contractions with the Milne metric
    diag(1, -1, -1, -tau^2) in matrix form."""
    g, _ = metric(t)
    pipi = np.einsum("ab,ab", g @ pi @ g, pi)
    s = np.sqrt(abs(pipi))
    tr = np.einsum("ab,ab", g, pi)
    piu = pi @ g @ u
    cands = [s / np.sqrt(e * e + 3 * p * p), tr / (0.1 * s)] + list(piu / (0.1 * s))
    rho = max(cands)
    return (1.0 if abs(rho) < 1e-7 else np.tanh(rho) / rho), int(np.argmax(cands))


def test_regulation_factor_general_pi():
    """kto_regulate_factor on random symmetric pi that is neither traceless
    nor transverse: every one of the six ratios is the maximum in some
    draws, and the factor equals the covariant form; above e_reg it is 1."""
    rng = np.random.default_rng(23)
    winners = set()
    for _ in range(3000):
        t = rng.uniform(0.4, 3.0)
        u = unit_u(t, *rng.uniform(-0.8, 0.8, 2), rng.uniform(-0.4, 0.4) / t)
        e = 10 ** rng.uniform(-1, 1.5)
        p = e / 3 * rng.uniform(0.5, 1.0)
        base = transverse_traceless(t, u, rng, rng.uniform(0.01, 1.0) * e)
        A = rng.standard_normal((4, 4))
        A = 0.5 * (A + A.T) * e * 10 ** rng.uniform(-4, -1)
        A[:, 3] /= t
        A[3, :] /= t
        pi = base + A
        want, w = regulation_factor_covariant(t, pi, e, p, u)
        winners.add(w)
        got = O.regulate_factor(t, pi_components(pi), e, p, u, 2 * e)
        assert got == pytest.approx(want, rel=1e-11, abs=1e-15)
        assert O.regulate_factor(t, pi_components(pi), e, p, u, 0.5 * e) == 1.0
    assert winners == set(range(6)), winners


def test_newton_damped_branch():
    """The damping of EnergyMomentumTensor.cu:35-37 (step x 0.9 from
    iteration MAX_ITERS/2 = 500000 on).  A fluid at rest (M = 0, Pi = 0)
    with p = c e and a c_s^2 column linear in e, c_s^2 = c + beta e, makes
    the update exactly linear: D = -M0, f = e0 - M0 and
    fp = 1 + beta (1 - beta) M0^2 / H with H = M0 (1 + c), a constant.
    beta is chosen so that fp = 1/2: every undamped step maps
    e - M0 -> -(e - M0), a 2-cycle that never meets the stop rule; from
    iteration 500000 on the damped map e - M0 -> (1 - 0.9/fp)(e - M0)
    = -0.8 (e - M0) converges.  So the oracle stops after 500000 + m
    iterations at M0 + d0 (-0.8)^m, m the first damped step with
    |de| = 1.8 |d| <= 1e-3 |e|."""
    c, M0 = 0.3, 10.0
    H = M0 * (1 + c)
    phi = 0.5
    beta = (1 - np.sqrt(1 + 4 * (1 - phi) * H / M0 ** 2)) / 2  # beta (1 - beta) = (phi-1) H/M0^2
    tab = LinearTable(c)
    e_nodes = tab.p / c
    tab.cs2 = c + beta * e_nodes
    P = table_params(tab, bulk=False)
    q = np.zeros(15)
    q[0] = M0
    for d0 in (0.5, -0.5, 0.8):
        # even number of undamped steps: the cycle is back at d0 at 500000
        m, d = 0, d0
        while True:
            d_new = -0.8 * d
            m += 1
            if abs(d_new - d) <= 1e-3 * abs(M0 + d_new):
                break
            d = d_new
        out = O.inferred(P, 1.0, q, M0 + d0)
        assert O.lib().kto_last_newton_iters() == 500000 + m
        assert out[0] == pytest.approx(M0 + d_new, rel=1e-12)
