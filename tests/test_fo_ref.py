"""Pins of the freeze-out glue restatement (tests/fo_ref.py) by closed forms.
CPU only."""
import numpy as np
import pytest

import fo_ref as F


def test_interp_reproduces_multilinear_fields():
    """The interpolation is exact on fields a + b t + c x + ... + d t x y z
    (its own function space) and returns corner values at corners."""
    rng = np.random.default_rng(5)
    for dim, terms in ((4, F.TERMS4), (3, F.TERMS3)):
        coef = rng.standard_normal(1 << dim)

        def f(p):
            # multilinear: sum over subsets S of coef[S] prod_{k in S} p_k
            s = 0.0
            for m in range(1 << dim):
                w = coef[m]
                for k in range(dim):
                    if (m >> k) & 1:
                        w *= p[k]
                s += w
            return s

        for _ in range(50):
            fr = rng.uniform(0, 1, dim)
            got = F.interp(fr, lambda bits: f(bits), terms)
            assert got == pytest.approx(f(fr), rel=1e-12, abs=1e-12)
        for bits in terms:
            assert F.interp(list(map(float, bits)), lambda b: f(b), terms) == pytest.approx(
                f(bits), abs=1e-15)


def _uniform_ring(fofreq, dims, E_of_slot, consts):
    ring = F.Ring(fofreq)
    for k in range(fofreq + 1):
        arrs = [np.full(dims, c, dtype=np.float64) for c in consts]
        arrs[3] = np.full(dims, E_of_slot(k))
        ring.slots[k] = arrs
    return ring


@pytest.mark.parametrize("dim", [3, 4])
@pytest.mark.parametrize("n,flags", [(9, 0), (29, 0), (29, F.FO_EXACT_TAU), (29, F.FO_NORMAL_TAU)])
def test_uniform_decay_window(dim, n, flags):
    """e uniform in space, falling linearly through e_fo between two slices:
    every cube of that slab is one element at the crossing time with normal
    (t dx dy deta, 0, 0, 0); interpolated variables are the constants and
    interpolated e is e_fo."""
    fofreq = 10
    nx, ny, nz = (5, 4, 3) if dim == 4 else (5, 4, 1)
    g = dict(nx=nx, ny=ny, nz=nz, dt=0.05, dx=0.2, dy=0.3, dz=0.25 if dim == 4 else 0.1,
             t0=0.6, dim=dim)
    E0, dE = 10.0, 0.7
    e_fo = E0 - 5.3 * dE        # crossing between slices 5 and 6
    consts = [0.1, -0.2, 0.03, 0.0, 1.5, -0.5, 0.25, 2.0, -1.0, 0.4]
    ring = _uniform_ring(fofreq, (nx, ny, nz), lambda k: E0 - k * dE, consts)
    t = 1.234
    rows = F.window_rows(ring, n, t, g, e_fo, lambda e: (2.0 * e, 3.0 * e), flags)
    ncube = (nx - 1) * (ny - 1) * ((nz - 1) if dim == 4 else 1)
    assert rows.shape == (ncube, 20)
    it = 5
    if n <= fofreq or flags & F.FO_EXACT_TAU:
        cell_tau = g["t0"] + (n - fofreq + it - 1) * g["dt"]
    else:
        cell_tau = g["t0"] + (n - fofreq + it) * g["dt"]
    tau = cell_tau + 0.3 * g["dt"]
    vol = g["dx"] * g["dy"] * (g["dz"] if dim == 4 else 1.0)
    scale = tau if flags & F.FO_NORMAL_TAU else t
    k = 0
    for ix in range(nx - 1):
        for iy in range(ny - 1):
            for iz in range((nz - 1) if dim == 4 else 1):
                r = rows[k]
                k += 1
                assert r[0] == pytest.approx(tau, abs=1e-12)
                assert r[1] == pytest.approx(ix * g["dx"] - (nx - 1) / 2 * g["dx"] + g["dx"] / 2)
                assert r[2] == pytest.approx(iy * g["dy"] - (ny - 1) / 2 * g["dy"] + g["dy"] / 2)
                if dim == 4:
                    assert r[3] == pytest.approx(iz * g["dz"] - (nz - 1) / 2 * g["dz"] + g["dz"] / 2)
                else:
                    assert r[3] == 0.0
                assert r[4] == pytest.approx(scale * vol, rel=1e-12)
                assert np.abs(r[5:8]).max() < 1e-14
                assert r[8:11] == pytest.approx(consts[0:3], abs=1e-14)
                assert r[11] == pytest.approx(e_fo, rel=1e-12)
                assert r[12] == pytest.approx(2.0 * r[11], rel=1e-15)
                assert r[13] == pytest.approx(3.0 * r[11], rel=1e-15)
                assert r[14:20] == pytest.approx(consts[4:10], abs=1e-14)


def test_ring_slots_follow_the_reference_indexing():
    """Slice k of the window searched at n = 10m + 9 holds the state after
    10m + k - 2 steps (k >= 1), slice 0 the one after 10m - 2."""
    ring = F.Ring(10)
    for n in range(1, 40):
        ring.record(n, [None, None, None, n - 1] + [None] * 6)  # state index = steps done
        if n % 10 == 9:
            m = n // 10
            for k in range(2 if m == 0 else 0, 11):
                want = 10 * m + k - 2
                assert ring.slots[k][3] == want, (n, k)


def test_host_view_staleness():
    cur = [np.full(2, 1.0)] * 10
    prev = [np.full(2, 0.0)] * 10
    v = F.host_view(cur, prev, 3, 0)
    assert all(v[i][0] == 0.0 for i in F.U_VARS + F.PI_VARS) and v[3][0] == 1.0
    assert all(x[0] == 1.0 for x in F.host_view(cur, prev, 4, 0))
    assert all(x[0] == 1.0 for x in F.host_view(cur, prev, 3, F.FO_CURRENT_FIELDS))
