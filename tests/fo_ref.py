"""Restatement of the reference run loop's freeze-out glue and loop control
over the oracle (test infrastructure: the checker of gvh_freezeout_* and
gvh_run).

Follows rhic/src/HydroPlugin.cpp:107-434 and freezeout/FreezeOut.cpp:
  * history ring of FOFREQ+1 slices, slice 0 <- slice FOFREQ when
    n % FOFREQ == 0, the current state into slice n % FOFREQ + 1
    (HydroPlugin.cpp:260-269, FreezeOut.cpp:49-126);
  * window search when n % FOFREQ == FOFREQ - 1 over it = start..FOFREQ-1,
    start = 2 for n <= FOFREQ else 0, cubes in (it, ix, iy, iz) order
    (HydroPlugin.cpp:270-299);
  * rows: centroid + cell position, normal times the current t, multilinear
    interpolation of the ten variables at the centroid, T and P of the
    interpolated e (HydroPlugin.cpp:300-386, FreezeOut.cpp:5-47, 264-280);
  * cell tau t0 + (n - FOFREQ + it - 1) dt for n <= FOFREQ, else
    t0 + (n - FOFREQ + it) dt (HydroPlugin.cpp:304-305);
  * the host copies of u and pi are the previous step's after an odd number
    of steps (DynamicalVariables.cu:310-331 reads the first-allocated buffers,
    FullyDiscreteKurganovTadmorScheme.cu:132 and DynamicalVariables.cu:400-402
    swap the roles every step); e and p are current;
  * termination: the count of steps with no interior cell above e_fo is never
    reset; stop once it reaches FOFREQ+1, before stepping (:396-413); the loop
    runs n = 1..nt+1 (:239).
"""
import numpy as np

import cornelius_lib as CL
import oracle_lib as O

VARS = ["ux", "uy", "un", "e", "pixx", "pixy", "pixn", "piyy", "piyn", "Pi"]
VAR_FIELDS = [O.UX, O.UY, O.UN, O.E, O.PIXX, O.PIXY, O.PIXN, O.PIYY, O.PIYN, O.PI]
U_VARS = (0, 1, 2)
PI_VARS = (4, 5, 6, 7, 8, 9)

# FreezeOut.cpp:12-27 and 37-44: terms of the multilinear interpolation in
# the reference's order, bits (tau, x, y[, eta])
TERMS4 = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (1, 1, 0, 0),
          (1, 0, 1, 0), (1, 0, 0, 1), (0, 1, 1, 0), (0, 1, 0, 1), (0, 0, 1, 1), (1, 1, 1, 0),
          (1, 1, 0, 1), (1, 0, 1, 1), (0, 1, 1, 1), (1, 1, 1, 1)]
TERMS3 = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]

FO_EXACT_TAU, FO_NORMAL_TAU, FO_CURRENT_FIELDS = 1, 2, 4


def interp(fr, corner, terms):
    """sum over terms of prod_k (x_k or 1 - x_k) * a, products and sum left
    to right.  corner(bits) gives the value at that corner."""
    r = None
    for bits in terms:
        w = None
        for x, b in zip(fr, bits):
            f = x if b else (1 - x)
            w = f if w is None else w * f
        term = w * corner(bits)
        r = term if r is None else r + term
    return r


class Ring:
    """FOFREQ+1 slices of e and the ten variables, arrays (nx, ny, nz)."""

    def __init__(self, fofreq):
        self.fofreq = fofreq
        self.slots = [None] * (fofreq + 1)

    def record(self, n, fields):
        nfo = n % self.fofreq
        if nfo == 0:
            self.slots[0] = self.slots[self.fofreq]
            self.slots[1] = fields
        else:
            self.slots[nfo + 1] = fields


def window_rows(ring, n, t, geom, e_fo, eos_T_p, flags=0):
    """Rows of the window searched at loop index n and current time t.
    geom: dict nx, ny, nz, dt, dx, dy, dz, t0, dim."""
    fofreq = ring.fofreq
    nx, ny, nz, dim = geom["nx"], geom["ny"], geom["nz"], geom["dim"]
    dt, dx, dy, dz, t0 = geom["dt"], geom["dx"], geom["dy"], geom["dz"], geom["t0"]
    start = 2 if n <= fofreq else 0
    dimz = nz - 1 if dim == 4 else 1
    spacing = [dt, dx, dy, dz][:dim]
    E = np.stack([ring.slots[it][3] for it in range(start, fofreq + 1)])  # (nit+1, nx, ny, nz)
    # surface-cube predicate over the window, vectorised
    above = np.zeros((fofreq - start, nx - 1, ny - 1, dimz), dtype=np.int32)
    zs = [0, 1] if dim == 4 else [0]
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                for d in zs:
                    blk = E[a:a + fofreq - start, b:b + nx - 1, c:c + ny - 1, d:d + dimz]
                    above += (blk >= e_fo)
    ncorner = 1 << dim
    hit = np.argwhere((above > 0) & (above < ncorner))
    rows = []
    for itl, ix, iy, iz in hit:
        it = start + itl
        if dim == 4:
            corners = [ring.slots[it + a][3][ix + b, iy + c, iz + d]
                       for a in (0, 1) for b in (0, 1) for c in (0, 1) for d in (0, 1)]
        else:
            corners = [ring.slots[it + a][3][ix + b, iy + c, 0]
                       for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        cen, nrm = CL.find(dim, corners, spacing, e_fo)
        if n <= fofreq or (flags & FO_EXACT_TAU):
            cell_tau = t0 + float(n - fofreq + (it - 1)) * dt
        else:
            cell_tau = t0 + float(n - fofreq + it) * dt
        cell_x = float(ix) * dx - (float(nx - 1) / 2.0 * dx)
        cell_y = float(iy) * dy - (float(ny - 1) / 2.0 * dy)
        cell_z = float(iz) * dz - (float(nz - 1) / 2.0 * dz)
        for i in range(len(cen)):
            c = [float(v) for v in cen[i]]
            m = [float(v) for v in nrm[i]]
            fr = [c[k] / spacing[k] for k in range(dim)]
            row = [c[0] + cell_tau, c[1] + cell_x, c[2] + cell_y,
                   c[3] + cell_z if dim == 4 else cell_z]
            scale = row[0] if (flags & FO_NORMAL_TAU) else t
            row += [scale * m[0], scale * m[1], scale * m[2],
                    scale * m[3] if dim == 4 else 0.0]
            vals = []
            for v in range(10):
                if dim == 4:
                    vals.append(interp(fr, lambda bt: ring.slots[it + bt[0]][v][
                        ix + bt[1], iy + bt[2], iz + bt[3]], TERMS4))
                else:
                    vals.append(interp(fr, lambda bt: ring.slots[it + bt[0]][v][
                        ix + bt[1], iy + bt[2], 0], TERMS3))
            T, P = eos_T_p(vals[3])
            row += vals[0:3] + [vals[3], T, P] + vals[4:10]
            rows.append(row)
    return np.array(rows, dtype=np.float64).reshape(-1, 20)


def state_fields(orc, bulk):
    """Current (u^x, u^y, u^eta, e, pi^xx .. pi^yeta, Pi) of the oracle."""
    out = [orc.get(f) for f in VAR_FIELDS]
    if not bulk:
        out[9] = np.zeros_like(out[9])
    return out


def host_view(cur, prev, steps_done, flags):
    """The ten variables as the reference's host copy sees them."""
    if (flags & FO_CURRENT_FIELDS) or steps_done % 2 == 0 or prev is None:
        return list(cur)
    view = list(cur)
    for v in U_VARS + PI_VARS:
        view[v] = prev[v]
    return view


def run_loop(orc, geom, nt, e_fo, eos_T_p, bulk, fofreq=10, freq=10, flags=0, surface=True,
             snapshot_fields=(), validity=False):
    """The run loop over an oracle whose initial state is set.  Returns
    (rows, steps, frozen, snapshots) with snapshots {t: {field id: array}}
    of the listed oracle field ids (u and pi as the host copy sees them);
    with validity, snapshots[t]["validity"] holds the validity arrays as the
    reference's last checkValidity left them (HydroPlugin.cpp:222,
    FullyDiscreteKurganovTadmorScheme.cu:137)."""
    t0, dt = geom["t0"], geom["dt"]
    ring = Ring(fofreq)
    rows = []
    snaps = {}
    t = t0
    steps = 0
    frozen_count = 0
    frozen = False
    prev = None
    prev_full = None
    for n in range(1, nt + 2):
        cur = state_fields(orc, bulk)
        if (n - 1) % freq == 0 and snapshot_fields:
            stale = not (flags & FO_CURRENT_FIELDS) and steps % 2 == 1
            snaps[t] = {}
            for f in snapshot_fields:
                is_u_pi = f in (O.UT, O.UX, O.UY, O.UN) or O.PITT <= f <= O.PI
                snaps[t][f] = prev_full[f] if (stale and is_u_pi) else orc.get(f)
            if validity:
                snaps[t]["validity"] = orc.validity(t, prev_q=steps > 0)
        if surface:
            ring.record(n, host_view(cur, prev, steps, flags))
            if n % fofreq == fofreq - 1:
                rows.append(window_rows(ring, n, t, geom, e_fo, eos_T_p, flags))
        if not (cur[3] > e_fo).any():
            frozen_count += 1
        if frozen_count >= fofreq + 1:
            frozen = True
            break
        prev = cur
        if snapshot_fields:
            prev_full = {f: orc.get(f) for f in snapshot_fields}
        orc.step(t)
        steps += 1
        t = t0 + n * dt
    rows = np.concatenate(rows) if rows else np.zeros((0, 20))
    return rows, steps, frozen, snaps
