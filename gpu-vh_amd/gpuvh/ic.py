"""Seeded synthetic initial conditions on the reference's lattice.

The reference reads file initial conditions (type 10: e.dat plus flow and
shear files, rhic/src/InitialConditions.cpp:98-489): each value is parsed
with "%f" into a float, 1e-3 is added to e (:124) and p is set from the EoS
(:127).  ``quantise_like_reference`` reproduces exactly that, so a field
generated here and written to e.dat gives the same lattice state as feeding
the reference the file.

Pseudo-random numbers come from SplitMix64 (Steele, Lea & Flood 2014),
seeded per event, so every event is reproducible from (seed, event id).
"""
import numpy as np

HBARC = 0.197326938  # GeV fm, rhic/src/HydroPlugin.cpp:132

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        """Box-Muller, consuming two uniforms."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def lattice_coords(n, h):
    """Cell-centre coordinates of an odd, centred axis (x_i = (i - (n-1)/2) h)."""
    return (np.arange(n) - (n - 1) / 2.0) * h


def longitudinal_profile(nz, dz, eta_flat=5.9, eta_variance=0.16):
    """Plateau plus Gaussian tails in eta, as
    longitudinalEnergyDensityDistribution (InitialConditions.cpp:731-752);
    eta uses the reference's integer (nz-1)/2."""
    k = np.arange(nz)
    eta = (k - (nz - 1) // 2) * dz
    scaled = np.abs(eta) - eta_flat / 2
    arg = -scaled * scaled / eta_variance / 2 * (scaled >= 0)
    return np.exp(arg)


def quantise_like_reference(e_fm4):
    """float parse of the IC file plus 1e-3 (InitialConditions.cpp:122-124)."""
    return np.asarray(e_fm4, dtype=np.float32).astype(np.float64) + 1.0e-3


def hotspot_event_2d(nx, ny, dx, dy, seed, n_hot, hot_sigma, hot_amp_max,
                     base_peak=0.0, base_sigma=1.0, centre_radius=3.0):
    """Transverse energy density [fm^-4] on an nx*ny lattice.

    e(x,y) [GeV/fm^3] = base_peak exp(-r^2/(2 base_sigma^2))
                        + sum_k A_k exp(-|x - x_k|^2 / (2 hot_sigma^2)),
    A_k ~ U(0, hot_amp_max), centres uniform in the disc r < centre_radius.
    Draw order per hotspot: radius uniform, angle uniform, amplitude.
    """
    rng = SplitMix64(seed)
    x = lattice_coords(nx, dx)[:, None]
    y = lattice_coords(ny, dy)[None, :]
    e = base_peak * np.exp(-(x * x + y * y) / (2.0 * base_sigma ** 2))
    for _ in range(n_hot):
        r = centre_radius * np.sqrt(rng.uniform())
        phi = 2.0 * np.pi * rng.uniform()
        amp = hot_amp_max * rng.uniform()
        xk, yk = r * np.cos(phi), r * np.sin(phi)
        e = e + amp * np.exp(-((x - xk) ** 2 + (y - yk) ** 2) / (2.0 * hot_sigma ** 2))
    return e / HBARC


def c1_event(seed=1):
    """61x61x1 lattice, dx=dy=0.1 fm: 30 GeV/fm^3 Gaussian (sigma 2 fm) plus
    10 hotspots (sigma 0.4 fm, amplitude U(0,10) GeV/fm^3) in the disc r<3 fm.
    Returns e [fm^-4] quantised like the reference's file reader, shape
    (61, 61, 1)."""
    e2 = hotspot_event_2d(61, 61, 0.1, 0.1, seed, n_hot=10, hot_sigma=0.4,
                          hot_amp_max=10.0, base_peak=30.0, base_sigma=2.0,
                          centre_radius=3.0)
    return quantise_like_reference(e2)[:, :, None]


def c2_event(nx=201, ny=201, nz=101, dx=0.1, dy=0.1, dz=0.1, seed=2,
             n_hot=200, hot_sigma=0.5, hot_amp_max=5.0, centre_radius=4.0):
    """3+1D event: 200 transverse hotspots (sigma 0.5 fm, amplitude
    U(0,5) GeV/fm^3, centres uniform in r < 4 fm) times the reference's
    plateau-plus-Gaussian-tail profile in eta.  Shape (nx, ny, nz), fm^-4,
    quantised like the reference's file reader."""
    eT = hotspot_event_2d(nx, ny, dx, dy, seed, n_hot=n_hot, hot_sigma=hot_sigma,
                          hot_amp_max=hot_amp_max, centre_radius=centre_radius)
    eL = longitudinal_profile(nz, dz)
    return quantise_like_reference(eT[:, :, None] * eL[None, None, :])


def c2_dense_event(nx=201, ny=201, nz=101, dx=0.1, dy=0.1, dz=0.1, seed=2, eta_flat=10.2,
                   base_peak=15.0, base_sigma=6.0):
    """The 3+1D event's 200 hotspots (as c2_event) on a transverse Gaussian
    base (base_peak GeV/fm^3, sigma base_sigma fm) times the reference's
    eta profile with a plateau of width eta_flat, by default wider than the
    lattice, so that no cell starts near vacuum.  c2_event's eta tails fall
    to the reader's 1e-3 fm^-4 floor and the viscous evolution goes unstable
    there near step 90 (the reference scheme's own behaviour); on the
    201x201x101 lattice this form runs 500 steps (a plateau of width 8, with
    tails on the lattice, fails near step 125).  Shape (nx, ny, nz), fm^-4,
    quantised like the reference's file reader."""
    eT = hotspot_event_2d(nx, ny, dx, dy, seed, n_hot=200, hot_sigma=0.5, hot_amp_max=5.0,
                          base_peak=base_peak, base_sigma=base_sigma, centre_radius=4.0)
    eL = longitudinal_profile(nz, dz, eta_flat=eta_flat)
    return quantise_like_reference(eT[:, :, None] * eL[None, None, :])


def serving_event(nx, ny, nz, dx, dz, seed, n_hot, base_peak=15.0, hot_amp_max=5.0):
    """Event of the batched and serving workloads: a transverse Gaussian base
    (base_peak GeV/fm^3, sigma 0.45 of the lattice half-width, so no cell
    starts near vacuum, where viscous evolution without regulation breaks
    down within ~100-200 steps) plus n_hot seeded hotspots (sigma 0.5 fm,
    amplitude U(0, hot_amp_max) GeV/fm^3, centres in the central half-width),
    times the plateau-plus-tail eta profile for nz > 1.  Shape (nx, ny, nz),
    fm^-4, quantised like the reference's file reader."""
    half = 0.5 * (min(nx, ny) - 1) * dx
    eT = hotspot_event_2d(nx, ny, dx, dx, seed, n_hot=n_hot, hot_sigma=0.5,
                          hot_amp_max=hot_amp_max, base_peak=base_peak,
                          base_sigma=0.45 * half, centre_radius=0.5 * half)
    eL = longitudinal_profile(nz, dz) if nz > 1 else np.ones(1)
    return quantise_like_reference(eT[:, :, None] * eL[None, None, :])


def at_rest(shape):
    """u^mu = (1,0,0,0) and pi^{mu nu} = 0, Pi = 0."""
    ut = np.ones(shape)
    z = np.zeros(shape)
    return ut, z.copy(), z.copy(), z.copy(), [z.copy() for _ in range(10)], z.copy()
