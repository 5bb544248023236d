"""Workload definitions shared by the parity tests, smoke() and bench.py.

Meshes are the split-quad meshes of the benchmark configurations; initial conditions
are the reference's: the Gaussian pulse of the advection case and the isentropic vortex
of HopeFOAM-0.1/tutorials/DG/2D/isentropicVortex/dgEulerFoam/setNonUniformInlet.H
(r = squared distance to the centre, free stream (1,0), strength beta).
"""
import numpy as np

GAMMA = 1.4


def gl_min_spacing(N):
    """smallest spacing of the N+1 Gauss-Lobatto nodes on [-1,1] (Legendre.C:160-172)"""
    if N == 1:
        return 2.0
    from numpy.polynomial import legendre
    c = np.zeros(N + 1)
    c[-1] = 1
    x = np.sort(np.concatenate([[-1.0], legendre.legroots(legendre.legder(c)), [1.0]]))
    return float(np.min(np.diff(x)))


def inscribed_radius_split_quad(hx, hy):
    """inscribed-circle radius of the right triangle with legs hx, hy (2 area / perimeter)"""
    return hx * hy / (hx + hy + np.hypot(hx, hy))


def stable_dt(N, hx, hy, speed, cfl):
    """dt = CFL * (2/3) * min GL spacing * inscribed radius / max wave speed"""
    return cfl * (2.0 / 3.0) * gl_min_spacing(N) * inscribed_radius_split_quad(hx, hy) / speed


def gaussian_pulse(x, y):
    return np.exp(-40.0 * (x * x + y * y))


def vortex(x, y, t=0.0, x0=5.0, y0=5.0, beta=5.0, gamma=GAMMA):
    """exact isentropic vortex [rho, rhou, rhov, E] advected with the free stream (1, 0)"""
    dx = x - x0 - t
    dy = y - y0
    r = dx * dx + dy * dy
    rho = (1.0 - (gamma - 1.0) * beta ** 2 * np.exp(2.0 * (1.0 - r)) / (16.0 * gamma * np.pi ** 2)) ** (
        1.0 / (gamma - 1.0))
    u = 1.0 - beta * np.exp(1.0 - r) * dy / (2.0 * np.pi)
    v = beta * np.exp(1.0 - r) * dx / (2.0 * np.pi)
    E = rho ** gamma / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def vortex_periodic(x, y, t, L=10.0, **kw):
    """vortex on the periodic square [0,L]^2: centre advected and wrapped into the box"""
    xc = (5.0 + t) % L
    dxs = [-L, 0.0, L]
    best = None
    for sx in dxs:
        q = vortex(x, y, 0.0, x0=xc + sx, y0=5.0, **kw)
        if best is None:
            best = q
            dist = (x - xc - sx) ** 2
        else:
            d2 = (x - xc - sx) ** 2
            sel = d2 < dist
            best = np.where(sel[None], q, best)
            dist = np.minimum(dist, d2)
    return best


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def riemann4(x, y, seed, x0=0.0, x1=2.0, y0=0.0, y1=0.5, gamma=GAMMA):
    """Seeded four-quadrant 2D Riemann data [rho, rhou, rhov, E]: the split point is uniform in the
    middle half of the box, each quadrant's (rho, u, v, p) uniform in [0.1,2] x [-1,1]^2 x [0.1,2]."""
    rng = np.random.default_rng(seed)
    xs = x0 + (x1 - x0) * rng.uniform(0.25, 0.75)
    ys = y0 + (y1 - y0) * rng.uniform(0.25, 0.75)
    states = np.stack([rng.uniform(0.1, 2.0, 4), rng.uniform(-1.0, 1.0, 4), rng.uniform(-1.0, 1.0, 4),
                       rng.uniform(0.1, 2.0, 4)], axis=1)            # quadrant x (rho, u, v, p)
    quad = (x >= xs).astype(int) + 2 * (y >= ys).astype(int)
    rho, u, v, p = (states[quad, k] for k in range(4))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def smooth_random(x, y, seed, L=10.0, amp=0.3, gamma=GAMMA):
    """Seeded smooth periodic state on [0,L]^2 [rho, rhou, rhov, E]: each primitive variable is a constant
    plus three random Fourier modes with integer wavenumbers |k| <= 2 (rho, p in [1 - 3 amp, 1 + 3 amp])."""
    rng = np.random.default_rng(seed)

    def field(base, a):
        out = np.full_like(x, base)
        for _ in range(3):
            kx, ky = rng.integers(-2, 3, 2)
            if kx == 0 and ky == 0:
                kx = 1
            out += a * rng.uniform(-1, 1) * np.sin(2 * np.pi * (kx * x + ky * y) / L + rng.uniform(0, 2 * np.pi))
        return out

    rho, u, v, p = field(1.0, amp), field(0.0, 0.5), field(0.0, 0.5), field(1.0, amp)
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])
