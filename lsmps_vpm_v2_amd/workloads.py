"""Seeded synthetic particle fields for the benchmark configurations.

Every generator returns a dict of float64 arrays ``x, y, s, omega`` and uint8
``active``, plus ``q = -omega s^2 / (2 pi)`` (the reference's source packing,
velocity_calc.cpp:477).  Random numbers come from numpy's PCG64 with a fixed
seed, so the fields are identical on every machine.
"""
from __future__ import annotations

import numpy as np


def _pack(x, y, s, omega, active=None, **extra):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    omega = np.ascontiguousarray(omega, np.float64)
    active = np.ones(len(x), np.uint8) if active is None else np.ascontiguousarray(active, np.uint8)
    q = -omega * (s * s) / (2 * np.pi)
    return dict(x=x, y=y, s=s, omega=omega, active=active, q=q, n=len(x), **extra)


def lamb_oseen(x, y, cx, cy, gamma, core):
    """Vorticity of a Lamb-Oseen vortex: gamma / (pi a^2) exp(-r^2 / a^2)."""
    r2 = (x - cx) ** 2 + (y - cy) ** 2
    return gamma / (np.pi * core * core) * np.exp(-r2 / (core * core))


def uniform_vortices(seed: int = 1, h: float = 0.0632, half: float = 10.0, n_vort: int = 3,
                     core: float = 0.5, noise: float = 1e-3) -> dict:
    """Uniform cell-centred grid on [-half, half]^2 with spacing h (316^2 =
    99,856 particles at the defaults), vorticity = sum of n_vort Lamb-Oseen
    vortices (core 0.5, circulation U[-2, 2], centres U[-5, 5]^2) + N(0, noise)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = int(np.floor(2 * half / h))
    c = -half + (np.arange(m) + 0.5) * h
    X, Y = np.meshgrid(c, c, indexing="xy")
    x, y = X.ravel(), Y.ravel()
    gam = rng.uniform(-2.0, 2.0, n_vort)
    cen = rng.uniform(-5.0, 5.0, (n_vort, 2))
    w = np.zeros_like(x)
    for k in range(n_vort):
        w += lamb_oseen(x, y, cen[k, 0], cen[k, 1], gam[k], core)
    w += rng.normal(0.0, noise, len(x))
    return _pack(x, y, np.full(len(x), h), w, name="uniform_vortices_100k")


def boundary_layer(seed: int = 2, lx: float = 10.2, h0: float = 0.0025, band: float = 0.2,
                   levels: int = 4, half: float = 10.0, delta: float = 0.02, U: float = 1.0,
                   perturb: float = 0.05, drop: float = 1e-6) -> dict:
    """Impulsively started flat-plate boundary layer (plate length 1, thickness
    0.02, at the origin): rows of spacing h0 for |y| < band, then spacing doubling
    per band of doubled width for `levels` bands, out to |y| = half; columns span
    x in [-lx/2, lx/2] at the row spacing.  omega = sign(y) (U/delta) exp(-|y|/delta)
    (1 + perturb U[-1, 1]); particles with |omega| < drop are removed.  The
    defaults give about 780,000 particles in the strip |y| < 0.36."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs, ys, hs = [], [], []
    lo, hi, h = 0.0, band, h0
    for lev in range(levels + 1):
        if lev == levels:
            hi = half
        nrow = int(round((hi - lo) / h))
        yrow = lo + (np.arange(nrow) + 0.5) * h
        ncol = int(round(lx / h))
        xcol = -lx / 2 + (np.arange(ncol) + 0.5) * h
        X, Y = np.meshgrid(xcol, yrow, indexing="xy")
        for sgn in (1.0, -1.0):
            xs.append(X.ravel())
            ys.append(sgn * Y.ravel())
            hs.append(np.full(X.size, h))
        lo, hi, h = hi, 2 * hi, 2 * h
        if lo >= half:
            break
    x, y, s = np.concatenate(xs), np.concatenate(ys), np.concatenate(hs)
    w = np.sign(y) * (U / delta) * np.exp(-np.abs(y) / delta)
    keep = np.abs(w) >= drop
    x, y, s, w = x[keep], y[keep], s[keep], w[keep]
    w = w * (1.0 + perturb * rng.uniform(-1.0, 1.0, len(w)))
    return _pack(x, y, s, w, name="flat_plate_boundary_layer_780k")


def _square_patch(cx, cy, half_width, n_side):
    c = (np.arange(n_side) + 0.5) * (2 * half_width / n_side) - half_width
    X, Y = np.meshgrid(cx + c, cy + c, indexing="xy")
    return X.ravel(), Y.ravel(), 2 * half_width / n_side


def shed_vortices(seed: int = 4, n_total: int = 2_000_000, n_vort: int = 8) -> dict:
    """The flat-plate boundary layer (``boundary_layer()``, about 780k
    particles) plus ``n_vort`` shed Lamb-Oseen vortices near the plate: core
    radius U[0.05, 0.3], |circulation| U[10, 100] with a random sign, centres at
    x U[-4.5, 4.5], |y| U[0.5, 2.0] (outside the |y| < 0.36 band, patches not
    overlapping).  Each vortex is a square patch of side 6 core radii on its own
    grid, sized so that the whole field holds about ``n_total`` particles, and
    carries that vortex's vorticity; the boundary-layer particles keep theirs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bl = boundary_layer()
    per = (n_total - bl["n"]) // n_vort
    side = int(np.sqrt(per))
    xs, ys, ss, ws = [bl["x"]], [bl["y"]], [bl["s"]], [bl["omega"]]
    placed = []
    while len(placed) < n_vort:
        a = rng.uniform(0.05, 0.3)
        cx, cy = rng.uniform(-4.5, 4.5), rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 2.0)
        if abs(cy) - 3 * a < 0.4:
            continue
        if any(max(abs(cx - px), abs(cy - py)) < 3 * (a + pa) for px, py, pa in placed):
            continue
        placed.append((cx, cy, a))
    for k, (cx, cy, a) in enumerate(placed):
        gam = rng.choice([-1.0, 1.0]) * rng.uniform(10.0, 100.0)
        x, y, h = _square_patch(cx, cy, 3 * a, side)
        xs.append(x)
        ys.append(y)
        ss.append(np.full(len(x), h))
        ws.append(lamb_oseen(x, y, cx, cy, gam, a))
    x, y, s, w = (np.concatenate(v) for v in (xs, ys, ss, ws))
    return _pack(x, y, s, w, name=f"boundary_layer_shed_vortices_{len(x) // 1000}k",
                 vortices=np.array(placed))


def von_karman(seed: int = 5, n_total: int = 8_000_000, half: float = 20.0, n_vort: int = 40,
               core: float = 0.3, gamma: float = 1.5, radius: float = 0.5, h_body: float = 0.002,
               layer: float = 0.1) -> dict:
    """von Karman-street-like field on [-half, half]^2: ``n_vort`` Lamb-Oseen
    vortices (core 0.3, |circulation| 1.5, alternating sign) in two staggered
    rows y = +-0.6 behind a cylinder of radius 0.5 at the origin (spacing
    half / (n_vort / 2) in x from x = 1.5), on a uniform background grid outside
    r = radius + layer, plus a body-surface layer of rings at spacing h_body for
    radius <= r < radius + layer with omega = sign(y) 10 exp(-(r - radius) / 0.02).
    Every particle is kept (no vorticity cut), the background spacing is chosen
    so that the total is about ``n_total``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    # body layer: rings of equal arc spacing
    rr = radius + (np.arange(int(round(layer / h_body))) + 0.5) * h_body
    bx, by = [], []
    for r in rr:
        m = int(round(2 * np.pi * r / h_body))
        th = (np.arange(m) + 0.5) * (2 * np.pi / m) + rng.uniform(0, 2 * np.pi / m)
        bx.append(r * np.cos(th))
        by.append(r * np.sin(th))
    bx, by = np.concatenate(bx), np.concatenate(by)
    rb = np.hypot(bx, by)
    wb = np.sign(by) * 10.0 * np.exp(-(rb - radius) / 0.02)
    # background grid
    n_bg = n_total - len(bx)
    m = int(round(np.sqrt(n_bg / (1.0 - np.pi * (radius + layer) ** 2 / (2 * half) ** 2))))
    h = 2 * half / m
    c = -half + (np.arange(m) + 0.5) * h
    X, Y = np.meshgrid(c, c, indexing="xy")
    gx, gy = X.ravel(), Y.ravel()
    keep = gx * gx + gy * gy >= (radius + layer) ** 2
    gx, gy = gx[keep], gy[keep]
    dxv = half / (n_vort / 2)
    centres = [(1.5 + (k // 2) * dxv + (0.5 * dxv if k % 2 else 0.0), 0.6 if k % 2 == 0 else -0.6)
               for k in range(n_vort)]
    wg = np.zeros_like(gx)
    for k, (cx, cy) in enumerate(centres):
        sgn = 1.0 if k % 2 == 0 else -1.0
        wg += lamb_oseen(gx, gy, cx, cy, -sgn * gamma, core)
    x, y = np.concatenate([bx, gx]), np.concatenate([by, gy])
    s = np.concatenate([np.full(len(bx), h_body), np.full(len(gx), h)])
    w = np.concatenate([wb, wg])
    return _pack(x, y, s, w, name=f"von_karman_street_{len(x) // 1000}k")


def perlman(h: float = 0.02, half: float = 1.2) -> dict:
    """Perlman vortex omega = (1 - r^2)^7 (r <= 1) on a uniform grid
    (reference src/initialization/initVor.cpp:10-53) with its closed-form
    velocity u = -y/(16 r^2)(1 - (1 - r^2)^8), v = x/(16 r^2)(...) inside r <= 1
    and without the bracket outside (initVor.cpp:62-184)."""
    m = int(round(2 * half / h))
    c = -half + (np.arange(m) + 0.5) * h
    X, Y = np.meshgrid(c, c, indexing="xy")
    x, y = X.ravel(), Y.ravel()
    r2 = x * x + y * y
    w = np.where(r2 <= 1.0, np.clip(1.0 - r2, 0.0, None) ** 7, 0.0)
    brk = np.where(r2 <= 1.0, 1.0 - np.clip(1.0 - r2, 0.0, None) ** 8, 1.0)
    u = -y / (16 * r2) * brk
    v = x / (16 * r2) * brk
    return _pack(x, y, np.full(len(x), h), w, u_exact=u, v_exact=v, name="perlman")


def random_clusters(n: int, seed: int = 0, active_frac: float = 1.0) -> dict:
    """Half clustered Gaussian blob, half uniform background (test input)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n // 2
    x = np.concatenate([rng.normal(0.0, 0.3, m), rng.uniform(-3, 3, n - m)])
    y = np.concatenate([rng.normal(0.5, 0.2, m), rng.uniform(-2, 2, n - m)])
    s = np.full(n, 0.01)
    w = rng.normal(0.0, 1.0, n) * 100.0
    a = (rng.uniform(0, 1, n) < active_frac).astype(np.uint8)
    return _pack(x, y, s, w, active=a, name=f"clusters_{n}")


def subsample(field: dict, n_targets: int, seed: int = 7) -> np.ndarray:
    """Seeded sorted sample of target indices for accuracy checks."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = field["n"]
    if n_targets >= n:
        return np.arange(n, dtype=np.int32)
    return np.sort(rng.choice(n, n_targets, replace=False)).astype(np.int32)


def vortex_ring(n_target: int = 780_000, radius: float = 1.0, sigma: float = 0.1,
                gamma: float = 1.0, cutoff: float = 3.0, jitter: float = 0.0, seed: int = 6,
                active_frac: float = 1.0) -> dict:
    """3-D vortex ring: particles on a cubic lattice (spacing h, size s = h)
    inside the torus of ring radius ``radius`` and tube radius cutoff*sigma
    about the z = 0 circle, with the Gaussian core vorticity
    gamma / (pi sigma^2) exp(-d^2 / sigma^2) along the azimuth (d = distance
    to the core circle).  h is chosen so that about ``n_target`` particles
    fall inside; ``jitter`` (in units of h) perturbs the lattice.  Returns x,
    y, z, s, wx, wy, wz, active and n."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = cutoff * sigma
    vol = 2.0 * np.pi * radius * np.pi * a * a
    h = (vol / n_target) ** (1.0 / 3.0)
    ext = radius + a
    g = np.arange(-ext, ext + 0.5 * h, h)
    gz = np.arange(-a, a + 0.5 * h, h)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rho = np.hypot(X, Y)
    ring = np.abs(rho - radius) <= a
    xs, ys = X[ring], Y[ring]
    xs, zs = np.broadcast_arrays(xs[:, None], gz[None, :])
    ys = np.broadcast_to(ys[:, None], xs.shape)
    x, y, z = xs.ravel().copy(), ys.ravel().copy(), zs.ravel().copy()
    d2 = (np.hypot(x, y) - radius) ** 2 + z * z
    keep = d2 <= a * a
    x, y, z, d2 = x[keep], y[keep], z[keep], d2[keep]
    if jitter:
        x += jitter * h * rng.uniform(-0.5, 0.5, len(x))
        y += jitter * h * rng.uniform(-0.5, 0.5, len(x))
        z += jitter * h * rng.uniform(-0.5, 0.5, len(x))
    r = np.hypot(x, y)
    om = gamma / (np.pi * sigma * sigma) * np.exp(-d2 / (sigma * sigma))
    n = len(x)
    act = (rng.uniform(0, 1, n) < active_frac).astype(np.uint8)
    return dict(x=x, y=y, z=z, s=np.full(n, h), wx=-om * y / r, wy=om * x / r,
                wz=np.zeros(n), active=act, n=n, h=h, name=f"vortex_ring_{n}")
