"""Synthetic problem generators for the workloads the engine is measured on.

Each returns ``(cfg, state)`` with ``state`` a dict of host numpy arrays in the
reference's Field record layout (x, v: (n, dim); m, rho, eps: (n,); H: (n, dim, dim)).
Lattice H follows GenerateNodeDistribution3d (NodeGenerators/GenerateNodeDistribution3d.py:585-597):
H = I/(nPerh*dx).  Seeds use numpy's PCG64 so the same seed gives the same inputs on
any host.
"""
import math

import numpy as np

from . import abi


def _iso_H(n, dim, hinv):
    H = np.zeros((n, dim, dim))
    hinv = np.broadcast_to(np.asarray(hinv, np.float64), (n,))
    for a in range(dim):
        H[:, a, a] = hinv
    return H


def taylor_green_3d(n1d=128, jitter=0.1, seed=12345, nperh=2.0, rho0=1.0, P0=100.0, gamma=5.0 / 3.0,
                    vamp=0.1, kernel=abi.SPX_KERNEL_WENDLANDC4, precision=abi.SPX_PREC_FP64):
    """3-D periodic Taylor-Green box: n1d^3 equal-mass particles on a unit-cube lattice
    with uniform jitter of +-jitter spacing, rho=1, P=100, gamma=5/3,
    v = vamp*(sin X cos Y cos Z, -cos X sin Y cos Z, 0) with X = 2 pi x."""
    dx = 1.0 / n1d
    rng = np.random.Generator(np.random.PCG64(seed))
    g = (np.arange(n1d) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    x += rng.uniform(-jitter, jitter, size=x.shape) * dx
    x = np.mod(x, 1.0)
    n = x.shape[0]
    tp = 2.0 * math.pi
    v = np.zeros_like(x)
    v[:, 0] = vamp * np.sin(tp * x[:, 0]) * np.cos(tp * x[:, 1]) * np.cos(tp * x[:, 2])
    v[:, 1] = -vamp * np.cos(tp * x[:, 0]) * np.sin(tp * x[:, 1]) * np.cos(tp * x[:, 2])
    m = np.full(n, rho0 * dx ** 3)
    rho = np.full(n, rho0)
    eps = np.full(n, P0 / ((gamma - 1.0) * rho0))
    H = _iso_H(n, 3, 1.0 / (nperh * dx))
    cfg = abi.default_config(3, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    abi.set_periodic(cfg, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    return cfg, dict(x=x, v=v, m=m, rho=rho, eps=eps, H=H)


def sod_1d(n=400, nperh=1.51, gamma=1.4, rho1=1.0, P1=1.0, rho2=0.125, P2=0.1, x0=-0.5, x1=0.0, x2=0.5,
           kernel=abi.SPX_KERNEL_BSPLINE):
    """1-D Sod tube, n equal-mass particles on [x0, x2] placed by inverting the mass
    coordinate of the step profile (as GenerateNodeProfile1d does), reflecting walls."""
    Mleft = rho1 * (x1 - x0)
    Mright = rho2 * (x2 - x1)
    Mtot = Mleft + Mright
    mi = Mtot / n
    mc = (np.arange(n) + 0.5) * mi
    x = np.where(mc < Mleft, x0 + mc / rho1, x1 + (mc - Mleft) / rho2)
    rho = np.where(x < x1, rho1, rho2)
    P = np.where(x < x1, P1, P2)
    eps = P / ((gamma - 1.0) * rho)
    dxl = mi / rho
    H = _iso_H(n, 1, 1.0 / (nperh * dxl))
    cfg = abi.default_config(1, kernel=kernel, nperh=nperh, gamma=gamma)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    cfg.bc_lo[0] = abi.SPX_BC_REFLECTING
    cfg.bc_hi[0] = abi.SPX_BC_REFLECTING
    cfg.xmin[0] = x0
    cfg.xmax[0] = x2
    cfg.last_dt = 1.0e-6
    cfg.dt_min = 1.0e-8
    return cfg, dict(x=x.reshape(n, 1), v=np.zeros((n, 1)), m=np.full(n, mi), rho=rho, eps=eps, H=H)


def noh_planar_1d(nx1=100, nperh=1.35, order=5, gamma=5.0 / 3.0, rho1=1.0, eps1=0.0, x0=0.0, x1=1.0, vr0=-1.0):
    """The reference's planar Noh test as Noh-planar-1d.py sets it up with its defaults
    (tests/functional/Hydro/Noh/Noh-planar-1d.py:98-180, 330-560): NBSplineKernel(5) in a
    1000-point TableKernel, nx1 nodes from GenerateNodeDistribution1d (x_i = x0 + (i+1/2)dx,
    m = rho dx, h = nPerh dx), eps = 0, v = vr0 toward a reflecting wall at x0, the SPH()
    factory defaults (LimitedMonaghanGingold Q with Cl = extent, Cq = extent^2/2,
    SPHHydros.py:88-94), cfl 0.5, compatible energy, grad-h, corrected velocity gradient,
    CheapSynchronousRK2 with lastDt 1e-4, dtMin 1e-5, dtMax 0.1, dtGrowth 2, hmin 1e-4,
    hmax 0.1.  The run starts with SpheralController's iterateIdealH (50 passes, 1e-4)."""
    dx = (x1 - x0) / nx1
    h0 = nperh * dx
    x = x0 + (np.arange(nx1) + 0.5) * dx
    m = np.full(nx1, rho1 * dx)
    H = _iso_H(nx1, 1, 1.0 / h0)
    v = np.where(x > x0, vr0, -vr0)
    extent = abi.kernel_extent(abi.SPX_KERNEL_NBSPLINE, order)
    cfg = abi.default_config(1, kernel=abi.SPX_KERNEL_NBSPLINE, kernel_order=order, nperh=nperh, gamma=gamma,
                             hmin=1.0e-4, hmax=0.1, cfl=0.5, Cl=2.0 * (extent / 2.0), Cq=2.0 * (extent / 2.0) ** 2,
                             viscosity=abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD, dt_min=1.0e-5, dt_max=0.1,
                             dt_growth=2.0, last_dt=1.0e-4)
    cfg.kernel = abi.SPX_KERNEL_NBSPLINE
    cfg.neighbor_extent = extent
    cfg.bc_lo[0] = abi.SPX_BC_REFLECTING
    cfg.xmin[0] = x0
    cfg.bc_hi[0] = abi.SPX_BC_NONE
    cfg.xmax[0] = 1.0e100           # open side: no plane
    return cfg, dict(x=x.reshape(nx1, 1), v=v.reshape(nx1, 1), m=m, rho=np.full(nx1, rho1),
                     eps=np.full(nx1, eps1), H=H)


def collisionless_sphere_3d(nr=10, r0_au=1.0, M0_earth=1.0, plummer_au=1.0e-2, opening=1.0, fdt=0.1):
    """The reference's N-body collapse test as CollisionlessSphereCollapse.py sets it up
    (tests/functional/Gravity/CollisionlessSphereCollapse.py:40-190) in MKS units: a
    (2 nr)^3 lattice on [-r0, r0]^3 clipped to r <= r0 (GenerateNodeDistribution3d lattice,
    NodeGenerators/GenerateNodeDistribution3d.py:573-624), masses renormalised to M0,
    OctTreeGravity(G, plummer softening, opening, ftimestep, AccelerationRatio) as the only
    package, SynchronousRK2 with lastDt = dtMin = 0.1 dt, dtMax = 100 dt, dtGrowth 2, where
    dt = collapse time / 100.  Returns (cfg, state, collapse_time)."""
    AU = 149597870700.0
    Mearth = 5.9722e24
    G = 6.67430e-11                      # PhysicalConstants::GMKS (Material/PhysicalConstants.cc:48)
    r0 = r0_au * AU
    M0 = M0_earth * Mearth
    plummer = plummer_au * AU
    rho0 = M0 / (4.0 / 3.0 * math.pi * r0 * r0 * r0)
    tdyn = math.sqrt(3.0 * math.pi / (16 * G * rho0))
    collapse = tdyn / math.sqrt(2.0)
    dt = collapse / 100
    n = 2 * nr
    dx = (r0 - (-r0)) / n
    xs, ms = [], []
    for ig in range(n * n * n):
        i, j, k = ig % n, (ig // n) % n, ig // (n * n)
        xx, yy, zz = -r0 + (i + 0.5) * dx, -r0 + (j + 0.5) * dx, -r0 + (k + 0.5) * dx
        if (xx - 0.0) ** 2 + (yy - 0.0) ** 2 + (zz - 0.0) ** 2 <= r0 ** 2:
            xs.append((xx, yy, zz))
            ms.append(dx * dx * dx * rho0)
    msum = 0.0
    for mi in ms:
        msum += mi
    fmass = M0 / msum
    m = np.array(ms) * fmass
    x = np.array(xs)
    N = len(m)
    cfg = abi.default_config(3, kernel=abi.SPX_KERNEL_BSPLINE, nperh=2.01)
    cfg.gravity = 1
    cfg.gravity_only = 1
    cfg.integrator = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
    cfg.G = G
    cfg.softening = plummer
    cfg.opening = opening
    cfg.ftimestep = fdt
    cfg.dt_min, cfg.dt_max, cfg.dt_growth, cfg.last_dt = 0.1 * dt, 100.0 * dt, 2.0, 0.1 * dt
    st = dict(x=x, v=np.zeros((N, 3)), m=m, rho=np.full(N, rho0), eps=np.zeros(N),
              H=_iso_H(N, 3, 1.0 / (2.01 * dx)))
    return cfg, st, collapse


def kelvin_helmholtz_2d(n_target=4096, seed=7, nperh=2.0, gamma=5.0 / 3.0, P0=2.5,
                        kernel=abi.SPX_KERNEL_BSPLINE, precision=abi.SPX_PREC_MIXED):
    """2-D periodic KH: lattice of spacing sqrt(1/N) with the band |y-0.5|<0.25 at twice the
    particle density (rho=2, vx=+0.5), outside rho=1, vx=-0.5; P=2.5; vy=0.01 sin(4 pi x)."""
    nx = max(4, int(round(math.sqrt(n_target / 1.5))))
    dx = 1.0 / nx
    g = (np.arange(nx) + 0.5) * dx
    X, Y = np.meshgrid(g, g, indexing="ij")
    outer = np.abs(Y.ravel() - 0.5) >= 0.25
    xo = np.stack([X.ravel()[outer], Y.ravel()[outer]], 1)
    # inner band: spacing dx/sqrt(2)
    ni = int(round(nx * math.sqrt(2.0)))
    dxi = 1.0 / ni
    gi = (np.arange(ni) + 0.5) * dxi
    nyi = int(round(0.5 / dxi))
    gy = 0.25 + (np.arange(nyi) + 0.5) * (0.5 / nyi)
    Xi, Yi = np.meshgrid(gi, gy, indexing="ij")
    xi = np.stack([Xi.ravel(), Yi.ravel()], 1)
    x = np.concatenate([xo, xi], 0)
    n = x.shape[0]
    band = np.abs(x[:, 1] - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    vol = np.where(band, dxi * (0.5 / nyi), dx * dx)
    m = rho * vol
    v = np.zeros((n, 2))
    v[:, 0] = np.where(band, 0.5, -0.5)
    v[:, 1] = 0.01 * np.sin(4.0 * math.pi * x[:, 0])
    eps = P0 / ((gamma - 1.0) * rho)
    hloc = np.sqrt(vol)
    H = _iso_H(n, 2, 1.0 / (nperh * hloc))
    cfg = abi.default_config(2, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    abi.set_periodic(cfg, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    return cfg, dict(x=x, v=v, m=m, rho=rho, eps=eps, H=H)


def evrard_3d(n=250_000, seed=4242, nperh=1.51, gamma=5.0 / 3.0, u0=0.05, M=1.0, R=1.0,
              kernel=abi.SPX_KERNEL_BSPLINE, precision=abi.SPX_PREC_MIXED):
    """Evrard collapse: rho(r) = M/(2 pi R^2 r) sampled by r = R*U (mass fraction M(r)/M = (r/R)^2
    -> r = R sqrt(u)), isotropic directions, u=0.05, v=0; TreeGravity G=1, opening 0.5,
    softening 0.01."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = R * np.sqrt(rng.uniform(0.0, 1.0, n))
    mu = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    st = np.sqrt(1.0 - mu * mu)
    x = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * mu], 1)
    m = np.full(n, M / n)
    rho = M / (2.0 * math.pi * R * R * np.maximum(r, 1e-3 * R))
    eps = np.full(n, u0)
    hloc = (m / rho) ** (1.0 / 3.0)
    H = _iso_H(n, 3, 1.0 / (nperh * hloc))
    cfg = abi.default_config(3, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    cfg.gravity = 1
    cfg.G = 1.0
    cfg.opening = 0.5
    cfg.softening = 0.01
    cfg.ftimestep = 0.1
    return cfg, dict(x=x, v=np.zeros((n, 3)), m=m, rho=rho, eps=eps, H=H)


def gravity_disk_2d(n=20_000, seed=4243, nperh=1.51, gamma=5.0 / 3.0, u0=0.05, M=1.0, R=1.0,
                    kernel=abi.SPX_KERNEL_BSPLINE, precision=abi.SPX_PREC_FP64):
    """The Evrard set-up in 2-D (TreeGravity's quadtree, Gravity_PYB11.py QuadTreeGravity): a disk
    with surface density M/(2 pi R r) sampled by r = R*u (uniform in u), isotropic angles,
    u = 0.05, v = 0; TreeGravity G = 1, opening 0.5, softening 0.01 (2-D force law M/sqrt(r^2 + s^2),
    potential -M log(sqrt(r^2 + s^2)), TreeGravity.cc:54-64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = R * rng.uniform(0.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    x = np.stack([r * np.cos(ph), r * np.sin(ph)], 1)
    m = np.full(n, M / n)
    rho = M / (2.0 * math.pi * R * np.maximum(r, 1e-3 * R))
    eps = np.full(n, u0)
    hloc = np.sqrt(m / rho)
    H = _iso_H(n, 2, 1.0 / (nperh * hloc))
    cfg = abi.default_config(2, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    cfg.gravity = 1
    cfg.G = 1.0
    cfg.opening = 0.5
    cfg.softening = 0.01
    cfg.ftimestep = 0.1
    return cfg, dict(x=x, v=np.zeros((n, 2)), m=m, rho=rho, eps=eps, H=H)


def uniform_lattice(dim, n1d, nperh=2.0, kernel=abi.SPX_KERNEL_WENDLANDC4, rho0=1.0, P0=1.0, gamma=5.0 / 3.0,
                    periodic=True):
    """Perfect periodic lattice at rest (uniform pressure): a closed-form test state."""
    dx = 1.0 / n1d
    g = (np.arange(n1d) + 0.5) * dx
    grids = np.meshgrid(*([g] * dim), indexing="ij")
    x = np.stack([gg.ravel() for gg in grids], 1)
    n = x.shape[0]
    cfg = abi.default_config(dim, kernel=kernel, nperh=nperh, gamma=gamma)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    if periodic:
        abi.set_periodic(cfg, (0.0,) * 3, (1.0,) * 3)
    return cfg, dict(x=x, v=np.zeros((n, dim)), m=np.full(n, rho0 * dx ** dim), rho=np.full(n, rho0),
                     eps=np.full(n, P0 / ((gamma - 1.0) * rho0)), H=_iso_H(n, dim, 1.0 / (nperh * dx)))


def _lattice(lo, hi, spacing):
    """Cell-centred lattice filling the box [lo, hi) with the spacing rounded to fit."""
    axes, vol = [], 1.0
    for a, b in zip(lo, hi):
        m = max(1, int(round((b - a) / spacing)))
        d = (b - a) / m
        axes.append(a + (np.arange(m) + 0.5) * d)
        vol *= d
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel() for g in grids], 1), vol


def sod_tube(dim, n_target, width=0.25, nperh=1.51, gamma=1.4, rho1=1.0, P1=1.0, rho2=0.125, P2=0.1,
             kernel=abi.SPX_KERNEL_BSPLINE, precision=abi.SPX_PREC_MIXED):
    """2-D or 3-D Sod tube: x in [-0.5, 0.5], cross-section [0, width]^(dim-1), reflecting walls on
    every side; left rho1/P1, right rho2/P2 on lattices whose spacing ratio (rho1/rho2)^(1/dim) gives
    (nearly) equal masses; about n_target particles."""
    ratio = (rho1 / rho2) ** (1.0 / dim)
    dx = (0.5 * width ** (dim - 1) * (1.0 + rho2 / rho1) / n_target) ** (1.0 / dim)
    lo_t, hi_t = [0.0] * (dim - 1), [width] * (dim - 1)
    xl, vl = _lattice([-0.5] + lo_t, [0.0] + hi_t, dx)
    xr, vr = _lattice([0.0] + lo_t, [0.5] + hi_t, dx * ratio)
    x = np.concatenate([xl, xr], 0)
    n = x.shape[0]
    left = np.arange(n) < len(xl)
    rho = np.where(left, rho1, rho2)
    vol = np.where(left, vl, vr)
    P = np.where(left, P1, P2)
    cfg = abi.default_config(dim, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    for a in range(dim):
        cfg.bc_lo[a] = cfg.bc_hi[a] = abi.SPX_BC_REFLECTING
        cfg.xmin[a], cfg.xmax[a] = (-0.5, 0.5) if a == 0 else (0.0, width)
    cfg.last_dt = 1.0e-6
    cfg.dt_min = 1.0e-8
    return cfg, dict(x=x, v=np.zeros((n, dim)), m=rho * vol, rho=rho, eps=P / ((gamma - 1.0) * rho),
                     H=_iso_H(n, dim, 1.0 / (nperh * vol ** (1.0 / dim))))


def gresho_2d(n_target, nperh=1.51, gamma=5.0 / 3.0, kernel=abi.SPX_KERNEL_BSPLINE, precision=abi.SPX_PREC_MIXED):
    """Gresho vortex on the periodic box [-0.5, 0.5]^2 (tests/functional/Hydro/GreshoVortex/
    GreshoVortex.py:273-297): rho = 1, v_phi = 5r (r < 0.2), 2 - 5r (0.2 <= r < 0.4), 0 beyond,
    pressure in radial balance (P = 5 + 12.5 r^2, 9 + 12.5 r^2 - 20 r + 4 ln 5r, 3 + 4 ln 2)."""
    x, vol = _lattice([-0.5, -0.5], [0.5, 0.5], 1.0 / math.sqrt(n_target))
    n = x.shape[0]
    r = np.hypot(x[:, 0], x[:, 1])
    r2 = r * r
    inner, mid = r < 0.2, (r >= 0.2) & (r < 0.4)
    vphi = np.where(inner, 5.0 * r, np.where(mid, 2.0 - 5.0 * r, 0.0))
    P = np.where(inner, 5.0 + 12.5 * r2,
                 np.where(mid, 9.0 + 12.5 * r2 - 20.0 * r + 4.0 * np.log(5.0 * np.maximum(r, 1e-300)),
                          3.0 + 4.0 * math.log(2.0)))
    rs = np.maximum(r, 1e-300)
    v = np.stack([-vphi * x[:, 1] / rs, vphi * x[:, 0] / rs], 1)
    cfg = abi.default_config(2, kernel=kernel, nperh=nperh, gamma=gamma, precision=precision)
    cfg.kernel = kernel
    cfg.neighbor_extent = abi.KERNEL_EXTENT[kernel]
    abi.set_periodic(cfg, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    return cfg, dict(x=x, v=v, m=np.full(n, vol), rho=np.ones(n), eps=P / (gamma - 1.0),
                     H=_iso_H(n, 2, 1.0 / (nperh * math.sqrt(vol))))
