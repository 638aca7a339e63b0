"""Workloads of the benchmark configurations, built on any C-ABI library.

* Sod shock tube: 2D, cas-test 3, [0,1] x [0, ny/nx], reflecting walls,
  options of mahyco/data/Data.1.arc (deltat-init 1e-5, min 1e-8, max 0.01,
  final-time 0.2, cfl 0.3, ADI order 2, minmod).
* Triple point: 2D [0,7] x [0,3], three perfect gases (rho, p, gamma):
  (1, 1, 1.5) for x<1, (1, 0.1, 1.4) for x>1, y<1.5, (0.125, 0.1, 1.5) for
  x>1, y>1.5, at rest, reflecting walls, cfl 0.3, run for a fixed number of
  steps (final-time and deltat-max out of reach).
* Riemann batch: independent 2D x-Riemann problems of uneven sizes, each
  run to its own final time (batching form).
"""
import numpy as np

from . import abi

TP_RHO = (1.0, 1.0, 0.125)
TP_P = (1.0, 0.1, 0.1)
TP_GAMMA = (1.5, 1.4, 1.5)
TP_DOMAIN = (7.0, 3.0)


def sod(lib, nx=400, ny=20, cas_test=3, **optkw):
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(2))
    m = abi.mesh_desc(2, [nx, ny], [0, 0], [1.0, ny / nx])
    s = abi.Simulation(lib, m, opt)
    s.init_sod(cas_test)
    return s


def triple_point(lib, nx=1792, ny=768, **optkw):
    if nx % 7 or ny % 2:
        raise ValueError("nx must be a multiple of 7 and ny even so x=1 and y=1.5 lie on faces")
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(2))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, g, 0.) for g in TP_GAMMA])
    opt.final_time = 1e30
    opt.deltat_max = 1e30
    m = abi.mesh_desc(2, [nx, ny], [0, 0], list(TP_DOMAIN))
    s = abi.Simulation(lib, m, opt)
    s.init_triple_point(TP_RHO, TP_P)
    return s


# 3D random-field case of the precision form: unit cube, one perfect gas
# gamma = 5/3, rho = exp(G) with G a Gaussian random field (correlation
# length 8 cells, standard deviation 0.3), p = 1, a solenoidal random
# velocity field with rms 0.1, reflecting walls, 200 steps.
P3D_GAMMA = 5.0 / 3.0
P3D_SEED = 20261203


def _spectral_filter(n, corr_cells):
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    return np.exp(-0.5 * (kx * kx + ky * ky + kz * kz) * corr_cells * corr_cells), (kx, ky, kz)


def p3d_fields(n=256, seed=P3D_SEED, corr_cells=8.0, sigma=0.3, vrms=0.1):
    """Seeded fields of the 3D case: cell density (n^3, i fastest) and nodal
    velocity (3, (n+1)^3): Gaussian filtering of white noise in Fourier space
    (periodic), the velocity projected on its divergence-free part; the
    nodal velocity repeats the periodic field on the boundary nodes."""
    rng = np.random.default_rng(seed)
    filt, (kx, ky, kz) = _spectral_filter(n, corr_cells)
    g = np.fft.ifftn(np.fft.fftn(rng.standard_normal((n, n, n))) * filt).real
    g = (g - g.mean()) / g.std() * sigma
    rho = np.exp(g)  # index [i, j, k]
    vh = [np.fft.fftn(rng.standard_normal((n, n, n))) * filt for _ in range(3)]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    div = (kx * vh[0] + ky * vh[1] + kz * vh[2]) / k2
    v = [np.fft.ifftn(vh[d] - (kx, ky, kz)[d] * div).real for d in range(3)]
    rms = np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    v = [vd * (vrms / rms) for vd in v]
    idx = np.arange(n + 1) % n
    vel = np.stack([vd[np.ix_(idx, idx, idx)].transpose(2, 1, 0).reshape(-1) for vd in v])
    return rho.transpose(2, 1, 0).reshape(-1).copy(), vel


def p3d(lib, n=256, fields=None, **optkw):
    """The 3D case on an n^3 mesh (fields from p3d_fields unless given)."""
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(3))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, P3D_GAMMA, 0.)])
    opt.final_time = 1e30
    opt.deltat_max = 1e30
    rho, vel = fields if fields is not None else p3d_fields(n)
    s = abi.Simulation(lib, abi.mesh_desc(3, [n, n, n], [0, 0, 0], [1.0, 1.0, 1.0]), opt)
    s.init_cells(np.zeros(rho.size, dtype=np.int32), rho, np.ones(rho.size), vel)
    return s


# Batch of 2D Riemann problems of the batching form: nx ~ U{64..1024},
# ny ~ U{8..64} on [0, 1] x [0, ny/nx] (square cells); left/right density
# log-uniform in [0.1, 10], pressure log-uniform in [0.01, 10], x velocity
# uniform in [-1, 1]; interface x0 ~ U[0.3, 0.7]; one perfect gas with gamma
# ~ U[1.2, 1.67]; reflecting walls; each problem runs to t_end = 0.15.
BATCH_SEED = 20261204
BATCH_TEND = 0.15


def riemann_batch_specs(n=1024, seed=BATCH_SEED, nx_range=(64, 1024), ny_range=(8, 64)):
    """Seeded problem parameters, one dict per problem."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
        p = np.exp(rng.uniform(np.log(0.01), np.log(10.0), 2))
        u = rng.uniform(-1.0, 1.0, 2)
        specs.append(dict(nx=nx, ny=ny, rho=(float(rho[0]), float(rho[1])), p=(float(p[0]), float(p[1])),
                          u=(float(u[0]), float(u[1])), x0=float(rng.uniform(0.3, 0.7)),
                          gamma=float(rng.uniform(1.2, 1.67))))
    return specs


def riemann_mesh_options(lib, spec, t_end=BATCH_TEND, **optkw):
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(2))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, spec["gamma"], 0.)])
    opt.final_time = t_end
    m = abi.mesh_desc(2, [spec["nx"], spec["ny"]], [0, 0], [1.0, spec["ny"] / spec["nx"]])
    return m, opt


def riemann_fields(spec):
    """Cell env/density/pressure (i fastest) and nodal velocity (2, N_n):
    left state for cell centres and nodes with x < x0."""
    nx, ny = spec["nx"], spec["ny"]
    xc = (np.arange(nx) + 0.5) / nx
    left = np.tile(xc < spec["x0"], ny)
    rho = np.where(left, spec["rho"][0], spec["rho"][1])
    p = np.where(left, spec["p"][0], spec["p"][1])
    xn = np.arange(nx + 1) / nx
    ux = np.tile(np.where(xn < spec["x0"], spec["u"][0], spec["u"][1]), ny + 1)
    vel = np.stack([ux, np.zeros_like(ux)])
    return np.zeros(nx * ny, dtype=np.int32), rho, p, vel


def riemann_problem(lib, spec, t_end=BATCH_TEND, **optkw):
    """One problem of the batch as a context of its own."""
    m, opt = riemann_mesh_options(lib, spec, t_end, **optkw)
    s = abi.Simulation(lib, m, opt)
    s.init_cells(*riemann_fields(spec))
    return s


def riemann_batch(lib, specs, t_end=BATCH_TEND, **optkw):
    """The problems of `specs` as one batch (mhg_batch_*), initialised."""
    mo = [riemann_mesh_options(lib, sp, t_end, **optkw) for sp in specs]
    b = abi.Batch(lib, [m for m, _ in mo], [o for _, o in mo])
    for i, sp in enumerate(specs):
        b.problem(i).init_cells(*riemann_fields(sp))
    return b
