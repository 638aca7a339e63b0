"""Workloads of the benchmark configurations, built on any C-ABI library.

* Sod shock tube: 2D, cas-test 3, [0,1] x [0, ny/nx], reflecting walls,
  options of mahyco/data/Data.1.arc (deltat-init 1e-5, min 1e-8, max 0.01,
  final-time 0.2, cfl 0.3, ADI order 2, minmod).
* Triple point: 2D [0,7] x [0,3], three perfect gases (rho, p, gamma):
  (1, 1, 1.5) for x<1, (1, 0.1, 1.4) for x>1, y<1.5, (0.125, 0.1, 1.5) for
  x>1, y>1.5, at rest, reflecting walls, cfl 0.3, run for a fixed number of
  steps (final-time and deltat-max out of reach).
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
