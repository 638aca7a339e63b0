"""Workloads of the benchmark configurations, built on any C-ABI library.

* Sod shock tube: 2D, cas-test 3, [0,1] x [0, ny/nx], reflecting walls,
  options of mahyco/data/Data.1.arc (deltat-init 1e-5, min 1e-8, max 0.01,
  final-time 0.2, cfl 0.3, ADI order 2, minmod).
* Triple point: 2D [0,7] x [0,3], three perfect gases (rho, p, gamma):
  (1, 1, 1.5) for x<1, (1, 0.1, 1.4) for x>1, y<1.5, (0.125, 0.1, 1.5) for
  x>1, y>1.5, at rest, reflecting walls, cfl 0.3, run for a fixed number of
  steps (final-time and deltat-max out of reach).
"""
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
