"""Generalised limiters of the oracle (fluxLimiterG, UtilesRemap.cc:127-173;
cell widths RemapADIService.cc:223-255; dual widths UtilesRemap.cc:46-57).

No test case of the reference selects these limiters, so the restatement is
pinned by what the mathematics fixes:
  - on equal widths, minmodG, superBeeG and vanLeerG give the slopes of the
    classical minmod, superbee and van Leer limiters (themselves pinned by
    the reference's golden runs);
  - on unequal widths, vanLeerG and arithmeticG reproduce a linear profile
    exactly, and every limiter keeps the cell's face values between its
    neighbours' values (the simple-slope bound);
  - a slope break (gradients of opposite signs) gives a zero slope;
  - whole runs with minmodG / superBeeG / vanLeerG stay close to the
    classical limiters' runs, while a wrong sign or neighbour in the wiring
    would make them diverge.
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H
from mahyco_amd import abi


@pytest.fixture(scope="module")
def L():
    return H.oracle()


def _lib(L):
    lib = L.lib
    lib.mho_probe_limiter.restype = C.c_double
    lib.mho_probe_limiter.argtypes = [C.c_int32, C.c_double]
    lib.mho_probe_limiter_g.restype = C.c_double
    lib.mho_probe_limiter_g.argtypes = [C.c_int32] + [C.c_double] * 7
    return lib


def _classical_slope(lib, lid, gb, gf):
    g = 0.0
    if gb != 0.0:
        g += 0.5 * (lib.mho_probe_limiter(lid, gf / gb) * gb)
    if gf != 0.0:
        g += 0.5 * (lib.mho_probe_limiter(lid, gb / gf) * gf)
    return g


def test_equal_widths_reduce_to_classical(L):
    lib = _lib(L)
    rng = np.random.default_rng(11)
    for _ in range(400):
        h = rng.uniform(0.01, 2.0)
        sgn = rng.choice([-1.0, 1.0])
        gm, gp = sgn * rng.uniform(0.0, 3.0, 2)
        y0 = rng.uniform(-1, 1)
        ym, yp = y0 - gm * h, y0 + gp * h   # cell values of a 1D mesh of width h
        for gid, cid in ((abi.LIM_MINMODG, abi.LIM_MINMOD), (abi.LIM_SUPERBEEG, abi.LIM_SUPERBEE),
                         (abi.LIM_VANLEERG, abi.LIM_VANLEER)):
            got = lib.mho_probe_limiter_g(gid, gp, gm, yp, ym, h, h, h)
            exp = _classical_slope(lib, cid, gm, gp)
            assert abs(got - exp) <= 1e-13 * max(1.0, abs(exp)), (gid, gp, gm, got, exp)


def test_slope_break_gives_zero(L):
    lib = _lib(L)
    for gid in range(abi.LIM_MINMODG, abi.LIM_ULTRABEEG + 1):
        assert lib.mho_probe_limiter_g(gid, 1.0, -2.0, 0.5, 0.2, 1.0, 0.5, 2.0) == 0.0
        assert lib.mho_probe_limiter_g(gid, -0.3, 0.7, 0.5, 0.2, 1.0, 0.5, 2.0) == 0.0


def _mesh3(rng):
    """Widths of three adjacent cells and their centres."""
    hm, h0, hp = rng.uniform(0.05, 2.0, 3)
    xm, x0, xp = -0.5 * (hm + h0), 0.0, 0.5 * (h0 + hp)
    return hm, h0, hp, xm, x0, xp


def test_linear_profiles_exact_on_unequal_widths(L):
    lib = _lib(L)
    rng = np.random.default_rng(12)
    for _ in range(400):
        hm, h0, hp, xm, x0, xp = _mesh3(rng)
        a, s = rng.uniform(-1, 1, 2)
        ym, y0, yp = a + s * xm, a + s * x0, a + s * xp
        gm, gp = (y0 - ym) / (x0 - xm), (yp - y0) / (xp - x0)
        for gid in (abi.LIM_VANLEERG, abi.LIM_ARITHMETICG, abi.LIM_MINMODG, abi.LIM_SUPERBEEG):
            got = lib.mho_probe_limiter_g(gid, gp, gm, yp, ym, h0, hp, hm)
            assert abs(got - s) <= 1e-12 * max(1.0, abs(s)), (gid, got, s)


def test_face_values_stay_between_neighbours(L):
    lib = _lib(L)
    rng = np.random.default_rng(13)
    for _ in range(2000):
        hm, h0, hp, xm, x0, xp = _mesh3(rng)
        ym, y0, yp = rng.uniform(-1, 1, 3)
        gm, gp = (y0 - ym) / (x0 - xm), (yp - y0) / (xp - x0)
        for gid in range(abi.LIM_MINMODG, abi.LIM_ULTRABEEG + 1):
            g = lib.mho_probe_limiter_g(gid, gp, gm, yp, ym, h0, hp, hm)
            front, back = y0 + 0.5 * h0 * g, y0 - 0.5 * h0 * g
            eps = 1e-12
            assert min(y0, yp) - eps <= front <= max(y0, yp) + eps, (gid, front, y0, yp)
            assert min(y0, ym) - eps <= back <= max(y0, ym) + eps, (gid, back, y0, ym)
            if gm * gp < 0:
                assert g == 0.0


def _sod_density(L, lim, steps=60):
    s = H.sod_setup(L, 80, 3, projection_limiteur_id=lim)
    s.run(steps)
    rho = s.download(["density"])["density"].copy()
    s.close()
    return rho


@pytest.mark.parametrize("gid,cid", [(abi.LIM_MINMODG, abi.LIM_MINMOD), (abi.LIM_SUPERBEEG, abi.LIM_SUPERBEE),
                                     (abi.LIM_VANLEERG, abi.LIM_VANLEER)])
def test_sod_close_to_classical(L, gid, cid):
    """The Lagrange phase leaves the widths almost equal on Sod, so a
    generalised limiter's run stays within a small distance of its classical
    counterpart's, and much closer to it than to the order-1 run."""
    g, c = _sod_density(L, gid), _sod_density(L, cid)
    s = H.sod_setup(L, 80, 3, ordre_projection=1)
    s.run(60)
    o1 = s.download(["density"])["density"].copy()
    s.close()
    assert np.all(np.isfinite(g))
    d_gc = np.abs(g - c).mean()
    d_c1 = np.abs(c - o1).mean()
    assert d_gc < 0.05 * d_c1, (d_gc, d_c1)


@pytest.mark.parametrize("gid", [abi.LIM_ARITHMETICG, abi.LIM_ULTRABEEG])
def test_sod_runs_bounded(L, gid):
    """arithmeticG and ultrabeeG keep the Sod densities within the initial
    range [0.125, 1] (the simple-slope bound keeps the remap monotone)."""
    g = _sod_density(L, gid)
    assert np.all(np.isfinite(g))
    assert g.min() >= 0.125 - 1e-9 and g.max() <= 1.0 + 1e-9, (g.min(), g.max())


def test_cell_widths_are_the_cell_extents(L):
    """After the Lagrange phase of some Sod steps (rectangular cells of
    unequal widths along x) the width along x of an inner cell is its extent
    x_{i+1} - x_i, and half of it for a cell on the mesh boundary, whose
    boundary face is left out; along y every width is the same."""
    lib = L.lib
    lib.mho_probe_cell_widths.restype = C.c_int
    lib.mho_probe_cell_widths.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
    nx, ny = 30, 3
    s = H.sod_setup(L, nx, ny, projection_limiteur_id=abi.LIM_MINMODG)
    s.run(20)
    for name in ["begin_iteration", "save_values_at_n", "update_velocity", "apply_boundary_condition",
                 "update_position", "compute_geometric_values"]:
        s.entry(name)
    x = np.asarray(s.download(["node_coord"])["node_coord"])[0].reshape(ny + 1, nx + 1)
    ext = np.diff(x, axis=1)[0]                      # same on every row
    assert np.ptp(ext) > 1e-6 * ext.mean()           # the widths are unequal
    for idir, n in ((0, nx), (1, ny)):
        h = np.zeros(nx * ny)
        assert lib.mho_probe_cell_widths(s.ctx, idir, h.ctypes.data_as(C.POINTER(C.c_double))) == 0
        h = h.reshape(ny, nx)
        if idir == 0:
            exp = ext.copy()
            exp[0] *= 0.5
            exp[-1] *= 0.5
            assert np.allclose(h, exp[None, :], rtol=1e-13, atol=0)
        else:
            dy = 1.0 / nx
            exp = np.full(ny, dy)
            exp[0] *= 0.5
            exp[-1] *= 0.5
            assert np.allclose(h, exp[:, None], rtol=1e-12, atol=0)
    s.close()
