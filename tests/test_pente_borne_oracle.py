"""The oracle's pente-borne remap (projection-pente-borne, UtilesRemap.cc:73-814,
RemapADIService.cc:319-431, 549-581, 733-872), pinned by what the
mathematics fixes:

* INTY is the integral of a clipped linear function (trapezoid areas);
* the three-piece reconstruction reproduces a linear profile, so the
  integrated face flux of every limiter that keeps the linear slope is the
  exact integral over the swept length (ultrabee, whose slope is twice the
  linear one, stays within the neighbours' bounds);
* at a local extremum the flux is the first-order y0 ppv (formula 16);
* one step of a uniform translation advects a linear density profile
  exactly with debar fix 0 and 1, and as first-order upwind with fix 2
  (mass flux = the donor cell's density times the volume flux);
* each material's mass is kept in a closed box, up to the reference's
  fraction-threshold leak, and the volume fractions stay in [0, 1];
* the Sod tube converges to the exact Riemann solution;
* -mixte's generic gradient loop equals the specialised one where the
  pure limiter is the limiter.
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H
from mahyco_amd import abi, riemann

THR = 1e-10
LINEAR_EXACT = [abi.LIM_MINMOD, abi.LIM_SUPERBEE, abi.LIM_VANLEER, abi.LIM_MINMODG, abi.LIM_SUPERBEEG,
                abi.LIM_VANLEERG, abi.LIM_ARITHMETICG]


@pytest.fixture(scope="module")
def L():
    return H.oracle()


@pytest.fixture(scope="module")
def probes(L):
    lib = L.lib
    lib.mho_probe_pb_inty.restype = C.c_double
    lib.mho_probe_pb_inty.argtypes = [C.c_double] * 6
    lib.mho_probe_pb_flux.restype = C.c_double
    lib.mho_probe_pb_flux.argtypes = [C.c_int32, C.c_int32] + [C.c_double] * 9
    lib.mho_probe_limiter.restype = C.c_double
    lib.mho_probe_limiter.argtypes = [C.c_int32, C.c_double]
    lib.mho_probe_limiter_g.restype = C.c_double
    lib.mho_probe_limiter_g.argtypes = [C.c_int32] + [C.c_double] * 7
    return lib


def trapezoid_integral(X, x0, y0, x1, y1):
    """Area under the segment (x0, y0)-(x1, y1) left of X (x0 < x1)."""
    b = min(max(X, x0), x1)
    yb = y0 + (y1 - y0) * (b - x0) / (x1 - x0)
    return 0.5 * (b - x0) * (y0 + yb)


def test_inty_is_the_clipped_linear_integral(probes):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        x0 = rng.uniform(-1, 1)
        x1 = x0 + rng.uniform(1e-3, 2)
        y0, y1 = rng.uniform(-3, 3, 2)
        X = rng.uniform(x0 - 1, x1 + 1)
        got = probes.mho_probe_pb_inty(X, x0, y0, x1, y1, THR)
        ref = trapezoid_integral(X, x0, y0, x1, y1)
        assert abs(got - ref) <= 1e-13 * (1 + abs(ref)), (X, x0, y0, x1, y1)
    # a degenerate interval carries nothing
    assert probes.mho_probe_pb_inty(0.3, 0.2, 1.0, 0.2 + 1e-11, 5.0, THR) == 0.0


def limited_slope(probes, lim, ym, y0, yp, h):
    gb, gf = (y0 - ym) / h, (yp - y0) / h
    if lim < abi.LIM_MINMODG:
        g = 0.0
        if gb != 0.0:
            g += 0.5 * (probes.mho_probe_limiter(lim, gf / gb) * gb)
        if gf != 0.0:
            g += 0.5 * (probes.mho_probe_limiter(lim, gb / gf) * gf)
        return g
    return probes.mho_probe_limiter_g(lim, gf, gb, yp, ym, h, h, h)


@pytest.mark.parametrize("lim", LINEAR_EXACT + [abi.LIM_ULTRABEEG])
def test_flux_of_a_linear_profile(probes, lim):
    rng = np.random.default_rng(lim)
    for _ in range(200):
        h = rng.uniform(0.1, 2.0)
        y0 = rng.uniform(0.5, 3.0)
        s = rng.uniform(-0.9, 0.9) * y0 / h  # positive values (the flux is floored at zero)
        ym, yp = y0 - s * h, y0 + s * h
        g = limited_slope(probes, lim, ym, y0, yp, h)
        for frac in [0.05, 0.3, 0.5, 0.8, 1.0]:
            ppv = frac * h
            # exact integrals of y0 + s x over [-h/2, -h/2 + ppv] and [h/2 - ppv, h/2]
            a, b = -h / 2, -h / 2 + ppv
            ex0 = y0 * ppv + s * (b * b - a * a) / 2
            a, b = h / 2 - ppv, h / 2
            ex1 = y0 * ppv + s * (b * b - a * a) / 2
            f0 = probes.mho_probe_pb_flux(0, lim, ppv, h, h, h, y0, yp, ym, g, THR)
            f1 = probes.mho_probe_pb_flux(1, lim, ppv, h, h, h, y0, yp, ym, g, THR)
            if lim in LINEAR_EXACT:
                assert abs(f0 - ex0) < 1e-13 * y0 * h and abs(f1 - ex1) < 1e-13 * y0 * h, (lim, h, y0, s, frac)
            else:
                # the reconstruction stays within the neighbours' values
                lo, hi = min(ym, yp) * ppv, max(ym, yp) * ppv
                assert lo - 1e-13 <= f0 <= hi + 1e-13 and lo - 1e-13 <= f1 <= hi + 1e-13


@pytest.mark.parametrize("lim", LINEAR_EXACT + [abi.LIM_ULTRABEEG])
def test_flux_at_an_extremum_is_first_order(probes, lim):
    rng = np.random.default_rng(100 + lim)
    for _ in range(200):
        h = rng.uniform(0.1, 2.0)
        y0 = rng.uniform(1.0, 2.0)
        ym, yp = y0 - rng.uniform(0, 1), y0 - rng.uniform(0, 1)  # local maximum
        if rng.uniform() < 0.5:
            ym, yp = 2 * y0 - ym, 2 * y0 - yp  # local minimum
        g = limited_slope(probes, lim, ym, y0, yp, h)
        ppv = rng.uniform(0, 1) * h
        for t in (0, 1):
            assert probes.mho_probe_pb_flux(t, lim, ppv, h, h, h, y0, yp, ym, g, THR) == y0 * ppv


@pytest.mark.parametrize("lim", LINEAR_EXACT + [abi.LIM_ULTRABEEG])
def test_flux_of_a_monotone_step_is_bounded(probes, lim):
    """A volume fraction rising 0 -> 1 across the cell: the swept volume
    fraction never leaves [0, 1], whatever the widths."""
    rng = np.random.default_rng(200 + lim)
    for _ in range(500):
        h0, hp, hm = rng.uniform(0.2, 2.0, 3)
        y0 = rng.uniform(0, 1)
        g = limited_slope(probes, lim, 0.0, y0, 1.0, h0) if lim < abi.LIM_MINMODG else \
            probes.mho_probe_limiter_g(lim, (1 - y0) / (0.5 * (h0 + hp)), y0 / (0.5 * (h0 + hm)), 1.0, 0.0, h0, hp, hm)
        ppv = rng.uniform(0, 1) * h0
        for t in (0, 1):
            f = probes.mho_probe_pb_flux(t, lim, ppv, h0, hp, hm, y0, 1.0, 0.0, g, THR)
            assert -1e-15 <= f <= ppv * (1 + 1e-12), (lim, t, h0, hp, hm, y0, ppv, f)


def translation(L, fix, lim, nx=40, u0=0.5, dt=1e-3):
    """One step of a uniform x-translation of rho = 1 + 0.5 x at p = 1."""
    ny = 2
    opt = L.options(projection_pente_borne=1, projection_pente_borne_debar_fix=fix, projection_limiteur_id=lim,
                    deltat_init=dt)
    abi.set_bcs(opt, [("XMIN", abi.BC_VX, u0), ("XMAX", abi.BC_VX, u0), ("YMIN", abi.BC_VY, 0.0),
                      ("YMAX", abi.BC_VY, 0.0)])
    s = abi.Simulation(L, abi.mesh_desc(2, [nx, ny], [0, 0], [1.0, ny / nx]), opt)
    xc = (np.arange(nx) + 0.5) / nx
    rho = np.tile(1.0 + 0.5 * xc, ny)
    vel = np.zeros((2, s.nn))
    vel[0] = u0
    s.init_cells(np.zeros(nx * ny, np.int32), rho, np.ones(nx * ny), vel)
    s.run(1)
    st = s.download()
    return st, rho.reshape(ny, nx)[0], xc


@pytest.mark.parametrize("fix", [0, 1])
@pytest.mark.parametrize("lim", [abi.LIM_MINMOD, abi.LIM_SUPERBEE, abi.LIM_VANLEER, abi.LIM_MINMODG,
                                 abi.LIM_ARITHMETICG])
def test_translation_of_a_linear_profile_is_exact(L, fix, lim):
    st, rho0, xc = translation(L, fix, lim)
    nx = xc.size
    shift = 0.5 * 1e-3  # u0 times the first step's dt (deltat-init)
    rho = st["density"].reshape(2, nx)
    exact = 1.0 + 0.5 * (xc - shift)
    inner = slice(3, nx - 3)
    err = np.abs(rho[:, inner] - exact[inner]).max()
    print("pente-borne translation", fix, lim, err)
    assert err < 1e-13


def test_translation_with_fix_2_is_first_order_upwind(L):
    """Debar fix 2 takes the mass flux as the donor cell's density times the
    (exact) volume flux: rho_i - (u dt / h)(rho_i - rho_{i-1})."""
    st, rho0, xc = translation(L, 2, abi.LIM_MINMOD)
    nx = xc.size
    c = 0.5 * 1e-3 * nx
    rho = st["density"].reshape(2, nx)
    exact = rho0[1:] - c * (rho0[1:] - rho0[:-1])
    err = np.abs(rho[:, 3:nx - 3] - exact[2:nx - 4]).max()
    assert err < 1e-13, err


def pb_tp(L, fix, lim=abi.LIM_MINMOD, **kw):
    return H.tp_setup(L, 56, 24, projection_pente_borne=1, projection_pente_borne_debar_fix=fix,
                      projection_limiteur_id=lim, **kw)


def material_masses(st, ne):
    mask = st["env_mask"]
    out = []
    for k in range(ne):
        pure = mask == (1 << k)
        mixed = ((mask >> k) & 1).astype(bool) & ~pure
        out.append(st["cell_mass"][pure].sum() + st["cell_mass_env"][k][mixed].sum())
    return np.array(out)


@pytest.mark.parametrize("fix,lim", [(1, abi.LIM_MINMOD), (2, abi.LIM_SUPERBEE), (1, abi.LIM_VANLEERG)])
def test_triple_point_masses_kept_and_fractions_bounded(L, fix, lim):
    steps = 120
    s = pb_tp(L, fix, lim)
    m0 = material_masses(s.download(), 3)
    n_mixed = 0
    for _ in range(steps // 30):
        s.run(30)
        st = s.download()
        multi = np.array([bin(int(v)).count("1") for v in st["env_mask"]]) >= 2
        n_mixed = max(n_mixed, int(multi.sum()))
        f = st["fracvol_env"][:, multi]
        assert f.min() >= 0.0 and f.max() <= 1.0 + 1e-14
        assert np.all(st["density"] > 0) and np.all(np.isfinite(st["pressure"]))
    assert n_mixed > 20
    m1 = material_masses(st, 3)
    V = (7.0 / 56) * (3.0 / 24)
    bound = steps * n_mixed * 1e-10 * V * st["density"].max()
    print("pente-borne triple point mass change", fix, lim, m1 - m0, "bound", bound)
    assert np.abs(m1 - m0).max() < bound


def test_sod_converges(L):
    errs = []
    for nx in [100, 200, 400]:
        opt = L.options(projection_pente_borne=1, projection_pente_borne_debar_fix=1)
        abi.set_bcs(opt, abi.reflecting_walls(2))
        s = abi.Simulation(L, abi.mesh_desc(2, [nx, 2], [0, 0], [1.0, 2 / nx]), opt)
        s.init_sod(3)
        s.run(100000)
        st = s.download()
        rho = st["density"].reshape(2, nx)
        assert np.abs(rho - rho[0]).max() < 1e-12
        xc = (np.arange(nx) + 0.5) / nx
        rex, _, _ = riemann.sample(xc, st["global_time"], 0.5, (1, 0, 1), (0.125, 0, 0.1), 1.4)
        errs.append(float(np.mean(np.abs(rho[0] - rex))))
    print("pente-borne Sod L1(rho)", errs)
    assert errs[2] < 6e-3
    assert np.all(np.log2(np.array(errs[:-1]) / np.array(errs[1:])) > 0.5)


@pytest.mark.parametrize("pb", [0, 1])
def test_mixte_generic_loop_on_one_material(L, pb):
    """The generic gradient loop of -mixte equals the specialised one: with
    the pure limiter equal to the limiter (superbee), a one-material Sod runs
    bit for bit as without -mixte (debar fix 0, where the pure and mixed
    flux forms coincide).  With minmod as the limiter, the pure cells switch
    their cell slopes to superbee (the dual sweeps keep minmod): the run
    differs from both plain runs and keeps the mass."""
    def run(**kw):
        s = H.sod_setup(L, 60, 2, projection_pente_borne=pb, **kw)
        m0 = s.download()["cell_mass"].sum()
        s.run(40)
        st = s.download()
        assert abs(st["cell_mass"].sum() - m0) < 1e-13 * m0
        return st
    a = run(projection_limiteur_id=abi.LIM_SUPERBEE, projection_pente_borne_mixte=1, projection_limiteur_pure_id=1)
    b = run(projection_limiteur_id=abi.LIM_SUPERBEE)
    for f in ["density", "internal_energy", "velocity"]:
        assert np.array_equal(a[f], b[f]), f
    c = run(projection_limiteur_id=abi.LIM_MINMOD, projection_pente_borne_mixte=1, projection_limiteur_pure_id=1)
    d = run(projection_limiteur_id=abi.LIM_MINMOD)
    assert not np.array_equal(c["density"], d["density"])
    assert not np.array_equal(c["density"], b["density"])
