"""Learned limiter (MHG_LIM_LEARNED) in the oracle: the slope rule against
closed forms, and the fallback path against the minmod trajectory."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
from mahyco_amd import abi


@pytest.fixture(scope="module")
def lib():
    L = H.oracle()
    f = L.lib.mho_probe_learned_grad
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 4 + [C.c_int32] * 2 + [C.c_double] * 4
    return L


def lg(lib, alpha, phim, phi, phip, back_in, front_in, gb, gf, db=-0.5, df=0.5):
    return lib.lib.mho_probe_learned_grad(alpha, phim, phi, phip, back_in, front_in, gb, gf, db, df)


def test_learned_grad_closed_forms(lib):
    # linear profile 1, 2, 3 on unit cells: central slope 1, minmod slope 1
    assert lg(lib, 0.0, 1, 2, 3, 1, 1, 1, 1) == 0.0
    assert lg(lib, 1.0, 1, 2, 3, 1, 1, 1, 1) == 1.0
    assert lg(lib, 2.0, 1, 2, 3, 1, 1, 1, 1) == 2.0     # face values 1 and 3: on the bounds
    assert lg(lib, 2.5, 1, 2, 3, 1, 1, 1, 1) == 1.0     # 0.75 and 3.25: outside, minmod
    assert lg(lib, 0.5, 1, 2, 3, 1, 1, 1, 1) == 0.5
    # local maximum 1, 3, 2: central slope 0.5 overshoots 3; minmod of (2, -1) is 0
    assert lg(lib, 1.0, 1, 3, 2, 1, 1, 2, -1) == 0.0
    assert lg(lib, 0.0, 1, 3, 2, 1, 1, 2, -1) == 0.0
    # unequal slopes 1, 2, 5: minmod 1, central 2; alpha 1 gives faces 1 and 3
    assert lg(lib, 1.0, 1, 2, 5, 1, 1, 1, 3) == 2.0
    assert lg(lib, 4.0, 1, 2, 5, 1, 1, 1, 3) == 1.0    # faces -2, 6: outside
    # domain boundary behind the cell: gb = 0, bounds from the cell and its front
    assert lg(lib, 1.0, 0, 2, 3, 0, 1, 0, 1) == 0.0
    # NaN coefficient: the comparison fails, minmod
    assert lg(lib, float("nan"), 1, 2, 3, 1, 1, 1, 1) == 1.0
    # negative coefficient: minmod outright (unequal slopes 1, 3: minmod 1)
    assert lg(lib, -1.0, 1, 2, 5, 1, 1, 1, 3) == 1.0
    assert lg(lib, -1.0, 1, 3, 2, 1, 1, 2, -1) == 0.0
    # non-uniform faces: db = -0.25, df = 0.75, linear data: accepted up to the bounds
    assert lg(lib, 1.0, 1, 2, 3, 1, 1, 1, 1, -0.25, 0.75) == 1.0
    assert lg(lib, 1.5, 1, 2, 3, 1, 1, 1, 1, -0.25, 0.75) == 1.0   # 2 + 1.125 > 3


def test_consistent_env_slopes_closed_forms(lib):
    """Mixed cells of the learned mode: an environment's mass and energy
    slopes follow its kept volume slope (g_m = (phi_m / phi_V) g_V, g_E =
    (phi_E / phi_V) g_V), so its mass and energy face values are its volume
    face value times rho and rho e; without a kept volume slope, or with
    phi_V at or below the remap threshold (1e-10), they take the classical
    slopes."""
    f = lib.lib.mho_probe_consistent_slopes
    f.restype = None
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    f.argtypes = [C.c_int32, dp, dp, dp, np.ctypeslib.ndpointer(np.int32, flags="C")]
    # three environments: volumes, masses (rho 2, 5, 3), energies (e 3, 4, 7)
    V = np.array([0.6, 0.4, 0.0])
    m = V * np.array([2.0, 5.0, 3.0])
    E = m * np.array([3.0, 4.0, 7.0])
    phi = np.concatenate([V, m, E])
    gV = np.array([0.1, -0.2, 0.05])
    g = np.concatenate([gV, [9.0] * 6])
    gcl = np.arange(9, dtype=float) + 100.0
    f(3, phi, g, gcl, np.array([1, 0, 1], dtype=np.int32))
    assert g[3] == (m[0] / V[0]) * gV[0] and g[6] == (E[0] / V[0]) * gV[0]  # env 0 follows its volume
    assert g[4] == gcl[4] and g[7] == gcl[7]                            # env 1: volume slope not kept
    assert g[5] == gcl[5] and g[8] == gcl[8]                            # env 2: phi_V = 0
    for d in (-0.5, 0.37):  # face values: mass = rho * volume, energy = rho e * volume
        assert abs((m[0] + g[3] * d) - 2.0 * (V[0] + gV[0] * d)) < 1e-15
        assert abs((E[0] + g[6] * d) - 6.0 * (V[0] + gV[0] * d)) < 1e-14
    # a volume fraction at the remap threshold: classical; just above: follows
    phi2 = phi.copy()
    phi2[0] = 1e-10
    g2 = np.concatenate([gV, [9.0] * 6])
    f(3, phi2, g2, gcl, np.array([1, 1, 1], dtype=np.int32))
    assert g2[3] == gcl[3] and g2[6] == gcl[6]
    phi2[0] = 2e-10
    g2 = np.concatenate([gV, [9.0] * 6])
    f(3, phi2, g2, gcl, np.array([1, 1, 1], dtype=np.int32))
    assert g2[3] == (m[0] / 2e-10) * gV[0] and g2[6] == (E[0] / 2e-10) * gV[0]


def _state(s):
    return s.download(H.CELL_FIELDS + ["velocity", "node_coord", "env_mask"] + H.ENV_FIELDS)


@pytest.mark.parametrize("case", ["sod", "tp"])
def test_learned_with_huge_alpha_is_minmod(lib, case):
    """alpha = 1e30 puts every non-zero central slope outside the bounds, so
    every cell falls back to minmod: the trajectory is minmod's bit for bit."""
    if case == "sod":
        mk = lambda **kw: H.sod_setup(lib, 40, 4, **kw)  # noqa: E731
        steps = 40
    else:
        mk = lambda **kw: H.tp_setup(lib, 28, 12, **kw)  # noqa: E731
        steps = 30
    ref = mk()
    ref.run(steps)
    s = mk(projection_limiteur_id=abi.LIM_LEARNED)
    s.set_learned_alpha(np.full((2, s.nc), 1e30))
    s.run(steps)
    a, b = _state(s), _state(ref)
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k


def test_learned_changes_the_solution(lib):
    """alpha = 1 (central slopes where they keep the bounds) differs from
    minmod and still conserves mass."""
    ref = H.sod_setup(lib, 100, 2)
    ref.run(150)
    s = H.sod_setup(lib, 100, 2, projection_limiteur_id=abi.LIM_LEARNED)
    m0 = s.conservation()[0]
    s.set_learned_alpha(np.ones((2, s.nc)))
    s.run(150)
    d = np.max(np.abs(s.download(["density"])["density"] - ref.download(["density"])["density"]))
    assert d > 1e-6
    assert abs(s.conservation()[0] - m0) < 1e-13 * m0


def test_land_on_final_time(lib):
    """land_on_final_time: the run ends at final_time exactly (the last dt is
    shortened); without it the run stops before the step that would pass it."""
    a = H.sod_setup(lib, 100, 2, final_time=0.05)
    na = a.run(100000)
    b = H.sod_setup(lib, 100, 2, final_time=0.05, land_on_final_time=1)
    nb = b.run(100000)
    ta, tb = a.download(["density"])["global_time"], b.download(["density"])["global_time"]
    assert ta < 0.05 and nb == na + 1
    assert abs(tb - 0.05) <= 1e-15
