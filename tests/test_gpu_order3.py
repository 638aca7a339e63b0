"""Order-3 remap (ordre-projection 3: Remap-ordre-3.cc, RemapADIService.cc:
583-640) on the GPU against the oracle.

FP64, one step from the same uploaded state within 1e-12 relative per field
(Sod along x and y; the three-material triple point, whose mixed cells take
the mass and energy rescaling of non-pure neighbourhoods), with the minmod
and a generalised limiter for the dual sweeps; a 40-step triple-point
trajectory from the initialisation (the mixed / pure codes then come from
the zero-initialised flags and the final cell phase, not from an upload)
within 1e-12; 3D one steps on Sod along each axis and the three-material box
and a 15-step box trajectory within 1e-12; FP32-bulk within 1e-5 (L1) of
FP64.  The learned mode and batches refuse order 3."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, cases, load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12
O3 = dict(ordre_projection=3)


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def _sod_y(L, nx, ny, **kw):
    return H.sod_setup(L, nx, ny, cas=4, **kw)


CASES = {"sod_x": (H.sod_setup, 40, 4, 1, 9), "sod_y": (_sod_y, 4, 40, 1, 9), "tp": (H.tp_setup, 56, 24, 3, 40)}


@pytest.mark.parametrize("lim", [abi.LIM_MINMOD, abi.LIM_SUPERBEEG])
@pytest.mark.parametrize("case", list(CASES))
def test_one_step_matches_oracle(gpu, case, lim):
    setup, nx, ny, ne, warm = CASES[case]
    o = setup(H.oracle(), nx, ny, projection_limiteur_id=lim, **O3)
    o.run(warm)
    st = o.download()
    g = setup(gpu, nx, ny, projection_limiteur_id=lim, **O3)
    g.upload(st)
    o.upload(st)  # the same mixed / pure codes on both sides (from the fractions)
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), ne, nx, ny)
    print(case, lim, {k: f"{v:.1e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad
    g.close()
    o.close()


def test_triple_point_trajectory(gpu):
    o = H.tp_setup(H.oracle(), 56, 24, **O3)
    g = H.tp_setup(gpu, 56, 24, **O3)
    o.run(40)
    g.run(40)
    a, b = g.download(), o.download()
    d = H.compare_states(a, b, 3, 56, 24, metric=H.rel_max)
    print("tp trajectory", {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-12, d
    ref = H.tp_setup(gpu, 56, 24)
    ref.run(40)
    assert not np.array_equal(a["density"], ref.download()["density"])  # order 3 is not order 2
    for s in (o, g, ref):
        s.close()


def test_fp32_close_to_fp64(gpu):
    res = {}
    for prec in (abi.PREC_FP64, abi.PREC_FP32_BULK):
        g = H.tp_setup(gpu, 112, 48, precision=prec, **O3)
        g.run(100)
        res[prec] = np.asarray(g.download(["density"])["density"]).copy()
        g.close()
    l1 = np.abs(res[1] - res[0]).sum() / np.abs(res[0]).sum()
    print("fp32 vs fp64 L1", l1)
    assert l1 < 1e-5, l1


@pytest.mark.parametrize("case", ["sod0", "sod1", "sod2", "box"])
def test_3d_one_step_matches_oracle(gpu, case):
    """3D: the x lane and y / z tile sweeps with three halo cells a side,
    hexahedral widths, and the codes through the three sweeps."""
    from test_gpu_3d import box3, compare3
    if case.startswith("sod"):
        axis = int(case[3])
        setup = lambda L: H.sod3d_setup(L, 24, 3, axis, **O3)  # noqa: E731
        nenv, n, warm = 1, [3, 3, 3], 6
        n[axis] = 24
    else:
        setup = lambda L: box3(L, **O3)  # noqa: E731
        nenv, n, warm = 3, [14, 10, 8], 8
    o = setup(H.oracle())
    o.run(warm)
    st = o.download()
    g = setup(gpu)
    g.upload(st)
    o.upload(st)
    o.step()
    g.step()
    d = compare3(g.download(), o.download(), nenv, n)
    print(case, {k: f"{v:.1e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad
    g.close()
    o.close()


def test_3d_box_trajectory(gpu):
    from test_gpu_3d import box3, compare3
    o = box3(H.oracle(), **O3)
    g = box3(gpu, **O3)
    o.run(15)
    g.run(15)
    d = compare3(g.download(), o.download(), 3, [14, 10, 8], metric=H.rel_max)
    print("box trajectory", {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-12, d


def test_refused(gpu):
    with pytest.raises(abi.MahycoError):
        H.tp_setup(gpu, 56, 24, projection_limiteur_id=abi.LIM_LEARNED, **O3)
    specs = cases.riemann_batch_specs(2, seed=3)
    for s in specs:
        s["nx"], s["ny"] = 16, 4
    with pytest.raises(abi.MahycoError):
        cases.riemann_batch(gpu, specs, 0.01, **O3)


def test_combined_options_match_oracle(gpu):
    """Order 3 with the vanLeerG dual sweeps and the total-energy fix-up
    together: one triple-point step within 1e-12 of the oracle."""
    kw = dict(projection_limiteur_id=abi.LIM_VANLEERG, conservation_energie_totale=1, **O3)
    o = H.tp_setup(H.oracle(), 56, 24, **kw)
    o.run(30)
    st = o.download()
    g = H.tp_setup(gpu, 56, 24, **kw)
    g.upload(st)
    o.upload(st)
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), 3, 56, 24)
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad


def test_3d_combined_fp32_close_to_fp64(gpu):
    """3D FP32-bulk with order 3, superBeeG dual sweeps and the fix-up stays
    within 1e-5 (L1, relative) of FP64 on the three-material box."""
    from test_gpu_3d import box3
    kw = dict(projection_limiteur_id=abi.LIM_SUPERBEEG, conservation_energie_totale=1, **O3)
    res = {}
    for prec in (abi.PREC_FP64, abi.PREC_FP32_BULK):
        s = box3(gpu, precision=prec, **kw)
        s.run(30)
        res[prec] = np.asarray(s.download(["density"])["density"]).copy()
        s.close()
    l1 = np.abs(res[1] - res[0]).sum() / np.abs(res[0]).sum()
    print("3d fp32 vs fp64 L1", l1)
    assert np.all(np.isfinite(res[1])) and l1 < 1e-5, l1
