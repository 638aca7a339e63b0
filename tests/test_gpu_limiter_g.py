"""Generalised limiters (minmodG .. ultrabeeG, fluxLimiterG UtilesRemap.cc:
127-173) on the GPU against the CPU oracle.

FP64: one step from the same state within 1e-12 relative, per field, on Sod
along x and along y and on the three-material triple point (mixed cells, both
sweep directions, the dual sweeps), on 3D Sod along each axis and a
three-material 3D box, and 60-step triple-point trajectories within 1e-12.  The
batched form runs the same block bodies on short meshes (whole columns per
wavefront) and tall ones (64-row tiles) and must follow each problem's own
context bit for bit.  FP32-bulk stays within 1e-5 (L1, relative) of FP64."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, cases, load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12
GLIMS = [abi.LIM_MINMODG, abi.LIM_SUPERBEEG, abi.LIM_VANLEERG, abi.LIM_ARITHMETICG, abi.LIM_ULTRABEEG]


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def _sod_y(L, nx, ny, **kw):
    return H.sod_setup(L, nx, ny, cas=4, **kw)


CASES = {"sod_x": (H.sod_setup, 40, 4, 1, 7), "sod_y": (_sod_y, 4, 40, 1, 7), "tp": (H.tp_setup, 56, 24, 3, 40)}


@pytest.mark.parametrize("lim", GLIMS)
@pytest.mark.parametrize("case", list(CASES))
def test_one_step_matches_oracle(gpu, case, lim):
    setup, nx, ny, ne, warm = CASES[case]
    o = setup(H.oracle(), nx, ny, projection_limiteur_id=lim)
    o.run(warm)
    st = o.download()
    g = setup(gpu, nx, ny, projection_limiteur_id=lim)
    g.upload(st)
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), ne, nx, ny)
    print(case, lim, {k: f"{v:.1e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad
    g.close()
    o.close()


@pytest.mark.parametrize("lim", GLIMS)
def test_triple_point_trajectory(gpu, lim):
    """60 steps within 1e-12.  Over longer runs the compressive limiters
    (superBeeG, arithmeticG) amplify the one-ulp differences of pow(V, 1/dim)
    (1e-10 .. 1e-9 after 150-200 steps at 56x24, where classical superbee
    itself goes non-finite by step 200), while every single step stays
    bit-identical (test_one_step_matches_oracle)."""
    o = H.tp_setup(H.oracle(), 56, 24, projection_limiteur_id=lim)
    g = H.tp_setup(gpu, 56, 24, projection_limiteur_id=lim)
    o.run(60)
    g.run(60)
    d = H.compare_states(g.download(), o.download(), 3, 56, 24, metric=H.rel_max)
    print("tp trajectory", lim, {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-12, d
    g.close()
    o.close()


def test_batch_matches_single_contexts(gpu):
    shapes = [(5, 3), (17, 8), (23, 30), (12, 59), (9, 60), (7, 64), (6, 130)]
    specs = cases.riemann_batch_specs(len(shapes), seed=21)
    for s, (nx, ny) in zip(specs, shapes):
        s["nx"], s["ny"] = nx, ny
    t_end = 0.02
    for lim in (abi.LIM_VANLEERG, abi.LIM_ULTRABEEG):
        b = cases.riemann_batch(gpu, specs, t_end, projection_limiteur_id=lim)
        b.run()
        for i, sp in enumerate(specs):
            s = cases.riemann_problem(gpu, sp, t_end, projection_limiteur_id=lim)
            n = s.run(100000)
            assert b.status(i)["iterations"] == n, (lim, i)
            a, r = b.problem(i).download(), s.download()
            for f in ["density", "internal_energy", "pressure", "velocity", "node_coord", "dual_grad_phi"]:
                assert np.array_equal(a[f], r[f]), (lim, i, sp["nx"], sp["ny"], f)
            s.close()
        b.close()


def test_fp32_close_to_fp64(gpu):
    res = {}
    for prec in (abi.PREC_FP64, abi.PREC_FP32_BULK):
        g = H.tp_setup(gpu, 112, 48, projection_limiteur_id=abi.LIM_VANLEERG, precision=prec)
        g.run(100)
        res[prec] = np.asarray(g.download(["density"])["density"]).copy()
        g.close()
    l1 = np.abs(res[1] - res[0]).sum() / np.abs(res[0]).sum()
    print("fp32 vs fp64 L1", l1)
    assert l1 < 1e-5, l1


@pytest.mark.parametrize("lim", [abi.LIM_SUPERBEEG, abi.LIM_VANLEERG, abi.LIM_ULTRABEEG])
@pytest.mark.parametrize("case", ["sod0", "sod1", "sod2", "box"])
def test_3d_one_step_matches_oracle(gpu, case, lim):
    """3D: the x lane sweeps and the y / z tile sweeps, cell and dual, on Sod
    along each axis and on the three-material box (hexahedral face centres
    from four nodes, cell centres from eight)."""
    from test_gpu_3d import box3, compare3
    if case.startswith("sod"):
        axis = int(case[3])
        setup = lambda L: H.sod3d_setup(L, 24, 3, axis, projection_limiteur_id=lim)  # noqa: E731
        nenv, n, warm = 1, [3, 3, 3], 6
        n[axis] = 24
    else:
        setup = lambda L: box3(L, projection_limiteur_id=lim)  # noqa: E731
        nenv, n, warm = 3, [14, 10, 8], 8
    o = setup(H.oracle())
    o.run(warm)
    g = setup(gpu)
    g.upload(o.download())
    o.step()
    g.step()
    d = compare3(g.download(), o.download(), nenv, n)
    print(case, lim, {k: f"{v:.1e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad
    g.close()
    o.close()
