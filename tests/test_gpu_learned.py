"""Learned limiter (MHG_LIM_LEARNED) on the GPU vs the oracle, with fixed
seeded coefficients: one step from oracle states (1e-12 per field, as for the
classical limiters), trajectories (1e-10), the fallback path (coefficient 1e30
gives the minmod trajectory bit for bit) and the FP32 form."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12
LEARNED = {"projection_limiteur_id": abi.LIM_LEARNED}


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def alphas(nc, seed):
    return np.random.default_rng(seed).uniform(0.0, 2.0, (2, nc))


@pytest.mark.parametrize("case,warm", [("sod", 0), ("sod", 30), ("tp", 0), ("tp", 10), ("tp", 40)])
def test_learned_one_step(gpu, case, warm):
    setup, nx, ny = (H.sod_setup, 100, 3) if case == "sod" else (H.tp_setup, 56, 24)
    o = setup(H.oracle(), nx, ny, **LEARNED)
    a = alphas(o.nc, 100 + warm)
    o.set_learned_alpha(a)
    if warm:
        o.run(warm)
    g = setup(gpu, nx, ny, **LEARNED)
    g.set_learned_alpha(a)
    g.upload(o.download())
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), o.nenv, nx, ny)
    print(case, warm, {k: f"{v:.2e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad


def test_learned_trajectories(gpu):
    for setup, nx, ny, steps in [(H.sod_setup, 100, 2, 150), (H.tp_setup, 56, 24, 60)]:
        o = setup(H.oracle(), nx, ny, **LEARNED)
        g = setup(gpu, nx, ny, **LEARNED)
        a = alphas(o.nc, 7)
        o.set_learned_alpha(a)
        g.set_learned_alpha(a)
        o.run(steps)
        g.run(steps)
        d = H.compare_states(g.download(), o.download(), o.nenv, nx, ny, metric=H.rel_max)
        print(setup.__name__, {k: f"{v:.1e}" for k, v in d.items()})
        assert max(d.values()) < 1e-10, d


def test_learned_huge_alpha_is_minmod_on_gpu(gpu):
    ref = H.tp_setup(gpu, 112, 48)
    ref.run(50)
    s = H.tp_setup(gpu, 112, 48, **LEARNED)
    s.set_learned_alpha(np.full((2, s.nc), 1e30))
    s.run(50)
    a, b = s.download(), ref.download()
    for k in H.CELL_FIELDS + ["velocity", "node_coord"]:
        assert np.array_equal(a[k], b[k]), k


def test_learned_fp32(gpu):
    a64 = H.tp_setup(gpu, 112, 48, **LEARNED)
    a32 = H.tp_setup(gpu, 112, 48, precision=abi.PREC_FP32_BULK, **LEARNED)
    al = alphas(a64.nc, 3)
    a64.set_learned_alpha(al)
    a32.set_learned_alpha(al)
    a64.run(100)
    a32.run(100)
    r64 = a64.download(["density"])["density"]
    r32 = a32.download(["density"])["density"]
    l1 = float(np.sum(np.abs(r32 - r64)) / np.sum(np.abs(r64)))
    print("learned fp32 L1", l1)
    assert l1 < 1e-5


def test_learned_options_rejected_where_unsupported(gpu):
    opt = gpu.options(projection_limiteur_id=abi.LIM_LEARNED)
    abi.set_bcs(opt, abi.reflecting_walls(3))
    with pytest.raises(abi.MahycoError) as e:
        abi.Simulation(gpu, abi.mesh_desc(3, [4, 4, 4], [0, 0, 0], [1, 1, 1]), opt)
    assert e.value.code == abi.E_UNSUPPORTED
    s = H.sod_setup(gpu, 20, 2)
    with pytest.raises(abi.MahycoError) as e:
        s.set_learned_alpha(np.zeros((2, s.nc)))
    assert e.value.code == abi.E_UNSUPPORTED


def test_land_on_final_time_matches_oracle(gpu):
    o = H.tp_setup(H.oracle(), 56, 24, final_time=0.2, land_on_final_time=1)
    g = H.tp_setup(gpu, 56, 24, final_time=0.2, land_on_final_time=1)
    no, ng = o.run(100000), g.run(100000)
    assert no == ng
    a, b = g.download(), o.download()
    assert abs(a["global_time"] - 0.2) <= 1e-15 and a["global_time"] == b["global_time"]
    d = H.compare_states(a, b, 3, 56, 24, metric=H.rel_max)
    assert max(d.values()) < 1e-10, d
