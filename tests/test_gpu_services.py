"""The reference's service interfaces through the C-ABI, against the oracle:

* the remap split as MahycoModule::remap calls it (PrepareRemap.cc:448-563):
  computeVariablesForRemap + computeFaceQuantitesForRemap, IRemap::appliRemap
  (ending in remapVariables), IEquationOfState::applyEOS per environment,
  computePressionMoyenne, then computeDeltaT; the state after every call
  equals the oracle's, and the split equals the fused mhg_remap bit for bit;
* IRemap::remapVariables, resizeRemapVariables, synchronizeUremap and the
  getters; IEquationOfState::initEOS / applyEOS / applyOneCellEOS and its
  constants;
* the per-entry-point calls of Mahyco.config:30-41 on 3D meshes;
* the Lagrange-only mode (with-projection = false, MahycoModule.cc:1468-1476).
"""
import numpy as np
import pytest

import helpers as H
import test_gpu_3d as T3
from mahyco_amd import abi, load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12
LAGRANGE = ["begin_iteration", "save_values_at_n", "update_velocity", "apply_boundary_condition",
            "update_position", "compute_geometric_values", "update_density",
            "compute_artificial_viscosity", "update_energy_and_pressure"]


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def case_setup(case, **kw):
    """(setup(L), nenv, compare(got, ref)) of a named case."""
    if case == "sod2d":
        return (lambda L: H.sod_setup(L, 40, 4, **kw)), 1, (lambda a, b: H.compare_states(a, b, 1, 40, 4))
    if case == "tp":
        return (lambda L: H.tp_setup(L, 56, 24, **kw)), 3, (lambda a, b: H.compare_states(a, b, 3, 56, 24))
    if case == "sod3d":
        n = [3, 24, 3]
        return (lambda L: H.sod3d_setup(L, 24, 3, 1, **kw)), 1, (lambda a, b: T3.compare3(a, b, 1, n))
    if case == "box3d":
        return (lambda L: T3.box3(L, **kw)), 3, (lambda a, b: T3.compare3(a, b, 3, [14, 10, 8]))
    raise ValueError(case)


def same(g, o, cmp, what, skip=()):
    d = cmp(g.download(), o.download())
    bad = {k: v for k, v in d.items() if v > TOL and k not in skip}
    assert not bad, (what, bad)


# Pressures and sound speeds between appliRemap and the end of the EOS /
# pressure-average calls: an environment that enters a cell in the remap has
# no pressure until applyEOS, and a mixed cell no average until
# computePressionMoyenne (their slots hold whatever they held before).
PC_FIELDS = ("pressure", "sound_speed", "pressure_env", "sound_speed_env")


WARM = {"sod2d": 5, "tp": 30, "sod3d": 6, "box3d": 12}


@pytest.mark.parametrize("case", ["sod2d", "tp", "sod3d", "box3d"])
def test_remap_split_matches_oracle(gpu, case):
    setup, ne, cmp = case_setup(case)
    o = setup(H.oracle())
    o.run(WARM[case])
    g = setup(gpu)
    g.upload(o.download())
    for it in range(2):
        for name in LAGRANGE:
            o.entry(name)
            g.entry(name)
        same(g, o, cmp, (it, "lagrange"))
        o.remap_prepare()
        g.remap_prepare()
        same(g, o, cmp, (it, "remap_prepare"))
        o.remap_appli()
        g.remap_appli()
        same(g, o, cmp, (it, "remap_appli"), PC_FIELDS)
        for k in range(ne):
            o.eos_apply(k)
            g.eos_apply(k)
        same(g, o, cmp, (it, "eos_apply"), ("pressure", "sound_speed"))
        o.compute_pression_moyenne()
        g.compute_pression_moyenne()
        same(g, o, cmp, (it, "compute_pression_moyenne"))
        dto, so = o.compute_delta_t()
        dtg, sg = g.compute_delta_t()
        assert abs(dtg - dto) <= TOL * dto and so == sg
    # a fused step continues from the split state
    o.step()
    g.step()
    same(g, o, cmp, "fused step after the split")


@pytest.mark.parametrize("case", ["tp", "box3d"])
def test_remap_split_equals_fused_remap(gpu, case):
    """The split calls and the fused mhg_remap give the same state bit for bit."""
    setup, ne, cmp = case_setup(case)
    o = setup(H.oracle())
    o.run(WARM[case])
    st = o.download()
    a, b = setup(gpu), setup(gpu)
    a.upload(st)
    b.upload(st)
    for _ in range(3):
        for name in LAGRANGE:
            a.entry(name)
            b.entry(name)
        a.entry("remap")
        b.remap_split()
        assert a.compute_delta_t() == b.compute_delta_t()
    sa, sb = a.download(), b.download()
    for f, v in sa.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, sb[f]), f


def test_remap_variables_again(gpu):
    """remapVariables called again after appliRemap derives the same state on
    the GPU and in the oracle."""
    setup, ne, cmp = case_setup("tp")
    o = setup(H.oracle())
    o.run(40)
    g = setup(gpu)
    g.upload(o.download())
    for s in (o, g):
        for name in LAGRANGE:
            s.entry(name)
        s.remap_prepare()
        s.remap_appli()
    before = g.download()
    o.remap_variables()
    g.remap_variables()
    same(g, o, cmp, "remap_variables", PC_FIELDS)
    after = g.download()
    for f in ["density", "internal_energy", "cell_mass", "env_mask", "fracvol_env", "density_env"]:
        assert np.array_equal(before[f], after[f]), f


def test_service_arguments_and_getters(gpu):
    g = H.tp_setup(gpu, 56, 24)
    L = gpu.optional
    assert L["remap_resize_variables"](g.ctx, 14, 3) == abi.OK
    assert L["remap_resize_variables"](g.ctx, 13, 3) == abi.E_ARG
    assert L["remap_synchronize_uremap"](g.ctx) == abi.OK
    assert g.eos_constants(0) == (1.5, 0.0) and g.eos_constants(1) == (1.4, 0.0)
    with pytest.raises(abi.MahycoError):
        g.eos_constants(3)
    for name in LAGRANGE:
        g.entry(name)
    with pytest.raises(abi.MahycoError) as e:  # appliRemap before the remap vectors
        g.remap_appli()
    assert e.value.code == abi.E_STATE
    g.remap_prepare()
    for bad in [(3, 1, 14, 3), (2, 0, 14, 3), (2, 1, 11, 3), (2, 1, 14, 2)]:
        with pytest.raises(abi.MahycoError) as e:
            g.remap_appli(bad)
        assert e.value.code == abi.E_ARG
    g.remap_appli()
    assert g.L.remap_get_ordre_projection(g.ctx) == 2 and g.L.remap_is_euler(g.ctx) == 1


@pytest.mark.parametrize("case", ["tp", "box3d"])
def test_eos_entry_points(gpu, case):
    """initEOS, applyEOS and applyOneCellEOS per environment on a state with
    mixed cells and perturbed pressures / energies (stiffened gas on
    environment 2 in the triple point)."""
    if case == "tp":
        def setup(L):
            opt = L.options()
            abi.set_bcs(opt, abi.reflecting_walls(2))
            abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, 1.5, 0.), (abi.EOS_PERFECT_GAS, 1.4, 0.),
                               (abi.EOS_STIFFENED_GAS, 1.5, 0.05)])
            opt.final_time = opt.deltat_max = 1e30
            s = abi.Simulation(L, abi.mesh_desc(2, [56, 24], [0, 0], [7.0, 3.0]), opt)
            s.init_triple_point(H.TP_RHO, H.TP_P)
            return s
        cmp = lambda a, b: H.compare_states(a, b, 3, 56, 24)  # noqa: E731
    else:
        setup, _, cmp = case_setup(case)
    o = setup(H.oracle())
    o.run(30)
    st = o.download()
    rng = np.random.default_rng(5)
    for f in ["pressure", "internal_energy"]:
        st[f] = st[f] * rng.uniform(0.9, 1.1, st[f].shape)
        st[f + "_env"] = st[f + "_env"] * rng.uniform(0.9, 1.1, st[f + "_env"].shape)
    o.upload(st)
    g = setup(gpu)
    g.upload(st)
    mixed = np.nonzero(np.array([bin(int(v)).count("1") for v in st["env_mask"]]) >= 2)[0]
    assert mixed.size > 0
    for k in range(3):
        o.eos_init(k)
        g.eos_init(k)
        same(g, o, cmp, ("eos_init", k))
    st = o.download()
    st["internal_energy"] = st["internal_energy"] * 1.01
    st["internal_energy_env"] = st["internal_energy_env"] * 1.01
    o.upload(st)
    g.upload(st)
    for cell in [int(mixed[0]), 0, int(mixed[-1])]:
        for k in range(3):
            o.eos_apply_one_cell(k, cell)
            g.eos_apply_one_cell(k, cell)
    same(g, o, cmp, "eos_apply_one_cell")
    for k in range(3):
        o.eos_apply(k)
        g.eos_apply(k)
        same(g, o, cmp, ("eos_apply", k))


@pytest.mark.parametrize("case", ["sod3d", "box3d"])
def test_3d_entry_points_match_oracle(gpu, case):
    """Mahyco.config:30-41 one call at a time on hexahedra: the GPU state
    equals the oracle's after every call of two iterations."""
    setup, ne, cmp = case_setup(case)
    o = setup(H.oracle())
    o.run(WARM[case])
    g = setup(gpu)
    g.upload(o.download())
    for it in range(2):
        for name in LAGRANGE + ["remap"]:
            o.entry(name)
            g.entry(name)
            same(g, o, cmp, (it, name))
        dto, _ = o.compute_delta_t()
        dtg, _ = g.compute_delta_t()
        assert abs(dtg - dto) <= TOL * dto
    o.step()
    g.step()
    same(g, o, cmp, "fused step after the entry points")


@pytest.mark.parametrize("case", ["sod2d", "tp", "sod3d", "box3d"])
def test_lagrange_only(gpu, case):
    """with-projection = false: every entry point of two iterations equals the
    oracle's; then 40 steps of mhg_run (entry-point kernels) within 1e-10."""
    setup, ne, cmp = case_setup(case, with_projection=0)
    o = setup(H.oracle())
    o.run(WARM[case])
    g = setup(gpu)
    g.upload(o.download())
    for it in range(2):
        for name in LAGRANGE + ["remap"]:
            o.entry(name)
            g.entry(name)
            same(g, o, cmp, (it, name))
        dto, _ = o.compute_delta_t()
        dtg, _ = g.compute_delta_t()
        assert abs(dtg - dto) <= TOL * dto
    no, ng = o.run(40), g.run(40)
    assert no == ng == 40
    d = {k: v for k, v in cmp(g.download(), o.download()).items()}
    print(case, "lagrange-only 40 steps", {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10


def test_lagrange_only_refuses_other_forms(gpu):
    m = abi.mesh_desc(2, [8, 4], [0, 0], [1, .5])
    for kw in [dict(with_projection=0, precision=abi.PREC_FP32_BULK),
               dict(with_projection=0, projection_limiteur_id=abi.LIM_LEARNED)]:
        with pytest.raises(abi.MahycoError) as e:
            abi.Simulation(gpu, m, gpu.options(**kw))
        assert e.value.code == abi.E_UNSUPPORTED
