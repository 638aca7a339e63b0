"""GPU path vs the CPU oracle.

Per step, from an identical state: every field within 1e-12 relative
(|gpu - oracle| / max(|oracle|, 1e-3 * max|field|)).  The only arithmetic that
is not bit-identical is the characteristic length V^(1/dim): glibc's pow
differs by one ulp from the correctly rounded square root (2D, the device's
sqrt) and from the device pow (3D) in ~0.1% of inputs, which reaches the
pseudo-viscosity and, through dt, everything else at the 1e-16 level.
Trajectories are compared in the max-norm relative metric |a-b|_inf / |b|_inf."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def _one_step_parity(gpu, setup, warm_steps, nx, ny, **kw):
    o = setup(H.oracle(), nx, ny, **kw)
    if warm_steps:
        o.run(warm_steps)
    st = o.download()
    g = setup(gpu, nx, ny, **kw)
    g.upload(st)
    o.step()
    g.step()
    ref = o.download()
    got = g.download()
    d = H.compare_states(got, ref, o.nenv, nx, ny)
    print(setup.__name__, warm_steps, {k: f"{v:.2e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad


@pytest.mark.parametrize("warm", [0, 1, 7, 60])
def test_sod_one_step(gpu, warm):
    _one_step_parity(gpu, H.sod_setup, warm, 40, 4)


@pytest.mark.parametrize("warm", [0, 3, 25, 120])
def test_triple_point_one_step(gpu, warm):
    _one_step_parity(gpu, H.tp_setup, warm, 56, 24)


def test_sod_trajectory_to_end(gpu):
    o = H.sod_setup(H.oracle(), 100, 4)
    g = H.sod_setup(gpu, 100, 4)
    no = o.run(100000)
    ng = g.run(100000)
    ref, got = o.download(), g.download()
    assert no == ng
    d = H.compare_states(got, ref, 1, 100, 4, metric=H.rel_max)
    print("sod trajectory", no, {k: f"{v:.2e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10


def test_triple_point_trajectory(gpu):
    o = H.tp_setup(H.oracle(), 56, 24)
    g = H.tp_setup(gpu, 56, 24)
    o.run(200)
    g.run(200)
    ref, got = o.download(), g.download()
    d = H.compare_states(got, ref, 3, 56, 24, metric=H.rel_max)
    print("tp trajectory", {k: f"{v:.2e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10


def test_mass_conservation(gpu):
    """Sod: mass conserved to round-off.  Triple point: the reference scheme
    clamps negative material masses and drops volume fractions below 1e-10
    (RemapADIService.cc:952-959, RemapADIFinal.cc:124-141), so its mass drifts
    by ~1e-12; the GPU drift must equal the oracle's."""
    g = H.sod_setup(gpu, 100, 4)
    c0 = g.conservation()
    g.run(100000)
    c1 = g.conservation()
    assert abs(c1[0] - c0[0]) / c0[0] < 1e-14
    g = H.tp_setup(gpu, 112, 48)
    o = H.tp_setup(H.oracle(), 112, 48)
    c0, o0 = g.conservation(), o.conservation()
    g.run(100)
    o.run(100)
    c1, o1 = g.conservation(), o.conservation()
    assert c0[0] == o0[0]
    drift_g = (c1[0] - c0[0]) / c0[0]
    drift_o = (o1[0] - o0[0]) / o0[0]
    print("TP mass drift gpu", drift_g, "oracle", drift_o)
    assert abs(drift_g - drift_o) < 1e-14
    # total energy: same drift as the oracle (the scheme does not conserve it)
    eg = (c1[1] + c1[2] - c0[1] - c0[2]) / (c0[1] + c0[2])
    eo = (o1[1] + o1[2] - o0[1] - o0[2]) / (o0[1] + o0[2])
    print("TP energy drift gpu", eg, "oracle", eo)
    assert abs(eg - eo) < 1e-12


@pytest.mark.parametrize("kw", [dict(projection_limiteur_id=1), dict(projection_limiteur_id=2),
                                dict(ordre_projection=1), dict(pseudo_centree=1),
                                dict(longueur_caracteristique=1)])
def test_triple_point_options_one_step(gpu, kw):
    _one_step_parity(gpu, H.tp_setup, 40, 56, 24, **kw)


def test_two_material_sod_one_step(gpu):
    def bisod(L, nx, ny):
        opt = L.options()
        H.abi.set_bcs(opt, H.abi.reflecting_walls(2))
        H.abi.set_envs(opt, [(H.abi.EOS_PERFECT_GAS, 1.4, 0.), (H.abi.EOS_PERFECT_GAS, 1.4, 0.)])
        s = H.abi.Simulation(L, H.abi.mesh_desc(2, [nx, ny], [0, 0], [1.0, ny / nx]), opt)
        s.init_sod(13)
        return s
    bisod.__name__ = "bisod"
    _one_step_parity(gpu, bisod, 30, 40, 4)


@pytest.mark.parametrize("kernel", ["cell_sweep_y", "node_sweep_x", "node_sweep_y"])
@pytest.mark.parametrize("variant", range(6))
def test_launch_variants_one_step(gpu, kernel, variant):
    """Every launch variant of the tunable kernels, on a mesh that no strip,
    row chunk or 60-cell segment divides (70 x 134), from a state where both
    sweep orders have run (odd and even warm-up)."""
    sim = H.tp_setup(gpu, 70, 134)
    n = sim.variant_count(kernel)
    if variant >= n:
        pytest.skip(f"{kernel}: {n} variants")

    def tp_variant(L, nx, ny):
        s = H.tp_setup(L, nx, ny)
        if L is gpu:
            s.select_variant(variant, kernel)
        return s
    tp_variant.__name__ = f"{kernel}_v{variant}"
    for warm in (20, 21):
        _one_step_parity(gpu, tp_variant, warm, 70, 134)


def test_graph_replay_matches_direct_launches(gpu):
    """Steps replayed from captured HIP graphs are bit-identical to steps
    launched one kernel at a time, across graph groups and a ragged tail."""
    a = H.tp_setup(gpu, 70, 134)
    b = H.tp_setup(gpu, 70, 134)
    a.select_variant(0, "step_graph")
    b.select_variant(1, "step_graph")
    assert a.run(37) == 37 and b.run(37) == 37
    da, db = a.download(), b.download()
    for k in da:
        assert np.array_equal(np.asarray(da[k]), np.asarray(db[k])), k


@pytest.mark.parametrize("prec", [H.abi.PREC_FP64, H.abi.PREC_FP32_BULK])
def test_folded_dt_matches_single_steps(gpu, prec):
    """Inside one run the dt update of a step runs in the next step's node
    phase (two scalar buffers alternate) and the run's last step keeps its dt
    kernel; single steps each run the dt kernel.  29 steps in one call (two
    12-step graphs and a tail of 5) and 29 calls of one step give the same
    state and scalars bit for bit, and so does a run that stops at its final
    time inside a graph."""
    a = H.tp_setup(gpu, 70, 134, precision=prec)
    b = H.tp_setup(gpu, 70, 134, precision=prec)
    assert a.run(29) == 29
    for _ in range(29):
        assert b.step() == 0
    da, db = a.download(), b.download()
    for k in da:
        assert np.array_equal(np.asarray(da[k]), np.asarray(db[k])), k
    t_end = da["global_time"] + 0.5 * da["global_deltat"] * 15  # stops some steps into the next graph
    c = H.tp_setup(gpu, 70, 134, precision=prec, final_time=t_end)
    d = H.tp_setup(gpu, 70, 134, precision=prec, final_time=t_end)
    n = c.run(1000)
    m = 0
    while d.step() == 0:
        m += 1
    dc, dd = c.download(), d.download()
    assert dc["global_iteration"] == dd["global_iteration"] and n > 29
    for k in dc:
        assert np.array_equal(np.asarray(dc[k]), np.asarray(dd[k])), k


ENTRY_SEQUENCE = ["begin_iteration", "save_values_at_n", "update_velocity",
                  "apply_boundary_condition", "update_position", "compute_geometric_values",
                  "update_density", "compute_artificial_viscosity", "update_energy_and_pressure",
                  "remap"]


@pytest.mark.parametrize("case", ["sod", "tp", "tp_pc"])
def test_entry_points_match_oracle(gpu, case):
    """The per-entry-point C-ABI (Mahyco.config:30-41 order): after every entry
    point of two iterations the downloaded GPU state equals the oracle's, then
    a fused step continues from the entry-point state."""
    if case == "sod":
        setup, nx, ny, ne, warm, kw = H.sod_setup, 40, 4, 1, 5, {}
    else:
        setup, nx, ny, ne, warm, kw = H.tp_setup, 56, 24, 3, 30, {}
        if case == "tp_pc":
            kw = dict(pseudo_centree=1)
    o = setup(H.oracle(), nx, ny, **kw)
    o.run(warm)
    g = setup(gpu, nx, ny, **kw)
    g.upload(o.download())
    for it in range(2):
        for name in ENTRY_SEQUENCE:
            o.entry(name)
            g.entry(name)
            d = H.compare_states(g.download(), o.download(), ne, nx, ny)
            bad = {k: v for k, v in d.items() if v > TOL}
            assert not bad, (it, name, bad)
        dto, stop_o = o.compute_delta_t()
        dtg, stop_g = g.compute_delta_t()
        assert abs(dtg - dto) <= TOL * dto and stop_o == stop_g, (dtg, dto)
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), ne, nx, ny)
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, ("fused step after entry points", bad)


def test_entry_point_order_errors(gpu):
    """Entry points called out of order fail with MHG_E_STATE instead of
    computing on stale buffers."""
    g = H.sod_setup(gpu, 40, 4)
    with pytest.raises(H.abi.MahycoError):
        g.entry("update_position")
    with pytest.raises(H.abi.MahycoError):
        g.compute_delta_t()
    g.entry("begin_iteration")
    g.entry("save_values_at_n")
    with pytest.raises(H.abi.MahycoError):
        g.step()


def test_fp64_bit_identity_fixture(gpu):
    """The FP64 path reproduces, bit for bit, the state that this library
    wrote for the same runs when the fixture was made
    (tests/golden/make_gpu_fp64_fixture.py): operation order is part of the
    contract of the drop-in path."""
    import os
    import importlib.util
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "golden", "gpu_fp64_state.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    spec = importlib.util.spec_from_file_location(
        "mkfix", os.path.join(here, "golden", "make_gpu_fp64_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    ref = np.load(path)
    got = mk.run_all(gpu)
    for k in ref.files:
        assert np.array_equal(ref[k], got[k]), k


def _rel_l1(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def _rel_linf(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_fp32_precision_form_triple_point(gpu):
    """Precision form (FP32 bulk, FP64 coordinates, geometry and dt): relative
    L1 density error vs the FP64 oracle within 1e-5 after 200 triple-point
    steps, per-field L1/Linf printed; the same step count and mass within
    FP32 round-off of the oracle's."""
    nx, ny, steps = 112, 48, 200
    o = H.tp_setup(H.oracle(), nx, ny)
    g = H.tp_setup(gpu, nx, ny, precision=H.abi.PREC_FP32_BULK)
    o.run(steps)
    g.run(steps)
    ref, got = o.download(), g.download()
    errs = {f: (_rel_l1(got[f], ref[f]), _rel_linf(got[f], ref[f]))
            for f in ["density", "internal_energy", "pressure", "velocity", "node_coord"]}
    print("fp32 triple point", {k: f"L1 {a:.2e} Linf {b:.2e}" for k, (a, b) in errs.items()})
    assert got["global_iteration"] == ref["global_iteration"] == steps
    assert errs["density"][0] < 1e-5
    assert abs(got["global_time"] - ref["global_time"]) / ref["global_time"] < 1e-5
    m_o = np.sum(ref["cell_mass"])
    m_g = np.sum(np.asarray(got["cell_mass"], dtype=np.float64))
    assert abs(m_g - m_o) / m_o < 1e-5


def test_fp32_precision_form_sod_exact(gpu):
    """Precision form on Sod to t = 0.2: L1 density error vs the exact Riemann
    solution within 1 % of the FP64 oracle's own error."""
    from mahyco_amd import riemann
    nx = 400
    o = H.sod_setup(H.oracle(), nx, 4)
    g = H.sod_setup(gpu, nx, 4, precision=H.abi.PREC_FP32_BULK)
    o.run(100000)
    g.run(100000)
    ref, got = o.download(), g.download()
    x = (np.arange(nx) + 0.5) / nx
    exact, _, _ = riemann.sample(x, ref["global_time"], 0.5, (1, 0, 1), (0.125, 0, 0.1), 1.4)
    row_o = np.asarray(ref["density"]).reshape(4, nx)[0]
    row_g = np.asarray(got["density"]).reshape(4, nx)[0]
    e_o = np.mean(np.abs(row_o - exact))
    e_g = np.mean(np.abs(row_g - exact))
    print("sod L1 vs exact: fp64 oracle", e_o, "fp32 gpu", e_g)
    assert e_g < 1.01 * e_o


def test_run_async_past_the_end(gpu):
    """Steps enqueued past final-time are skipped by the device, and the next
    call sees the state of the steps really done: equal to the oracle's run
    to the end, and a later step stops at once with the state unchanged."""
    o = H.sod_setup(H.oracle(), 40, 4, final_time=0.02)
    g = H.sod_setup(gpu, 40, 4, final_time=0.02)
    n = o.run(100000)
    assert 20 < n < 200
    r = g.L.optional["run_async"](g.ctx, n + 37)
    assert r == 0
    assert g.L.optional["synchronize"](g.ctx) == 0
    got = g.download()
    assert got["global_iteration"] == n
    d = H.compare_states(got, o.download(), 1, 40, 4, metric=H.rel_max)
    print("run_async past the end", n, {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10
    # a second async call from the stopped state: nothing runs, nothing moves
    assert g.L.optional["run_async"](g.ctx, 13) == 0
    again = g.download()
    for f in ["density", "velocity", "node_coord"]:
        assert np.array_equal(again[f], got[f]), f
    assert g.step()


def test_sod_quadratic_viscosity_one(gpu):
    """The Sod configuration as stated (artificial viscosity 0.5 / 1.0; the
    reference hard-codes (gamma + 1) / 2 = 1.2): one step from oracle states
    within 1e-12, the whole run to t = 0.2 within 1e-10."""
    _one_step_parity(gpu, H.sod_setup, 30, 40, 4, visc_quad_coeff=1.0)
    o = H.sod_setup(H.oracle(), 100, 4, visc_quad_coeff=1.0)
    g = H.sod_setup(gpu, 100, 4, visc_quad_coeff=1.0)
    assert o.run(100000) == g.run(100000)
    d = H.compare_states(g.download(), o.download(), 1, 100, 4, metric=H.rel_max)
    assert max(d.values()) < 1e-10


@pytest.mark.parametrize("fusion", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("prec", [H.abi.PREC_FP64, H.abi.PREC_FP32_BULK])
@pytest.mark.parametrize("case", ["sod", "tp", "tp_sb", "tp_o1"])
def test_fusion_bit_identical(gpu, case, prec, fusion):
    """The fused step's final cell phase inside the last cell sweep (its CFL
    candidates in a kernel after the last dual sweep) and its Lagrange cell
    phase inside the first cell sweep give the state of the separate kernels
    bit for bit, over both sweep orders."""
    kw = {"tp_sb": dict(projection_limiteur_id=1), "tp_o1": dict(ordre_projection=1)}.get(case, {})
    setup = H.sod_setup if case == "sod" else H.tp_setup
    nx, ny = (40, 4) if case == "sod" else (70, 134)
    a = setup(gpu, nx, ny, precision=prec, **kw)
    b = setup(gpu, nx, ny, precision=prec, **kw)
    a.select_variant(0, "final_fusion")
    a.select_variant(0, "lagrange_fusion")
    b.select_variant(fusion[0], "final_fusion")
    b.select_variant(fusion[1], "lagrange_fusion")
    for n in (1, 2, 25):
        a.run(n)
        b.run(n)
        sa, sb = a.download(), b.download()
        for f, v in sa.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, sb[f]), (n, f)
            else:
                assert v == sb[f], (n, f)



@pytest.mark.parametrize("prec", [H.abi.PREC_FP64, H.abi.PREC_FP32_BULK])
def test_store_policy_bit_identical(gpu, prec):
    """The step kernels' store policy (ordinary or streaming stores) changes
    where the written lines go, never a value: both give the same state bit
    for bit, FP64 and FP32 bulk."""
    states = []
    for v in (0, 1, 6, 15):
        s = H.tp_setup(gpu, 70, 134, precision=prec)
        s.select_variant(v, "store_policy")
        s.run(13)
        states.append(s.download())
    for other in states[1:]:
        for f, v in states[0].items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, other[f]), f
            else:
                assert v == other[f], f


def test_formed_coordinates_match_stored(gpu):
    """Fused steps whose kernels all form the node coordinates as x0 + dt u
    do not store them; a context whose Lagrange cell phase is a kernel of
    its own (it reads stored coordinates) stores them every step.  Both give the same
    state bit for bit, coordinates included (written on demand), through
    graph groups, a download between steps, an entry-point iteration and
    further fused steps."""
    a = H.tp_setup(gpu, 70, 134)
    b = H.tp_setup(gpu, 70, 134)
    b.select_variant(0, "lagrange_fusion")

    def same(n):
        sa, sb = a.download(), b.download()
        for f, v in sa.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, sb[f]), (n, f)
            else:
                assert v == sb[f], (n, f)
    for n in (1, 25, 2):
        a.run(n)
        b.run(n)
        same(n)
    for name in ENTRY_SEQUENCE:
        a.entry(name)
        b.entry(name)
    a.compute_delta_t()
    b.compute_delta_t()
    same("entry points")
    a.run(14)
    b.run(14)
    same("after")


@pytest.mark.parametrize("prec", [H.abi.PREC_FP64, H.abi.PREC_FP32_BULK])
def test_ytile_width_bit_identical(gpu, prec):
    """The fused y sweeps on tiles of 8 columns (512-thread blocks) give the
    state of the 16-column tiles bit for bit, each sweep kind on its own
    (bits 1, 2, 4: lag, final, dual sweeps) and all together."""
    states = []
    for v in (0, 1, 2, 4, 7):
        s = H.tp_setup(gpu, 70, 134, precision=prec)
        s.select_variant(v, "ytile_width")
        s.run(13)
        states.append(s.download())
    for other in states[1:]:
        for f, v in states[0].items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, other[f]), f
            else:
                assert v == other[f], f
