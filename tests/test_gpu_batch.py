"""Batching form (mhg_batch_*): many independent 2D problems in persistent
kernels.

Each problem of a batch must follow, bit for bit, the trajectory the same
problem follows in a context of its own (same block bodies, same arithmetic,
its own dt and stop test), in FP64 and in FP32-bulk; in FP64 it must also match
the CPU oracle.  The meshes cover both y-sweep forms: whole columns in
segments of ny + 4 (cells) or ny + 5 (nodes) lanes, several per wavefront,
and 64-row tiles (ny > 60 for cells, ny > 59 for nodes), with row-tile
counts of 1 to 3."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, cases, load_gpu

pytestmark = pytest.mark.gpu

FIELDS = ["density", "internal_energy", "pressure", "sound_speed", "pseudo_viscosity", "cell_mass",
          "velocity", "node_coord", "dual_grad_phi"]
# (nx, ny): ny = 3, 8 (8 and 7 columns per wavefront), 20, 30, 59 (one segment of
# 63 / 64 lanes), 60 (cells in segments, nodes in tiles), 61, 64, 130 (tiles)
SHAPES = [(5, 3), (17, 8), (40, 20), (23, 30), (12, 59), (9, 60), (21, 61), (7, 64), (6, 130), (70, 4)]


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def specs_for(shapes, seed=11):
    sp = cases.riemann_batch_specs(len(shapes), seed=seed)
    for s, (nx, ny) in zip(sp, shapes):
        s["nx"], s["ny"] = nx, ny
    return sp


def same(a, b):
    bad = [f for f in FIELDS if not np.array_equal(a[f], b[f])]
    for k in ["global_deltat", "old_deltat", "global_time", "global_iteration"]:
        if a[k] != b[k]:
            bad.append(k)
    return bad


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK])
def test_batch_matches_single_contexts(gpu, prec):
    t_end = 0.03
    specs = specs_for(SHAPES)
    b = cases.riemann_batch(gpu, specs, t_end, precision=prec)
    nrun, zc = b.run()
    assert nrun == 0
    total = 0
    for i, sp in enumerate(specs):
        s = cases.riemann_problem(gpu, sp, t_end, precision=prec)
        n = s.run(100000)
        st = b.status(i)
        assert st["state"] == 1 and st["iterations"] == n, (i, st, n)
        total += sp["nx"] * sp["ny"] * n
        bad = same(b.problem(i).download(), s.download())
        assert not bad, (i, sp["nx"], sp["ny"], bad)
        s.close()
    assert zc == total
    b.close()


def test_batch_fp64_matches_oracle(gpu):
    t_end = 0.02
    specs = specs_for(SHAPES[:6], seed=5)
    b = cases.riemann_batch(gpu, specs, t_end)
    b.run()
    for i, sp in enumerate(specs):
        o = cases.riemann_problem(H.oracle(), sp, t_end)
        n = o.run(100000)
        assert b.status(i)["iterations"] == n
        d = H.compare_states(b.problem(i).download(), o.download(), 1, metric=H.rel_max)
        print(i, sp["nx"], sp["ny"], n, {k: f"{v:.1e}" for k, v in d.items()})
        assert max(d.values()) < 1e-10, (i, d)
    b.close()


def test_batch_split_runs_and_autotune_keep_state(gpu):
    """Two calls (7 iterations, then the rest) and a prior autotune leave the
    trajectory of every problem unchanged."""
    t_end = 0.03
    specs = specs_for(SHAPES, seed=3)
    ref = cases.riemann_batch(gpu, specs, t_end, precision=abi.PREC_FP32_BULK)
    ref.run()
    b = cases.riemann_batch(gpu, specs, t_end, precision=abi.PREC_FP32_BULK)
    report = b.autotune(1)
    print(report)
    assert "grid256" in report and "ysweeps" in report
    nrun, _ = b.run(7)
    assert nrun == len(specs)
    report = b.autotune(1)  # on a running batch
    nrun, _ = b.run()
    assert nrun == 0
    for i in range(len(specs)):
        bad = same(b.problem(i).download(), ref.problem(i).download())
        assert not bad, (i, bad)
    b.close()
    ref.close()


def test_batch_problem_contexts_refuse_stepping(gpu):
    specs = specs_for(SHAPES[:2])
    b = cases.riemann_batch(gpu, specs, 0.01)
    p = b.problem(0)
    for call in (lambda: p.step(), lambda: p.run(3), lambda: p.entry("save_values_at_n"),
                 lambda: p.autotune()):
        with pytest.raises(abi.MahycoError) as e:
            call()
        assert e.value.code == abi.E_STATE
    b.run(3)
    a = p.download()
    with pytest.raises(abi.MahycoError) as e:
        p.upload(a)
    assert e.value.code == abi.E_STATE
    with pytest.raises(abi.MahycoError) as e:
        p.init_cells(*cases.riemann_fields(specs[0]))
    assert e.value.code == abi.E_STATE
    assert b.status(0)["iterations"] == 3
    b.close()


def test_batch_mass_conservation(gpu):
    """FP64: every problem's total mass drifts by round-off only."""
    specs = specs_for(SHAPES, seed=8)
    b = cases.riemann_batch(gpu, specs, 0.03)
    m0 = [b.problem(i).conservation()[0] for i in range(len(specs))]
    b.run()
    for i in range(len(specs)):
        m1 = b.problem(i).conservation()[0]
        print(i, (m1 - m0[i]) / m0[i])
        assert abs(m1 - m0[i]) <= 1e-11 * m0[i], (i, m0[i], m1)
    b.close()
