"""3D (hexahedral) GPU path vs the CPU oracle.

One-step parity from oracle states (1e-12 relative per field, as in 2D; the
dual gradients are compared on the domain surface, the only nodes whose value
is read back), trajectories of Sod tubes along each axis, a three-material box
with a seeded velocity field (mixed cells appear after a few steps), mass
conservation and the FP32 precision form."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, load_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def box3(L, n=(14, 10, 8), seed=7, **optkw):
    """Three materials in a box: x < 0.35 material 0; otherwise material 1
    below z = 0.5 and material 2 above; a seeded smooth velocity field."""
    opt = L.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(3))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, 1.5, 0.), (abi.EOS_PERFECT_GAS, 1.4, 0.),
                       (abi.EOS_PERFECT_GAS, 5.0 / 3.0, 0.)])
    opt.final_time = 1e30
    opt.deltat_max = 1e30
    length = [1.0, n[1] / n[0], n[2] / n[0]]
    s = abi.Simulation(L, abi.mesh_desc(3, list(n), [0, 0, 0], length), opt)
    k, j, i = np.indices((n[2], n[1], n[0]))
    xc, zc = (i + 0.5) / n[0], (k + 0.5) / n[2]
    env = np.where(xc < 0.35, 0, np.where(zc < 0.5, 1, 2)).astype(np.int32).reshape(-1)
    rho = np.array([1.0, 1.0, 0.125])[env]
    p = np.array([1.0, 0.1, 0.1])[env]
    rng = np.random.default_rng(seed)
    nn = (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
    kk, jj, ii = np.indices((n[2] + 1, n[1] + 1, n[0] + 1))
    vel = np.zeros((3, nn))
    for d in range(3):
        a, b, c = rng.uniform(0.5, 2.0, 3)
        vel[d] = 0.05 * (np.sin(a * ii + b * jj + c * kk)).reshape(-1)
    s.init_cells(env, rho, p, vel)
    return s


def surface_nodes(n):
    nx, ny, nz = n
    k, j, i = np.indices((nz + 1, ny + 1, nx + 1))
    s = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return s.reshape(-1)


def compare3(got, ref, nenv, n, metric=None):
    d = H.compare_states(got, ref, nenv, metric=metric)
    surf = surface_nodes(n)
    rd = metric or H.rel_diff
    d["dual_grad_phi(surface)"] = (rd(got["dual_grad_phi"][:, surf], ref["dual_grad_phi"][:, surf])
                                   if metric else H.rel_diff(got["dual_grad_phi"][:, surf],
                                                             ref["dual_grad_phi"][:, surf], scale=1.0))
    return d


@pytest.mark.parametrize("case,warm", [("sod0", 0), ("sod1", 5), ("sod2", 9), ("box", 0), ("box", 4),
                                       ("box", 25)])
def test_3d_one_step(gpu, case, warm):
    if case.startswith("sod"):
        axis = int(case[3])
        setup = lambda L: H.sod3d_setup(L, 24, 3, axis)  # noqa: E731
        nenv, n = 1, [3, 3, 3]
        n[axis] = 24
    else:
        setup = lambda L: box3(L)  # noqa: E731
        nenv, n = 3, [14, 10, 8]
    o = setup(H.oracle())
    if warm:
        o.run(warm)
    g = setup(gpu)
    g.upload(o.download())
    o.step()
    g.step()
    d = compare3(g.download(), o.download(), nenv, n)
    print(case, warm, {k: f"{v:.2e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > TOL}
    assert not bad, bad


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_3d_sod_trajectory(gpu, axis):
    o = H.sod3d_setup(H.oracle(), 40, 3, axis)
    g = H.sod3d_setup(gpu, 40, 3, axis)
    no, ng = o.run(100000), g.run(100000)
    assert no == ng
    n = [3, 3, 3]
    n[axis] = 40
    d = compare3(g.download(), o.download(), 1, n, metric=H.rel_max)
    print("3d sod", axis, no, {k: f"{v:.2e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10


def test_3d_box_trajectory_and_mass(gpu):
    o, g = box3(H.oracle()), box3(gpu)
    c0, o0 = g.conservation(), o.conservation()
    assert c0[0] == o0[0]
    o.run(60)
    g.run(60)
    d = compare3(g.download(), o.download(), 3, [14, 10, 8], metric=H.rel_max)
    print("3d box", {k: f"{v:.2e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10
    c1, o1 = g.conservation(), o.conservation()
    drift_g, drift_o = (c1[0] - c0[0]) / c0[0], (o1[0] - o0[0]) / o0[0]
    print("3d box mass drift gpu", drift_g, "oracle", drift_o)
    assert abs(drift_g - drift_o) < 1e-14


def test_3d_fp32_precision_form(gpu):
    """FP32 bulk in 3D: density within 1e-5 relative L1 of the FP64 oracle
    after 60 steps of the three-material box."""
    o = box3(H.oracle())
    g = box3(gpu, precision=abi.PREC_FP32_BULK)
    o.run(60)
    g.run(60)
    ref, got = o.download(), g.download()
    l1 = float(np.sum(np.abs(got["density"] - ref["density"])) / np.sum(np.abs(ref["density"])))
    linf = float(np.max(np.abs(got["density"] - ref["density"])) / np.max(np.abs(ref["density"])))
    print("3d fp32 density L1", l1, "Linf", linf)
    assert l1 < 1e-5


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK])
def test_3d_store_policy_bit_identical(gpu, prec):
    """Ordinary and streaming stores of the 3D step kernels give the same
    state bit for bit."""
    states = []
    for v in (0, 1, 6, 15):
        s = box3(gpu, precision=prec)
        s.select_variant(v, "store_policy")
        s.run(9)
        states.append(s.download())
    for other in states[1:]:
        for f, v in states[0].items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, other[f]), f
            else:
                assert v == other[f], f
