"""The 3D random-field case of the precision form at a size the oracle runs
in seconds: 32^3 cells of the unit cube, the same seed, sigma 0.3 and rms
velocity 0.1 as the 256^3 benchmark case, the correlation length scaled with
the mesh (8 cells at 256^3 -> 1 cell at 32^3), gamma 5/3, p = 1, reflecting
walls.

* FP64: one step from oracle states within 1e-12 per field; a 20-step
  trajectory within 1e-10 (max-norm relative).
* FP32 bulk: 200 steps against the FP64 oracle: relative L1 of the density
  within 1e-5; internal energy, pressure and velocity reported, their L1
  within 5e-5 and every L-infinity within 3e-3."""
import numpy as np
import pytest

import helpers as H
import test_gpu_3d as T3
from mahyco_amd import abi, cases, load_gpu

pytestmark = pytest.mark.gpu
N = 32


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


@pytest.fixture(scope="module")
def fields():
    return cases.p3d_fields(N, corr_cells=8.0 * N / 256)


@pytest.mark.parametrize("warm", [0, 7])
def test_p3d_one_step(gpu, fields, warm):
    o = cases.p3d(H.oracle(), N, fields=fields)
    if warm:
        o.run(warm)
    g = cases.p3d(gpu, N, fields=fields)
    g.upload(o.download())
    o.step()
    g.step()
    d = T3.compare3(g.download(), o.download(), 1, [N, N, N])
    print("p3d one step", warm, {k: f"{v:.1e}" for k, v in d.items()})
    bad = {k: v for k, v in d.items() if v > 1e-12}
    assert not bad, bad


def test_p3d_trajectory(gpu, fields):
    o = cases.p3d(H.oracle(), N, fields=fields)
    g = cases.p3d(gpu, N, fields=fields)
    o.run(20)
    g.run(20)
    d = T3.compare3(g.download(), o.download(), 1, [N, N, N], metric=H.rel_max)
    print("p3d 20 steps", {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10


def test_p3d_fp32_against_oracle(gpu, fields):
    steps = 200
    o = cases.p3d(H.oracle(), N, fields=fields)
    g = cases.p3d(gpu, N, fields=fields, precision=abi.PREC_FP32_BULK)
    o.run(steps)
    g.run(steps)
    ref, got = o.download(), g.download()
    assert ref["global_iteration"] == got["global_iteration"] == steps
    err = {}
    for f in ["density", "internal_energy", "pressure", "velocity"]:
        a, b = np.asarray(got[f], float), np.asarray(ref[f], float)
        err[f] = (float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))),
                  float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    print("p3d FP32 vs oracle after", steps, "steps (L1, Linf):", err)
    # the precision form's contract: density within 1e-5 relative L1; the
    # other fields are reported, with bounds a few times their measured
    # errors (measured on one MI355X: density 2.9e-6 / 1.8e-5; velocity,
    # the largest, 1.3e-5 L1 and 8.4e-4 L-infinity at a node of small speed)
    assert err["density"][0] < 1e-5
    for f, (l1, linf) in err.items():
        assert l1 < 5e-5, (f, l1)
        assert linf < 3e-3, (f, linf)
