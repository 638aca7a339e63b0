"""The reference's ctest input tests/golden/Data.1.arc (3D, two
environments, SOD cas-test 13) read by mahyco_amd.arc and run on the GPU at
a reduced 40x3x3 mesh: 30 steps within 1e-10 (max-norm, relative) of the
oracle's run of the same case."""
import os

import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import arc, load_gpu

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_data1_case_matches_oracle():
    from test_gpu_3d import compare3
    case = arc.read_case(os.path.join(HERE, "golden", "Data.1.arc"))
    n = [40, 3, 3]
    g = case.build(load_gpu(), n=n)
    o = case.build(H.oracle(), n=n)
    assert g.run(30) == o.run(30) == 30
    d = compare3(g.download(), o.download(), 2, n, metric=H.rel_max)
    print({k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) < 1e-10, d
    g.close()
    o.close()
