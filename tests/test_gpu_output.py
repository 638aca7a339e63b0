"""Checkpoint / restart of GPU contexts (mahyco_amd.output): an interrupted
triple-point run continued in a new context ends bit for bit where the
uninterrupted one ends, in FP64 and in FP32-bulk; a GPU checkpoint restarted
on the oracle continues one step within 1e-12 of the GPU's own step."""
import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, load_gpu, output

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK])
def test_restart_bit_for_bit(gpu, tmp_path, prec):
    a = H.tp_setup(gpu, 112, 48, precision=prec)
    a.run(40)
    path = str(tmp_path / "ck.npz")
    output.checkpoint(a, path)
    a.run(25)
    b = output.restart(gpu, path)
    b.run(25)
    da, db = a.download(), b.download()
    for f in ["density", "internal_energy", "pressure", "velocity", "fracvol_env", "dual_grad_phi", "env_mask"]:
        assert np.array_equal(da[f], db[f]), f
    for k in output.SCALARS:
        assert da[k] == db[k], k
    a.close()
    b.close()


def test_gpu_checkpoint_restarts_on_oracle(gpu, tmp_path):
    g = H.tp_setup(gpu, 56, 24)
    g.run(30)
    path = str(tmp_path / "ck.npz")
    output.checkpoint(g, path)
    o = output.restart(H.oracle(), path)
    g.step()
    o.step()
    d = H.compare_states(g.download(), o.download(), 3, 56, 24)
    bad = {k: v for k, v in d.items() if v > 1e-12}
    assert not bad, bad
    g.close()
    o.close()
