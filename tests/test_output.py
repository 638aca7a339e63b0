"""Checkpoint / restart and EnSight output (mahyco_amd.output).

Restart: a run interrupted by a checkpoint and continued in a new context
ends bit for bit where the uninterrupted run ends (the dump and
hydroContinueInit, MahycoModule.cc:277-305), in 2D with three environments
and in 3D.  EnSight: the files written for a few steps read back to the
downloaded fields (to the 6 digits of the format), with the counts and time
values of the case file."""
import os

import numpy as np
import pytest

import helpers as H
from mahyco_amd import output

FIELDS = ["density", "internal_energy", "pressure", "velocity", "node_coord", "fracvol_env",
          "dual_grad_phi", "env_mask"]


def _same(a, b):
    bad = [f for f in FIELDS if not np.array_equal(a[f], b[f])]
    bad += [k for k in output.SCALARS if a[k] != b[k]]
    return bad


@pytest.mark.parametrize("case", ["tp", "sod3d"])
def test_restart_continues_bit_for_bit(tmp_path, case):
    L = H.oracle()
    setup = (lambda: H.tp_setup(L, 28, 12)) if case == "tp" else (lambda: H.sod3d_setup(L, 20, 3))
    a = setup()
    a.run(15)
    path = str(tmp_path / "ck.npz")
    output.checkpoint(a, path)
    a.run(10)
    b = output.restart(L, path)
    b.run(10)
    bad = _same(a.download(), b.download())
    assert not bad, bad
    a.close()
    b.close()


def test_checkpoint_holds_no_pickles(tmp_path):
    L = H.oracle()
    s = H.sod_setup(L, 20, 2)
    path = str(tmp_path / "ck.npz")
    output.checkpoint(s, path)
    with np.load(path, allow_pickle=False) as z:
        assert all(z[k].dtype != object for k in z.files)
    s.close()


def _read_floats(path, skip):
    with open(path) as f:
        lines = f.read().split("\n")[skip:]
    return np.array([float(x) for x in lines if x.strip()])


def test_ensight_files(tmp_path):
    L = H.oracle()
    nx, ny = 20, 3
    s = H.sod_setup(L, nx, ny, cas=13, nb_env=2)
    w = output.EnsightWriter(s, str(tmp_path), "sod")
    for _ in range(3):
        s.run(4)
        w.write_step()
    st = s.download()
    d = tmp_path
    case = (d / "sod.case").read_text()
    assert "number of steps: 3" in case and "model: sod.geo" in case
    assert f"{st['global_time']:.9e}" in case
    geo = (d / "sod.geo").read_text().split("\n")
    assert geo[8].split() == [str(nx + 1), str(ny + 1), "1"]
    xyz = np.array([float(v) for v in geo[9:] if v.strip()]).reshape(3, -1)
    assert np.allclose(xyz[:2], np.asarray(st["node_coord"]), rtol=1e-5, atol=1e-7)
    rho = _read_floats(d / "sod.00002.density", 4)
    assert np.allclose(rho, st["density"], rtol=1e-5, atol=0)
    f1 = _read_floats(d / "sod.00002.fracvol_1", 4)
    assert np.allclose(f1, np.asarray(st["fracvol_env"])[1], rtol=1e-5, atol=1e-12)
    vel = _read_floats(d / "sod.00002.velocity", 4).reshape(3, -1)
    assert np.allclose(vel[:2], np.asarray(st["velocity"]), rtol=1e-5, atol=1e-9)
    assert os.path.exists(d / "sod.00000.pressure")
    s.close()
