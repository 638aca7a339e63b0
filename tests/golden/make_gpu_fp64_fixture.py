"""Writes the FP64 GPU state of two short runs (default launch variants):
the bit-identity fixture that guards the FP64 path against changes of
operation order.  Run on a GPU, from the repository root:

    python tests/golden/make_gpu_fp64_fixture.py tests/golden/gpu_fp64_state.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "arcane-benchs_amd"))
import helpers as H  # noqa: E402
from mahyco_amd import load_gpu  # noqa: E402

FIELDS = ["density", "internal_energy", "pressure", "sound_speed", "pseudo_viscosity",
          "cell_mass", "velocity", "node_coord", "caracteristic_length", "density_env",
          "internal_energy_env", "fracvol_env", "env_mask"]
RUNS = {"tp56x24_s40": (H.tp_setup, 56, 24, 40), "sod100x4_s120": (H.sod_setup, 100, 4, 120)}


def run_all(lib):
    out = {}
    for name, (setup, nx, ny, steps) in RUNS.items():
        s = setup(lib, nx, ny)
        s.run(steps)
        st = s.download()
        for f in FIELDS:
            out[f"{name}/{f}"] = np.asarray(st[f])
        out[f"{name}/global_deltat"] = np.asarray(st["global_deltat"])
    return out


if __name__ == "__main__":
    np.savez_compressed(sys.argv[1], **run_all(load_gpu()))
