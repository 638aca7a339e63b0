"""Checkpoint / restart and EnSight Gold output of a context.

Checkpoint (the dump and `hydroContinueInit`, MahycoModule.cc:277-305): one
`.npz` file (no pickled objects) holds the mesh description and the
`mhg_options` as their C structures' bytes, and the whole state that
`mhg_download_state` returns: cell and environment fields, environment
masks, node velocities, masses and coordinates, the dual gradients kept
between steps, the characteristic lengths, dt, the previous dt, the time, the
sweep order and the iteration count.  `restart` creates a context from the
stored mesh and options and uploads the state, so that the run continues as
it would have without the interruption (bit for bit on the same library and
precision).  Arcane's restart adds 1 to its iteration counter because its
dump is taken before the counter moves; the state here is taken after, so
the count is restored as saved.  What a context holds besides the state (a
learned-limiter network or coefficients, a forced launch variant) is set
again by the caller.

EnSight Gold (the reference's post-processing output, Arcane's EnSight7
writer): `EnsightWriter` writes an ASCII case file, the Euler mesh as one
structured block part, and per step the cell fields (density, pressure,
internal energy, sound speed, and each environment's volume fraction) and
the node velocity, readable by ParaView or EnSight.
"""
import json
import os

import numpy as np

from . import abi

SCALARS = ["global_deltat", "old_deltat", "global_time", "sens_projection", "global_iteration"]
FORMAT_VERSION = 1


def _struct_bytes(s):
    return np.frombuffer(bytes(s), dtype=np.uint8).copy()


def checkpoint(sim, path):
    """Write the context's mesh, options and state to `path` (.npz)."""
    st = sim.download()
    arrays = {k: v for k, v in st.items() if isinstance(v, np.ndarray)}
    meta = {"version": FORMAT_VERSION, "dim": sim.dim, "nenv": sim.nenv,
            "scalars": {k: st[k] for k in SCALARS}}
    np.savez(path, mesh=_struct_bytes(sim.mesh), options=_struct_bytes(sim.opt),
             meta=np.array(json.dumps(meta)), **{"state_" + k: v for k, v in arrays.items()})


def restart(lib, path, **override):
    """A new context on `lib` continuing the run stored in `path`.
    override: mhg_options fields to change (precision, ...)."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("version") != FORMAT_VERSION:
            raise ValueError(f"{path}: checkpoint format {meta.get('version')}, expected {FORMAT_VERSION}")
        mesh = abi.MeshDesc.from_buffer_copy(z["mesh"].tobytes())
        opt = abi.Options.from_buffer_copy(z["options"].tobytes())
        state = {k[len("state_"):]: z[k] for k in z.files if k.startswith("state_")}
    for k, v in override.items():
        setattr(opt, k, v)
    sim = abi.Simulation(lib, mesh, opt)
    # an initialisation prepares what the state does not carry (Euler
    # volumes, boundary node sets, EOS constants); the upload then replaces
    # every field it set
    mask = state["env_mask"].astype(np.int64)
    low = mask & -mask  # lowest set bit; a cell without environments takes environment 0 here
    first_env = np.where(low > 0, np.log2(np.maximum(low, 1)), 0).astype(np.int32)
    sim.init_cells(first_env, np.ascontiguousarray(state["density"]), np.ascontiguousarray(state["pressure"]))
    state.update(meta["scalars"])
    state = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in state.items()}
    sim.upload(state)
    return sim


class EnsightWriter:
    """EnSight Gold ASCII output of one context's steps under `directory`."""

    def __init__(self, sim, directory, name="mahyco"):
        self.sim, self.dir, self.name = sim, directory, name
        self.times = []
        os.makedirs(directory, exist_ok=True)
        st = sim.download(["node_coord"])
        m = sim.mesh
        self.dim = m.dim
        self.n = [m.n[d] for d in range(m.dim)]
        self.nn = [k + 1 for k in self.n] + ([1] if m.dim == 2 else [])
        self._write_geometry(np.asarray(st["node_coord"]))

    @staticmethod
    def _floats(f, values):
        for v in np.asarray(values, dtype=np.float64).ravel():
            f.write(f"{v:12.5e}\n")

    def _write_geometry(self, coord):
        with open(os.path.join(self.dir, self.name + ".geo"), "w") as f:
            f.write("MaHyCo Euler mesh\nstructured block\nnode id off\nelement id off\n")
            f.write(f"part\n{1:10d}\nmesh\nblock\n")
            f.write("".join(f"{k:10d}" for k in self.nn) + "\n")
            for d in range(3):
                self._floats(f, coord[d] if d < self.dim else np.zeros(coord.shape[1]))

    def _write_var(self, fname, desc, values_per_comp):
        with open(os.path.join(self.dir, fname), "w") as f:
            f.write(desc + "\npart\n" + f"{1:10d}\nblock\n")
            for comp in values_per_comp:
                self._floats(f, comp)

    def variables(self, nenv):
        cells = ["density", "pressure", "internal_energy", "sound_speed"]
        cells += [f"fracvol_{k}" for k in range(nenv)]
        return cells, ["velocity"]

    def write_step(self):
        """Write the context's present state as the next time step."""
        st = self.sim.download()
        i = len(self.times)
        self.times.append(float(st["global_time"]))
        cells, nodes = self.variables(self.sim.nenv)
        fv = np.asarray(st["fracvol_env"])
        for v in cells:
            vals = fv[int(v.split("_")[1])] if v.startswith("fracvol_") else np.asarray(st[v])
            self._write_var(f"{self.name}.{i:05d}.{v}", v, [vals])
        vel = np.asarray(st["velocity"])
        comps = [vel[d] if d < self.dim else np.zeros(vel.shape[1]) for d in range(3)]
        self._write_var(f"{self.name}.{i:05d}.velocity", "velocity", comps)
        self._write_case()

    def _write_case(self):
        cells, nodes = self.variables(self.sim.nenv)
        lines = ["FORMAT", "type: ensight gold", "", "GEOMETRY", f"model: {self.name}.geo", "", "VARIABLE"]
        lines += [f"scalar per element: 1 {v} {self.name}.*****.{v}" for v in cells]
        lines += [f"vector per node: 1 {v} {self.name}.*****.{v}" for v in nodes]
        lines += ["", "TIME", "time set: 1", f"number of steps: {len(self.times)}",
                  "filename start number: 0", "filename increment: 1", "time values:"]
        lines += [f"{t:.9e}" for t in self.times]
        with open(os.path.join(self.dir, self.name + ".case"), "w") as f:
            f.write("\n".join(lines) + "\n")
