"""Algorithmic HBM bytes of the time step (compulsory traffic, 8-byte words).

Per-zone word counts of the reference step over pure cells with nodes ~ cells
(2D, n_m environments):
  L-node  (velocity, boundary conditions, position)        13
  L-cell  (geometry, density, viscosity, energy, U build)  3 n_m + 9
  cell sweep, per direction                                6 n_m + 7
  node (dual) sweep, per direction                         10
  final cells (moy, EOS, pressure average)                 3 n_m + 7
  final nodes (velocity from the dual, node mass)          7
  dt                                                       1
  total                                                    71 + 18 n_m
giving 1000 B per zone-cycle on the 3-material triple point.  The kernels of
the library map onto these rows; the final-node words are carried by the last
node sweep (which writes the velocity) and the dt word by the final cell
kernel.  The fused step carries the Lagrange cell phase in its first cell
sweep and the final cell phase in its last (the remap vector then never
goes to memory before the first sweep or after the last) and the CFL
candidates in a kernel of their own.  A fused kernel's algorithmic bytes are
the words of the rows it runs; its measured traffic is lower by the remap
vector it keeps in registers.  The step's model stays 71 + 18 n_m.
"""

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def words_per_zone(n_env, dim=2):
    """71 + 18 n_m in 2D; 108 + 24 n_m in 3D (three cell sweeps of 6 n_m + 9
    words, three dual sweeps of 12, L-node 18, L-cell 3 n_m + 10)."""
    return 71 + 18 * n_env if dim == 2 else 108 + 24 * n_env


def bytes_per_zone_cycle(n_env, dim=2, word=8):
    return words_per_zone(n_env, dim) * word


# Coordinate words of the per-zone count (kept in FP64 by the precision form):
# 2D: L-node 6 (previous Lagrangian x, x0, new x), L-cell 2, cell sweeps 2 x 2,
# dual sweeps 2 x 1; 3D: 24.
COORD_WORDS = {2: 14, 3: 24}


def bytes_per_zone_cycle_fp32(n_env, dim=2):
    """Precision form: coordinate words at 8 bytes, every other word at 4
    (556 B on the triple point, 412 B for one material in 2D, 624 B in 3D)."""
    c = COORD_WORDS[dim]
    return c * 8 + (words_per_zone(n_env, dim) - c) * 4


def kernel_bytes(name, n_cells, n_nodes, n_env, word=8):
    """Algorithmic bytes of one launch of a library kernel."""
    w = {
        "node_lagrange": 13 * n_nodes,
        "cell_lagrange": (3 * n_env + 9) * n_cells,
        "cell_sweep_x": (6 * n_env + 7) * n_cells,
        "cell_sweep_y": (6 * n_env + 7) * n_cells,
        "node_sweep_x": 10 * n_nodes,
        "node_sweep_y": 10 * n_nodes,
        "cell_final": (3 * n_env + 7 + 1) * n_cells,
        # a kernel that runs several rows of the model counts their words:
        # the last cell sweep with the final cell phase (the remap vector
        # between them stays in registers, so the kernel moves less than
        # this), the first cell sweep with the Lagrange cell phase (likewise)
        "cell_sweep_x_final": (9 * n_env + 14) * n_cells,
        "cell_sweep_y_final": (9 * n_env + 14) * n_cells,
        "cell_sweep_x_lag": (9 * n_env + 16) * n_cells,
        "cell_sweep_y_lag": (9 * n_env + 16) * n_cells,
        # the dt row: the CFL candidates after the last dual sweep
        "cfl": 1 * n_cells,
    }.get(name)
    return None if w is None else w * word


def last_node_sweep_extra(n_nodes, word=8):
    """The last node sweep also carries the final-node words (velocity, mass)."""
    return 7 * n_nodes * word


TRAFFIC_PROFILE = "triple_point_1792x768_fp64_hbm_traffic.md"


# The fused step alternates the sweep order (x then y, y then x): its first
# cell sweep is cell_sweep_x_lag on one step and cell_sweep_y_lag on the next,
# likewise for the last; each of these pairs runs once per step on average.
FUSED_PAIRS = (("cell_sweep_x_lag", "cell_sweep_y_lag"), ("cell_sweep_x_final", "cell_sweep_y_final"))


def measured_step_bytes(profiles_dir):
    """HBM bytes of one triple-point step from the committed traffic profile
    (the kernels that launch once per step, plus the mean of each fused pair
    of the alternating sweep order); None if absent."""
    tab = _traffic_table(profiles_dir)
    if not tab:
        return None
    paired = {k for pair in FUSED_PAIRS for k in pair}
    tot = sum(v for k, v in tab.items() if k not in paired)
    for pair in FUSED_PAIRS:
        got = [tab[k] for k in pair if k in tab]
        if got:
            tot += sum(got) / len(got)
    return tot


def _traffic_table(profiles_dir):
    import os
    path = os.path.join(profiles_dir, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    tab, in_table = {}, False
    for line in open(path):
        if line.startswith("| label | HBM bytes per launch |"):
            in_table = True
            continue
        if in_table and line.startswith("| ") and not line.startswith("|---"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            tab[cells[0]] = float(cells[1])
    return tab


def measured_traffic(label, profiles_dir):
    """HBM bytes per launch of a step kernel on the 1792x768 triple point, as
    measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counters (copy-calibrated)
    and recorded in the committed traffic profile; None if absent."""
    tab = _traffic_table(profiles_dir)
    return None if not tab else tab.get(label)
