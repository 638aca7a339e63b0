// MhgArcane.cc -- see MhgArcane.h.  Compiled inside MaHyCo (Arcane), not in
// this repository.
#include "MhgArcane.h"

#include <arcane/ArcaneException.h>
#include <arcane/IItemFamily.h>
#include <arcane/materials/IMeshEnvironment.h>
#include <arcane/materials/IMeshMaterial.h>
#include <arcane/materials/MeshMaterialModifier.h>
#include <arcane/materials/MatItemEnumerator.h>

#include <algorithm>
#include <cmath>

namespace {
MhgArcane* g_instance = nullptr;

// Sorted distinct values of x (a Cartesian mesh's node abscissae), tolerance tol.
std::vector<double> distinct(std::vector<double> x, double tol) {
  std::sort(x.begin(), x.end());
  std::vector<double> out;
  for (double v : x)
    if (out.empty() || v - out.back() > tol) out.push_back(v);
  return out;
}
}  // namespace

MhgArcane::MhgArcane(IMesh* mesh, IMeshMaterialMng* mm) : m_mesh(mesh), m_mm(mm) {}

MhgArcane::~MhgArcane() {
  if (m_ctx) mhg_destroy(m_ctx);
  if (g_instance == this) g_instance = nullptr;
}

MhgArcane* MhgArcane::instance() { return g_instance; }

int MhgArcane::check(int rc, const char* what) const {
  if (rc < 0) {
    const char* msg = m_ctx ? mhg_last_error(m_ctx) : "";
    throw FatalErrorException(A_FUNCINFO, String::format("{0}: MI355X library error {1}: {2}", what, rc, msg));
  }
  return rc;
}

// The mesh description from the node coordinates (the Cartesian generator's
// uniform boxes) and the library index of every cell and node.
void MhgArcane::buildIndexMaps(const VariableNodeReal3& node_coord) {
  const Integer dim = m_mesh->dimension();
  std::vector<double> xs[3];
  ENUMERATE_NODE (inode, m_mesh->allNodes()) {
    const Real3 x = node_coord[inode];
    xs[0].push_back(x.x);
    xs[1].push_back(x.y);
    xs[2].push_back(x.z);
  }
  std::vector<double> axis[3];
  m_desc.dim = dim;
  for (int d = 0; d < dim; ++d) {
    const auto mm = std::minmax_element(xs[d].begin(), xs[d].end());
    const double extent = *mm.second - *mm.first;
    axis[d] = distinct(xs[d], 1e-9 * extent);
    m_desc.n[d] = (int32_t)axis[d].size() - 1;
    m_desc.origin[d] = *mm.first;
    m_desc.length[d] = extent;
  }
  const Int64 nnx = m_desc.n[0] + 1, nny = m_desc.n[1] + 1;
  m_nn = nnx * nny * (dim == 3 ? m_desc.n[2] + 1 : 1);
  m_nc = (Int64)m_desc.n[0] * m_desc.n[1] * (dim == 3 ? m_desc.n[2] : 1);
  auto index_of = [&](int d, double v) {
    const double h = m_desc.length[d] / m_desc.n[d];
    return (Int64)std::llround((v - m_desc.origin[d]) / h);
  };
  m_node_index.assign(m_mesh->nodeFamily()->maxLocalId(), -1);
  ENUMERATE_NODE (inode, m_mesh->allNodes()) {
    const Real3 x = node_coord[inode];
    const Int64 i = index_of(0, x.x), j = index_of(1, x.y), k = dim == 3 ? index_of(2, x.z) : 0;
    m_node_index[inode.localId()] = i + nnx * (j + nny * k);
  }
  m_cell_index.assign(m_mesh->cellFamily()->maxLocalId(), -1);
  ENUMERATE_CELL (icell, m_mesh->allCells()) {
    // the node with the smallest coordinates is the cell's (i, j, k) corner
    Int64 i = -1, j = -1, k = -1;
    for (Node n : (*icell).nodes()) {
      const Real3 x = node_coord[n];
      const Int64 a = index_of(0, x.x), b = index_of(1, x.y), c = dim == 3 ? index_of(2, x.z) : 0;
      if (i < 0 || a < i) i = a;
      if (j < 0 || b < j) j = b;
      if (k < 0 || c < k) k = c;
    }
    m_cell_index[icell.localId()] = i + (Int64)m_desc.n[0] * (j + (Int64)m_desc.n[1] * k);
  }
}

void MhgArcane::create(const mhg_options& opt, const VariableNodeReal3& node_coord) {
  buildIndexMaps(node_coord);
  m_nenv = m_mm->environments().size();
  if (m_nenv != opt.nb_env)
    throw FatalErrorException(A_FUNCINFO, "environment count differs between the options and the material manager");
  check(mhg_create(&m_desc, &opt, &m_ctx), "mhg_create");
  g_instance = this;
}

void MhgArcane::upload(MhgArcaneFields& f, Real global_time, Integer global_iteration) {
  const Int64 NC = m_nc, NN = m_nn, EN = (Int64)m_nenv * NC;
  const Integer dim = m_desc.dim;
  std::vector<double> cell[8], env[8];
  MaterialVariableCellReal* mv[8] = {&f.density, &f.internal_energy, &f.pressure, &f.sound_speed,
                                     &f.pseudo_viscosity, &f.cell_mass, &f.fracvol, &f.mass_fraction};
  for (int t = 0; t < 8; ++t) {
    cell[t].assign(NC, 0.);
    env[t].assign(EN, 0.);
    ENUMERATE_CELL (icell, m_mesh->allCells())
      cell[t][m_cell_index[icell.localId()]] = mv[t]->globalVariable()[icell];
  }
  std::vector<int32_t> mask(NC, 0);
  ENUMERATE_ENV (ienv, m_mm) {
    IMeshEnvironment* e = *ienv;
    const Integer k = e->id();
    ENUMERATE_ENVCELL (ienvcell, e) {
      EnvCell ev = *ienvcell;
      const Int64 c = m_cell_index[ev.globalCell().localId()];
      mask[c] |= 1 << k;
      for (int t = 0; t < 8; ++t) env[t][(size_t)k * NC + c] = (*mv[t])[ev];
    }
  }
  std::vector<double> vel((size_t)dim * NN, 0.), xyz((size_t)dim * NN, 0.), nm(NN, 0.), L(NC, 0.);
  ENUMERATE_NODE (inode, m_mesh->allNodes()) {
    const Int64 n = m_node_index[inode.localId()];
    const Real3 u = f.velocity[inode], x = f.node_coord[inode];
    for (int d = 0; d < dim; ++d) {
      vel[(size_t)d * NN + n] = d == 0 ? u.x : (d == 1 ? u.y : u.z);
      xyz[(size_t)d * NN + n] = d == 0 ? x.x : (d == 1 ? x.y : x.z);
    }
    nm[n] = f.node_mass[inode];
  }
  ENUMERATE_CELL (icell, m_mesh->allCells()) L[m_cell_index[icell.localId()]] = f.caracteristic_length[icell];
  mhg_state st{};
  st.density = cell[0].data();
  st.internal_energy = cell[1].data();
  st.pressure = cell[2].data();
  st.sound_speed = cell[3].data();
  st.pseudo_viscosity = cell[4].data();
  st.cell_mass = cell[5].data();
  st.fracvol = cell[6].data();
  st.mass_fraction = cell[7].data();
  st.env_mask = mask.data();
  st.density_env = env[0].data();
  st.internal_energy_env = env[1].data();
  st.pressure_env = env[2].data();
  st.sound_speed_env = env[3].data();
  st.pseudo_viscosity_env = env[4].data();
  st.cell_mass_env = env[5].data();
  st.fracvol_env = env[6].data();
  st.mass_fraction_env = env[7].data();
  st.velocity = vel.data();
  st.node_mass = nm.data();
  st.node_coord = xyz.data();
  st.caracteristic_length = L.data();
  st.global_deltat = f.global_deltat();
  st.old_deltat = f.old_deltat();
  st.global_time = global_time;
  st.sens_projection = f.sens_projection();
  st.global_iteration = global_iteration;
  check(mhg_upload_state(m_ctx, &st), "mhg_upload_state");
}

void MhgArcane::download(MhgArcaneFields& f) {
  const Int64 NC = m_nc, NN = m_nn, EN = (Int64)m_nenv * NC;
  const Integer dim = m_desc.dim;
  std::vector<double> cell[8], env[8];
  for (int t = 0; t < 8; ++t) {
    cell[t].assign(NC, 0.);
    env[t].assign(EN, 0.);
  }
  std::vector<int32_t> mask(NC, 0);
  std::vector<double> vel((size_t)dim * NN), xyz((size_t)dim * NN), nm(NN), L(NC);
  mhg_state st{};
  double* cp[8];
  double* ep[8];
  for (int t = 0; t < 8; ++t) {
    cp[t] = cell[t].data();
    ep[t] = env[t].data();
  }
  st.density = cp[0]; st.internal_energy = cp[1]; st.pressure = cp[2]; st.sound_speed = cp[3];
  st.pseudo_viscosity = cp[4]; st.cell_mass = cp[5]; st.fracvol = cp[6]; st.mass_fraction = cp[7];
  st.density_env = ep[0]; st.internal_energy_env = ep[1]; st.pressure_env = ep[2]; st.sound_speed_env = ep[3];
  st.pseudo_viscosity_env = ep[4]; st.cell_mass_env = ep[5]; st.fracvol_env = ep[6]; st.mass_fraction_env = ep[7];
  st.env_mask = mask.data();
  st.velocity = vel.data();
  st.node_mass = nm.data();
  st.node_coord = xyz.data();
  st.caracteristic_length = L.data();
  check(mhg_download_state(m_ctx, &st), "mhg_download_state");

  // environment membership after the remap (RemapADIFinal.cc:145-226)
  {
    MeshMaterialModifier modifier(m_mm);
    ENUMERATE_ENV (ienv, m_mm) {
      IMeshEnvironment* e = *ienv;
      const Integer k = e->id();
      std::vector<bool> member(NC, false);
      ENUMERATE_ENVCELL (ienvcell, e) member[m_cell_index[(*ienvcell).globalCell().localId()]] = true;
      Int32UniqueArray add, remove;
      ENUMERATE_CELL (icell, m_mesh->allCells()) {
        const Int64 c = m_cell_index[icell.localId()];
        const bool now = (mask[c] >> k) & 1;
        if (now && !member[c]) add.add(icell.localId());
        if (!now && member[c]) remove.add(icell.localId());
      }
      IMeshMaterial* mat = e->materials()[0];  // one material per environment in MaHyCo
      if (!add.empty()) modifier.addCells(mat, add);
      if (!remove.empty()) modifier.removeCells(mat, remove);
    }
  }
  MaterialVariableCellReal* mv[8] = {&f.density, &f.internal_energy, &f.pressure, &f.sound_speed,
                                     &f.pseudo_viscosity, &f.cell_mass, &f.fracvol, &f.mass_fraction};
  for (int t = 0; t < 8; ++t) {
    ENUMERATE_CELL (icell, m_mesh->allCells())
      mv[t]->globalVariable()[icell] = cell[t][m_cell_index[icell.localId()]];
    // values of the cells with two or more environments (a pure cell's
    // environment value is its cell value)
    ENUMERATE_ENV (ienv, m_mm) {
      IMeshEnvironment* e = *ienv;
      const Integer k = e->id();
      ENUMERATE_ENVCELL (ienvcell, e) {
        EnvCell ev = *ienvcell;
        const Int64 c = m_cell_index[ev.globalCell().localId()];
        if (__builtin_popcount((unsigned)mask[c]) >= 2) (*mv[t])[ev] = env[t][(size_t)k * NC + c];
      }
    }
  }
  ENUMERATE_NODE (inode, m_mesh->allNodes()) {
    const Int64 n = m_node_index[inode.localId()];
    f.velocity[inode] = Real3(vel[n], vel[NN + n], dim == 3 ? vel[2 * NN + n] : 0.);
    f.node_coord[inode] = Real3(xyz[n], xyz[NN + n], dim == 3 ? xyz[2 * NN + n] : 0.);
    f.node_mass[inode] = nm[n];
  }
  ENUMERATE_CELL (icell, m_mesh->allCells()) f.caracteristic_length[icell] = L[m_cell_index[icell.localId()]];
  f.global_deltat = st.global_deltat;
  f.old_deltat = st.old_deltat;
  f.sens_projection = st.sens_projection;
}

void MhgArcane::saveValuesAtN() {
  check(mhg_begin_iteration(m_ctx), "beginIteration");
  check(mhg_save_values_at_n(m_ctx), "saveValuesAtN");
}

Real MhgArcane::computeDeltaT(bool* stop) {
  double dt = 0.;
  int32_t s = 0;
  check(mhg_compute_delta_t(m_ctx, &dt, &s), "computeDeltaT");  // MHG_E_DT_MIN: dt below deltat-min
  *stop = s != 0;
  return dt;
}

bool MhgArcane::step() {
  int32_t stop = 0;
  check(mhg_step(m_ctx, &stop), "step");
  return stop != 0;
}
