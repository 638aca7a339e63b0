// MhgArcane.h -- Arcane-side owner of a MI355X context for MaHyCo.
//
// Compiled inside MaHyCo, where Arcane is installed (not built in this
// repository, which has no Arcane).  One MhgArcane per subdomain holds the
// mhg_ctx of the library (include/mahyco_gpu.h), the maps between Arcane's
// local ids and the library's Cartesian numbering (cells i + nx (j + ny k),
// nodes i + (nx+1) (j + (ny+1) k)), and moves the state between the
// module's variables (Mahyco.axl:6-147) and the context:
//   * upload(): after hydroStartInit / hydroContinueInit, or whenever
//     Arcane code has changed the state;
//   * download(): before an output, a checkpoint or any Arcane code that
//     reads the state; it also applies the environment membership changes of
//     the remap with a MeshMaterialModifier (as RemapADIFinal.cc:145-226).
// The entry points of MahycoModule forward to the methods below (the patch
// of MahycoModule.cc is listed in the integration notes); the IRemap and
// IEquationOfState services of this directory reach the same context through
// MhgArcane::instance().
#ifndef MHG_ARCANE_H
#define MHG_ARCANE_H

#include <arcane/IMesh.h>
#include <arcane/ItemTypes.h>
#include <arcane/MeshVariableScalarRef.h>
#include <arcane/materials/IMeshMaterialMng.h>
#include <arcane/materials/MeshMaterialVariableRef.h>

#include <vector>

#include "mahyco_gpu.h"

using namespace Arcane;
using namespace Arcane::Materials;

// The module variables the shim reads and writes (references to the
// ArcaneMahycoObject fields).
struct MhgArcaneFields {
  MaterialVariableCellReal& density;
  MaterialVariableCellReal& internal_energy;
  MaterialVariableCellReal& pressure;
  MaterialVariableCellReal& sound_speed;
  MaterialVariableCellReal& pseudo_viscosity;
  MaterialVariableCellReal& cell_mass;
  MaterialVariableCellReal& fracvol;
  MaterialVariableCellReal& mass_fraction;
  VariableNodeReal3& velocity;
  VariableNodeReal& node_mass;
  VariableNodeReal3& node_coord;
  VariableCellReal& caracteristic_length;
  VariableScalarReal& global_deltat;      // m_global_deltat
  VariableScalarReal& old_deltat;         // m_old_deltat
  VariableScalarInteger& sens_projection;  // m_sens_projection
};

class MhgArcane {
 public:
  MhgArcane(IMesh* mesh, IMeshMaterialMng* mm);
  ~MhgArcane();

  // Library context for the Cartesian mesh (extents from the node
  // coordinates, uniform spacing) and the options.
  void create(const mhg_options& opt, const VariableNodeReal3& node_coord);
  void upload(MhgArcaneFields& f, Real global_time, Integer global_iteration);
  void download(MhgArcaneFields& f);

  mhg_ctx* ctx() const { return m_ctx; }
  // library cell index of an Arcane cell local id
  Int64 cellIndex(Int32 local_id) const { return m_cell_index[local_id]; }
  // Fatal error with the library's message when rc < 0.
  int check(int rc, const char* what) const;

  // The context of the subdomain's mesh (set by create()), for the services.
  static MhgArcane* instance();

  // MahycoModule entry points (Mahyco.config:30-41), each a library call.
  void saveValuesAtN();
  void updateVelocity() { check(mhg_update_velocity(m_ctx), "updateVelocity"); }
  void applyBoundaryCondition() { check(mhg_apply_boundary_condition(m_ctx), "applyBoundaryCondition"); }
  void updatePosition() { check(mhg_update_position(m_ctx), "updatePosition"); }
  void computeGeometricValues() { check(mhg_compute_geometric_values(m_ctx), "computeGeometricValues"); }
  void updateDensity() { check(mhg_update_density(m_ctx), "updateDensity"); }
  void computeArtificialViscosity() { check(mhg_compute_artificial_viscosity(m_ctx), "computeArtificialViscosity"); }
  void updateEnergyAndPressure() { check(mhg_update_energy_and_pressure(m_ctx), "updateEnergyAndPressure"); }
  void computePressionMoyenne() { check(mhg_compute_pression_moyenne(m_ctx), "computePressionMoyenne"); }
  void remap() { check(mhg_remap(m_ctx), "remap"); }
  // computeVariablesForRemap + computeFaceQuantitesForRemap, for a module
  // that keeps remap()'s own sequence and calls the IRemap service
  void remapPrepare() { check(mhg_remap_prepare(m_ctx), "computeVariablesForRemap"); }
  // computeDeltaT: the new dt; *stop when the run reaches its final time.
  Real computeDeltaT(bool* stop);
  // The whole iteration in the fused kernels (MahycoLoop in one call).
  bool step();

 private:
  void buildIndexMaps(const VariableNodeReal3& node_coord);

  IMesh* m_mesh;
  IMeshMaterialMng* m_mm;
  mhg_ctx* m_ctx = nullptr;
  mhg_mesh_desc m_desc{};
  Integer m_nenv = 0;
  // library index of each Arcane local id (cells, nodes)
  std::vector<Int64> m_cell_index, m_node_index;
  Int64 m_nc = 0, m_nn = 0;
};

#endif
