// RemapMhgService.cc -- IRemap (mahyco/src/Remap/IRemap.h:34-65) on the
// MI355X library: each method is the library call of the same name on the
// subdomain's context (MhgArcane::instance()).  Registered like the
// reference's RemapADI service (RemapADIService.cc:1059), selected in a case
// with <remap name="RemapMhg">.  Compiled inside MaHyCo (Arcane), not in this
// repository.
//
// The module's remap() (PrepareRemap.cc:448-563) keeps its order: its
// computeVariablesForRemap / computeFaceQuantitesForRemap forward to
// mhg_remap_prepare, then appliRemap here (the sweeps and remapVariables),
// then applyEOS per environment (EquationOfStateMhgService) and
// computePressionMoyenne.
#include "IRemap.h"
#include "MhgArcane.h"
#include "RemapMhg_axl.h"

class RemapMhgService : public ArcaneRemapMhgObject {
 public:
  explicit RemapMhgService(const ServiceBuildInfo& sbi) : ArcaneRemapMhgObject(sbi) {}

  // The remap part of mhg_options, from this service's options (the module
  // fills the rest from Mahyco.axl's before MhgArcane::create).
  void fillOptions(mhg_options& o) {
    o.remap_threshold = options()->threshold();
    o.is_euler_scheme = options()->isEulerScheme() ? 1 : 0;
    o.ordre_projection = options()->ordreProjection();
    o.conservation_energie_totale = options()->conservationEnergieTotale() ? 1 : 0;
    o.projection_limiteur_id = options()->projectionLimiteurId();
    o.projection_pente_borne = options()->projectionPenteBorne() ? 1 : 0;
    o.projection_pente_borne_mixte = options()->projectionPenteBorneMixte() ? 1 : 0;
    o.projection_limiteur_pure_id = options()->projectionLimiteurPureId();
    o.projection_pente_borne_debar_fix = options()->projectionPenteBorneDebarFix();
    if (options()->calculFluxMasse() != 0)
      throw FatalErrorException(A_FUNCINFO, "calcul_flux_masse != 0 is not run by the MI355X library");
  }

  void appliRemap(Integer dimension, Integer withDualProjection, Integer nb_vars_to_project, Integer nb_env) override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_remap_appli(g->ctx(), dimension, withDualProjection, nb_vars_to_project, nb_env), "appliRemap");
  }
  void resizeRemapVariables(Integer nb_vars_to_project, Integer nb_env) override {
    // the context sizes its remap arrays at creation; the counts are checked
    // once it exists
    MhgArcane* g = MhgArcane::instance();
    if (g) g->check(mhg_remap_resize_variables(g->ctx(), nb_vars_to_project, nb_env), "resizeRemapVariables");
  }
  void synchronizeUremap() override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_remap_synchronize_uremap(g->ctx()), "synchronizeUremap");
  }
  void remapVariables(Integer dimension, Integer withDualProjection, Integer nb_vars_to_project,
                      Integer nb_env) override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_remap_variables(g->ctx(), dimension, withDualProjection, nb_vars_to_project, nb_env),
             "remapVariables");
  }
  Integer getOrdreProjection() override { return options()->ordreProjection(); }
  bool hasProjectionPenteBorne() override { return options()->projectionPenteBorne(); }
  bool hasConservationEnergieTotale() override { return options()->conservationEnergieTotale(); }
  bool isEuler() override { return options()->isEulerScheme(); }
};

ARCANE_REGISTER_SERVICE_REMAPMHG(RemapMhg, RemapMhgService);
