// EquationOfStateMhgService.cc -- IEquationOfState
// (mahyco/src/eos/IEquationOfState.h:32-64) on the MI355X library, one
// service instance per environment as MaHyCo's environment options create
// them.  Compiled inside MaHyCo (Arcane), not in this repository.
#include "IEquationOfState.h"
#include "MhgArcane.h"
#include "EquationOfStateMhg_axl.h"

class EquationOfStateMhgService : public ArcaneEquationOfStateMhgObject {
 public:
  explicit EquationOfStateMhgService(const ServiceBuildInfo& sbi) : ArcaneEquationOfStateMhgObject(sbi) {}

  // This environment's entry of mhg_options.env (before MhgArcane::create).
  void fillOptions(mhg_env_desc& e) {
    e.eos = options()->law() == EosStiffenedGas ? MHG_EOS_STIFFENED_GAS : MHG_EOS_PERFECT_GAS;
    e.adiabatic_cst = options()->adiabaticCst();
    e.limit_tension = options()->limitTension();
  }

  void initEOS(IMeshEnvironment* env) override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_eos_init(g->ctx(), env->id()), "initEOS");
  }
  void applyEOS(IMeshEnvironment* env) override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_eos_apply(g->ctx(), env->id()), "applyEOS");
  }
  void applyOneCellEOS(IMeshEnvironment* env, EnvCell ev) override {
    MhgArcane* g = MhgArcane::instance();
    g->check(mhg_eos_apply_one_cell(g->ctx(), env->id(), g->cellIndex(ev.globalCell().localId())),
             "applyOneCellEOS");
  }
  Real getAdiabaticCst(IMeshEnvironment* env) override {
    MhgArcane* g = MhgArcane::instance();
    double v = 0.;
    if (g) g->check(mhg_eos_get_adiabatic_cst(g->ctx(), env->id(), &v), "getAdiabaticCst");
    else v = options()->adiabaticCst();
    return v;
  }
  Real getTensionLimitCst(IMeshEnvironment* env) override {
    MhgArcane* g = MhgArcane::instance();
    double v = 0.;
    if (g) g->check(mhg_eos_get_tension_limit_cst(g->ctx(), env->id(), &v), "getTensionLimitCst");
    else v = options()->limitTension();
    return v;
  }
};

ARCANE_REGISTER_SERVICE_EQUATIONOFSTATEMHG(EquationOfStateMhg, EquationOfStateMhgService);
