// gpu_dem_model.h -- a model::DEMModel subclass that runs PeriDEM's explicit
// time step on one MI355X through the C-ABI of include/pdem.h.
//
// Add this header to the reference tree (next to src/model/dem/) and link
// peridem_amd/lib/libpdem_hip.so.  The subclass keeps the reference's init()
// (deck reading, particle creation, materials, fixity, setupContact) and
// restart(), hands the resulting ModelData to the engine, and replaces the
// time loop: integrate() (demModel.h:101, demModel.cpp:268-340) advances the
// engine by HIP-graph replays between the steps on which the host has work
// (compressive / two-particle test output every d_dtTestOut steps, output
// every d_dtOut steps, stop criteria every step), and keeps d_n and d_time
// equal to the engine's clock.  ppTwoParticleTest and ppCompressiveTest are
// not virtual in the reference (demModel.h:223,230): the loop calls this
// class's own versions, which read their quantities from the device.
//
// Field arrays stay on the device between host callbacks; output() and the
// stop criteria pull only what they need.  The single-step methods
// (integrateCD, integrateVerlet, computeForces, ...) are overridden as well,
// for callers that drive the step themselves.
#ifndef PDEM_INTEGRATION_GPU_DEM_MODEL_H
#define PDEM_INTEGRATION_GPU_DEM_MODEL_H

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "inp/decks/materialDeck.h"
#include "inp/decks/modelDeck.h"
#include "inp/decks/outputDeck.h"
#include "inp/decks/restartDeck.h"
#include "inp/input.h"
#include "inp/pdecks/contactDeck.h"
#include "inp/pdecks/particleDeck.h"
#include "loading/particleFLoading.h"
#include "loading/particleULoading.h"
#include "model/dem/demModel.h"
#include "particle/baseParticle.h"
#include "util/function.h"
#include "util/geomObjects.h"
#include "util/point.h"

#include "pdem.h"

namespace pdem_binding {

class GpuDEMModel : public model::DEMModel {
public:
  explicit GpuDEMModel(inp::Input *deck, int32_t precision = PDEM_PREC_FP64,
                       std::string name = "GpuDEMModel")
      : model::DEMModel(deck, name), d_precision(precision) {}

  ~GpuDEMModel() {
    if (d_gpu) pdem_destroy(d_gpu);
  }

  /*! Time loop of DEMModel::integrate (demModel.cpp:268-340) on the engine. */
  void integrate() override {
    if (d_n == 0 && d_outputDeck_p->d_performOut) model::DEMModel::output();  // host state is current
    if (d_n == 0) applyInitialCondition();
    createEngine();  // uploads the state after the initial condition or the restart
    if (d_n > 0) {   // DEMModel::restart read x, u, v into ModelData
      check(pdem_set_field(d_gpu, PDEM_FIELD_X, &d_x[0].d_x));
      check(pdem_set_field(d_gpu, PDEM_FIELD_U, &d_u[0].d_x));
      check(pdem_set_field(d_gpu, PDEM_FIELD_V, &d_v[0].d_x));
      check(pdem_restart(d_gpu, (int64_t)d_n));
    }
    // computeExternalDisplacementBC(); computeForces(); of the loop's head
    check(pdem_integrate_begin(d_gpu));
    const bool verlet = d_modelDeck_p->d_timeDiscretization == "velocity_verlet";
    const std::string &test = d_pDeck_p->d_testName;
    const bool pp = test == "two_particle" || test == "compressive_test";
    const bool out = d_outputDeck_p->d_performOut && d_outputDeck_p->d_dtOut > 0;
    const size_t test_every = d_outputDeck_p->d_dtTestOut;
    const size_t nt = d_modelDeck_p->d_Nt;
    while (d_n < nt) {
      // the next step after which the host has work
      size_t next = nt;
      auto upto = [&](size_t every) {
        if (every > 0) next = std::min(next, (d_n / every + 1) * every);
      };
      if (pp) upto(test_every);
      if (out) upto(d_outputDeck_p->d_dtOut);
      if (stopCriterion() != 0) upto(1);
      const int64_t k = (int64_t)(next - d_n);
      check(verlet ? pdem_integrate_verlet_graph(d_gpu, k) : pdem_integrate_cd_graph(d_gpu, k));
      syncClock();
      if (test == "two_particle")
        gpuTwoParticleTest();
      else if (test == "compressive_test")
        gpuCompressiveTest();
      if (out && d_n % d_outputDeck_p->d_dtOut == 0 && d_n >= d_outputDeck_p->d_dtOut) output();
      checkStop();
    }
    if (d_ppFile.is_open()) d_ppFile.close();
  }

  void integrateCD() override {
    createEngine();
    check(pdem_integrate_cd(d_gpu, 1));
    syncClock();
  }
  void integrateVerlet() override {
    createEngine();
    check(pdem_integrate_verlet(d_gpu, 1));
    syncClock();
  }
  void computeForces() override {
    createEngine();
    check(pdem_compute_forces(d_gpu));
  }
  void computePeridynamicForces() override {
    createEngine();
    check(pdem_compute_peridynamic_forces(d_gpu));
  }
  void computeContactForces() override {
    createEngine();
    check(pdem_compute_contact_forces(d_gpu));
  }
  void computeExternalForces() override {
    createEngine();
    check(pdem_compute_external_forces(d_gpu));
  }
  void computeExternalDisplacementBC() override {
    createEngine();
    check(pdem_compute_external_displacement_bc(d_gpu));
  }

  /*! DEMModel::checkStop (demModel.cpp:2151-2200) on the device state. */
  void checkStop() override {
    const int32_t k = stopCriterion();
    if (!k || !d_gpu) return;
    int32_t stop = 0;
    double value = 0.;
    int64_t node = -1;
    check(pdem_check_stop(d_gpu, k, d_outputDeck_p->d_outCriteriaParams[0], &stop, &value, &node));
    if (stop) {
      if (d_ppFile.is_open()) d_ppFile.close();
      exit(1);
    }
  }

  /*! Pull the fields the reference's writer reads, then DEMModel::output. */
  void output() override {
    if (d_gpu) {
      check(pdem_get_field(d_gpu, PDEM_FIELD_X, &d_x[0].d_x));
      check(pdem_get_field(d_gpu, PDEM_FIELD_U, &d_u[0].d_x));
      check(pdem_get_field(d_gpu, PDEM_FIELD_V, &d_v[0].d_x));
      check(pdem_get_field(d_gpu, PDEM_FIELD_F, &d_f[0].d_x));
      check(pdem_get_field(d_gpu, PDEM_FIELD_Z, d_Z.data()));
      check(pdem_get_field(d_gpu, PDEM_FIELD_THETA, d_thetaX.data()));
    }
    model::DEMModel::output();
  }

  /*! Upload ModelData to the engine (the device part of DEMModel::init,
   * demModel.cpp:109-266); called by integrate(), or by a caller that drives
   * the step itself, after init() and the initial condition. */
  void createEngine() {
    if (d_gpu) return;
    const size_t dim = d_modelDeck_p->d_dim;
    // the influence function is process-global: the material of the first
    // particle created wins (material.h:32, :275)
    const inp::MaterialDeck &md0 =
        d_pDeck_p->d_particleZones[d_particlesListTypeAll[0]->d_zoneId].d_matDeck;
    std::vector<pdem_particle> parts;
    for (particle::BaseParticle *p : d_particlesListTypeAll) {
      pdem_particle q{};
      q.glob_start = (int64_t)p->d_globStart;
      q.glob_end = (int64_t)p->d_globEnd;
      q.center_node = (int64_t)(p->d_globStart + p->getCenterNodeId());
      q.zone = (int32_t)p->d_zoneId;
      q.type_index = p->getTypeIndex();
      q.all_dofs_constrained = p->d_allDofsConstrained ? 1 : 0;
      q.compute_force = p->d_computeForce ? 1 : 0;
      const inp::MaterialDeck &md = d_pDeck_p->d_particleZones[p->d_zoneId].d_matDeck;
      q.material_type = materialType(md.d_materialType);
      q.irreversible = md.d_irreversibleBondBreak ? 1 : 0;
      q.density = p->getDensity();
      q.horizon = p->getHorizon();
      q.mesh_size = p->getMeshSize();
      q.bounding_radius = p->d_geom_p->boundingRadius();
      q.Kn = p->d_Kn;
      q.Rc = p->d_Rc;
      materialConstants(md, md0, (int32_t)dim, q.horizon, q.material_type, q.mat);
      parts.push_back(q);
    }
    std::vector<pdem_contact_pair> ct;  // setupContact() has scaled the deck values
    for (const auto &row : d_cDeck_p->d_data)
      for (const auto &c : row) ct.push_back({c.d_contactR, c.d_Kn, c.d_mu, c.d_betan, c.d_kappa});
    std::vector<int64_t> ptid(d_ptId.begin(), d_ptId.end());

    pdem_desc d{};
    d.dim = (int32_t)dim;
    d.precision = d_precision;
    d.multi_particle = d_input_p->isMultiParticle() ? 1 : 0;
    d.influence_fn_type = (int32_t)md0.d_influenceFnType;
    for (size_t k = 0; k < md0.d_influenceFnParams.size() && k < 2; k++)
      d.influence_fn_params[k] = md0.d_influenceFnParams[k];
    d.dt = d_modelDeck_p->d_dt;
    d.n_nodes = (int64_t)d_x.size();
    d.x_ref = &d_xRef[0].d_x;
    d.u = &d_u[0].d_x;
    d.v = &d_v[0].d_x;
    d.vol = d_vol.data();
    d.fix = d_fix.data();
    d.pt_id = ptid.data();
    d.n_particles = (int32_t)parts.size();
    d.particles = parts.data();
    d.n_zones = (int32_t)d_cDeck_p->d_data.size();
    d.contact = ct.empty() ? nullptr : ct.data();
    // computeExternalForces adds gravity when |g| > 1e-8 (demModel.cpp:797)
    const util::Point g = d_pDeck_p->d_gravity;
    if (g.length() > 1.0E-8) {
      d.gravity[0] = g.d_x;
      d.gravity[1] = g.d_y;
      d.gravity[2] = g.d_z;
    }
    d.search_factor = d_pDeck_p->d_pNeighDeck.d_sFactor;
    d.neigh_update_interval = (int64_t)d_pDeck_p->d_pNeighDeck.d_neighUpdateInterval;

    BcStore ub, fb;
    fillBcSets(d_pDeck_p->d_dispDeck, false, ub);
    fillBcSets(d_pDeck_p->d_forceDeck, true, fb);
    d.n_ubc = (int32_t)ub.sets.size();
    d.ubc = ub.sets.empty() ? nullptr : ub.sets.data();
    d.n_fbc = (int32_t)fb.sets.size();
    d.fbc = fb.sets.empty() ? nullptr : fb.sets.data();

    if (pdem_create(&d, &d_gpu) != 0) {
      std::cerr << "pdem_create: " << pdem_last_error(nullptr) << "\n";
      exit(EXIT_FAILURE);  // the reference's error style
    }
  }

  pdem_ctx *engine() { return d_gpu; }

private:
  pdem_ctx *d_gpu = nullptr;
  int32_t d_precision;

  // node lists and amplitudes the pdem_bc_set entries point into
  struct BcStore {
    std::vector<std::vector<int64_t>> nodes;
    std::vector<std::vector<double>> amp;
    std::vector<pdem_bc_set> sets;
  };

  void check(int rc) {
    if (rc != 0) {
      std::cerr << "pdem: " << pdem_last_error(d_gpu) << "\n";
      exit(EXIT_FAILURE);
    }
  }

  void syncClock() {
    int64_t n = 0;
    double t = 0.;
    check(pdem_get_time(d_gpu, &n, &t));
    d_n = (size_t)n;
    d_time = t;
  }

  // the criteria DEMModel::checkStop evaluates (demModel.cpp:2151-2200)
  int32_t stopCriterion() const {
    const std::string &c = d_outputDeck_p->d_outCriteria;
    if (c == "max_node_dist") return PDEM_STOP_MAX_NODE_DIST;
    if (c == "max_particle_dist" && d_pDeck_p->d_testName == "two_particle")
      return PDEM_STOP_MAX_PARTICLE_DIST;
    return 0;
  }

  static int32_t materialType(const std::string &name) {
    if (name == "RNPBond") return PDEM_MAT_RNP;
    if (name == "PDState") return PDEM_MAT_PDSTATE;
    if (name == "PMBBond") return PDEM_MAT_PMB;
    if (name == "PDElasticBond") return PDEM_MAT_PDELASTIC;
    std::cerr << "Error: material type " << name << " is not supported on the GPU path.\n";
    exit(EXIT_FAILURE);
  }

  // pdem_particle.mat from the zone's material deck: the constants the
  // reference's material constructors keep as private members
  void materialConstants(const inp::MaterialDeck &md, const inp::MaterialDeck &md0, int32_t dim,
                         double horizon, int32_t type, double *mat) {
    pdem_material_deck d{};
    d.material_type = type;
    d.dim = dim;
    d.compute_from_classical = md.d_computeParamsFromElastic ? 1 : 0;
    d.influence_fn_type = (int32_t)md0.d_influenceFnType;
    d.n_influence_fn_params = (int32_t)std::min<size_t>(md0.d_influenceFnParams.size(), 2);
    for (int k = 0; k < d.n_influence_fn_params; k++) d.influence_fn_params[k] = md0.d_influenceFnParams[k];
    d.E = md.d_matData.d_E;
    d.K = md.d_matData.d_K;
    d.G = md.d_matData.d_G;
    d.nu = md.d_matData.d_nu;
    d.Gc = md.d_matData.d_Gc;
    d.KIc = md.d_matData.d_KIc;
    d.horizon = horizon;
    d.check_sc_factor = md.d_checkScFactor;
    d.n_bond_potential_params = (int32_t)std::min<size_t>(md.d_bondPotentialParams.size(), 3);
    for (int k = 0; k < d.n_bond_potential_params; k++) d.bond_potential_params[k] = md.d_bondPotentialParams[k];
    double k_contact = 0.;
    if (pdem_material_constants(&d, mat, &k_contact) != 0) {
      std::cerr << "Error: " << pdem_material_last_error() << "\n";
      exit(EXIT_FAILURE);
    }
  }

  static int32_t uTimeFn(const std::string &t) {
    if (t == "constant") return PDEM_UFN_CONSTANT;
    if (t == "linear") return PDEM_UFN_LINEAR;
    if (t == "quadratic") return PDEM_UFN_QUADRATIC;
    if (t == "sin") return PDEM_UFN_SIN;
    if (t == "rotation") return PDEM_UFN_ROTATION;
    return PDEM_UFN_NONE;
  }
  static int32_t fTimeFn(const std::string &t) {
    if (t == "linear") return PDEM_FFN_LINEAR;
    if (t == "linear_step") return PDEM_FFN_LINEAR_STEP;
    if (t == "linear_slow_fast") return PDEM_FFN_LINEAR_SLOW_FAST;
    if (t == "sin") return PDEM_FFN_SIN;
    return PDEM_FFN_CONSTANT;  // ParticleFLoading::apply scales by the slope only
  }

  // The node selection and the spatial amplitude of ParticleULoading::apply
  // (particleULoading.cpp:115-230) and ParticleFLoading::apply
  // (particleFLoading.cpp:85-173) are static, so they are resolved here into
  // node lists; the time function runs on the device.
  void fillBcSets(std::vector<inp::PBCData> &decks, bool force, BcStore &st) {
    st.nodes.resize(decks.size());
    st.amp.resize(decks.size());
    for (size_t s = 0; s < decks.size(); s++) {
      const inp::PBCData &bc = decks[s];
      auto &nodes = st.nodes[s];
      auto &amp = st.amp[s];
      int32_t mask = 0;
      for (size_t d : bc.d_direction) mask |= 1 << (int)(d - 1);
      for (particle::BaseParticle *p : d_particlesListTypeAll) {
        // a zero-displacement set is marked applied on its first apply()
        // call, which is for particle 0: only particle 0 is ever written
        if (!force && bc.d_isDisplacementZero && p != d_particlesListTypeAll[0]) break;
        const bool take = force ? d_fLoading_p->needToProcessParticle(p->getId(), bc)
                                : d_uLoading_p->needToProcessParticle(p->getId(), bc);
        if (!take) continue;
        std::pair<util::Point, util::Point> box =
            bc.d_isRegionActive ? bc.d_regionGeomData.d_geom_p->box() : p->d_geom_p->box();
        for (size_t i = 0; i < p->getNumNodes(); i++) {
          const util::Point x = p->getXRefLocal(i);
          const bool dof = force ? d_fLoading_p->needToComputeDof(x, p->getId(), bc)
                                 : d_uLoading_p->needToComputeDof(x, p->getId(), bc);
          if (!dof) continue;
          nodes.push_back((int64_t)p->getNodeId(i));
          amp.push_back(spatialAmplitude(bc, x, box, force));
        }
      }
      pdem_bc_set b{};
      b.time_fn = force ? fTimeFn(bc.d_timeFnType) : uTimeFn(bc.d_timeFnType);
      b.dir_mask = mask;
      b.is_zero = (!force && bc.d_isDisplacementZero) ? 1 : 0;
      for (size_t k = 0; k < bc.d_timeFnParams.size() && k < 4; k++) b.time_params[k] = bc.d_timeFnParams[k];
      b.n_entries = (int64_t)nodes.size();
      b.nodes = nodes.data();
      b.amp = amp.data();
      st.sets.push_back(b);
    }
  }

  // umax (displacement: the slope d_timeFnParams[0] times the spatial
  // function) or fmax (force: the spatial function; the device multiplies by
  // the time function and the slope)
  static double spatialAmplitude(const inp::PBCData &bc, const util::Point &x,
                                 const std::pair<util::Point, util::Point> &box, bool force) {
    const std::string &t = bc.d_spatialFnType;
    const std::vector<double> &sp = bc.d_spatialFnParams;
    double a = force ? 1.0 : (bc.d_timeFnParams.empty() ? 0. : bc.d_timeFnParams[0]);
    if (t == "hat_x") return sp[0] * util::hatFunction(x.d_x, box.first.d_x, box.second.d_x);
    if (t == "hat_y") return sp[0] * util::hatFunction(x.d_y, box.first.d_y, box.second.d_y);
    const double base = force ? (sp.empty() ? 1.0 : sp[0]) : a;
    if (t == "sin_x") return base * std::sin(M_PI * sp[0] * x.d_x);
    if (t == "sin_y") return base * std::sin(M_PI * sp[0] * x.d_y);
    if (t == "linear_x") return base * sp[0] * x.d_x;
    if (t == "linear_y") return base * sp[0] * x.d_y;
    return a;
  }

  // ppCompressiveTest (demModel.cpp:2202-2250) on the device state; the
  // curve file gets numbers (the reference's format string writes literal
  // text, demModel.cpp:2240)
  void gpuCompressiveTest() {
    const size_t every = d_outputDeck_p->d_dtTestOut;
    if (every == 0 || d_n % every != 0 || d_n < every) return;
    double pen = 0., reac = 0.;
    check(pdem_pp_compressive_test(d_gpu, (int32_t)d_pDeck_p->d_particleIdCompressiveTest,
                                   (int32_t)d_pDeck_p->d_particleForceDirectionCompressiveTest - 1, &pen,
                                   &reac));
    if (!d_ppFile.is_open()) {
      const std::string tag = d_outputDeck_p->d_tagPPFile.empty() ? "0" : d_outputDeck_p->d_tagPPFile;
      const std::string fn = d_outputDeck_p->d_path + "pp_" + d_pDeck_p->d_testName + "_" + tag + ".csv";
      d_ppFile.open(fn.c_str(), std::ofstream::out | std::ofstream::app);
      d_ppFile << "t, delta, force \n";
    }
    char line[96];
    std::snprintf(line, sizeof(line), "%.6e, %.6e, %.6e\n", d_time, pen, reac);
    d_ppFile << line;
    setKeyData("wall_penetration", pen);
    setKeyData("tot_reaction_force", reac);
  }

  // ppTwoParticleTest (demModel.cpp:2096-2149) quantities on the device state
  void gpuTwoParticleTest() {
    const size_t every = d_outputDeck_p->d_dtTestOut;
    if (every == 0 || d_n % every != 0 || d_n < every) return;
    double o[5];
    check(pdem_pp_two_particle_test(d_gpu, o));
    setKeyData("pen_dist", o[0]);
    setKeyData("contact_area_radius", o[1]);
    setKeyData("max_y", o[2]);
    setKeyData("max_dist", o[3]);
    setKeyData("max_y_loc", o[4]);
  }
};

}  // namespace pdem_binding

#endif  // PDEM_INTEGRATION_GPU_DEM_MODEL_H
