/*
 * pdem.h -- C-ABI of the MI355X explicit time-step engine for PeriDEM models.
 *
 * The engine owns the per-node state of a PeriDEM model on one GPU and runs
 * the central-difference time step there: peridynamic bond forces (bond-based
 * RNP and state-based PD, with bond breaking and damage), the cell-list
 * contact search, the contact / damping / friction forces, external forces,
 * displacement boundary conditions and the node update.
 *
 * Every entry point below replaces one virtual method of the reference's
 * model::DEMModel (reference file src/model/dem/demModel.h:74-233); the
 * mapping is given per function.  Signatures hold only plain C types and
 * pointers: no HIP or torch types cross this boundary.
 *
 * Conventions
 *  - Return codes are int: 0 = ok, nonzero = error; pdem_last_error() gives
 *    the message.  No exception crosses the ABI.
 *  - Point arrays are AoS [n][3] doubles, the memory layout of the reference's
 *    std::vector<util::Point> (src/util/point.h), so ModelData vectors can be
 *    passed as &v[0].d_x.  In 2D the third component is ignored on input and
 *    written as zero on output.
 *  - The context owns all device memory; host buffers stay owned by the caller
 *    and are copied at create / get / set.  One host thread per context; calls
 *    are stream ordered and every pdem_get_* call synchronises.
 */
#ifndef PDEM_H
#define PDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* material laws (reference src/material/mparticle/material.h) */
#define PDEM_MAT_RNP 0       /* RnpMaterial      material.h:251-659   */
#define PDEM_MAT_PDSTATE 1   /* PdState          material.h:1330-1724 */
#define PDEM_MAT_PMB 2       /* PmbMaterial      material.h:665-1023  */
#define PDEM_MAT_PDELASTIC 3 /* PdElastic        material.h:1025-1324 */

/* influence functions (reference src/material/mparticle/influenceFn.cpp:14-122) */
#define PDEM_INF_CONST 0
#define PDEM_INF_LINEAR 1
#define PDEM_INF_GAUSSIAN 2

/* arithmetic modes */
#define PDEM_PREC_FP64 0   /* all FP64, no FMA contraction: parity mode */
#define PDEM_PREC_MIXED 1  /* FP32 bond/contact math, FP64 sums, FP64 theta/m, FP64 integration */

/* displacement BC time functions (reference src/loading/particleULoading.cpp:115-230) */
#define PDEM_UFN_CONSTANT 0
#define PDEM_UFN_LINEAR 1
#define PDEM_UFN_QUADRATIC 2
#define PDEM_UFN_SIN 3
#define PDEM_UFN_ROTATION 4
#define PDEM_UFN_NONE 5   /* unknown type: u = v = 0 on listed dofs */

/* force BC time functions (reference src/loading/particleFLoading.cpp:85-173) */
#define PDEM_FFN_CONSTANT 0
#define PDEM_FFN_LINEAR 1
#define PDEM_FFN_LINEAR_STEP 2
#define PDEM_FFN_LINEAR_SLOW_FAST 3
#define PDEM_FFN_SIN 4

/* fields for pdem_get_field / pdem_set_field */
#define PDEM_FIELD_X 0      /* double [n][3] current position  (ModelData::d_x)   */
#define PDEM_FIELD_U 1      /* double [n][3] displacement      (d_u)              */
#define PDEM_FIELD_V 2      /* double [n][3] velocity          (d_v)              */
#define PDEM_FIELD_F 3      /* double [n][3] force density     (d_f)              */
#define PDEM_FIELD_XREF 4   /* double [n][3] reference position (d_xRef)          */
#define PDEM_FIELD_VOL 5    /* double [n]    nodal volume      (d_vol)            */
#define PDEM_FIELD_Z 6      /* float  [n]    damage            (d_Z)              */
#define PDEM_FIELD_THETA 7  /* double [n]    dilation          (d_thetaX)         */
#define PDEM_FIELD_MX 8     /* double [n]    weighted volume   (d_mX)             */
#define PDEM_FIELD_VMAG 9   /* double [n]    |v|               (d_vMag)           */
#define PDEM_FIELD_BONDS 10 /* uint8  [nnz]  1 = broken, PD neighbor-list order    */

#define PDEM_NEIGH_PD 0       /* d_neighPd: peridynamic neighbors (same particle)    */
#define PDEM_NEIGH_CONTACT 1  /* d_neighC: contact candidates (other particles)      */

/* One particle or wall.  The reference builds these in
 * particle::BaseParticle::BaseParticle (src/particle/baseParticle.cpp:60-189). */
typedef struct pdem_particle {
  int64_t glob_start;        /* first node (d_globStart) */
  int64_t glob_end;          /* one past last node (d_globEnd) */
  int64_t center_node;       /* global id of the node closest to the geometric centre
                                (refParticle.cpp:36-45 + baseParticle.h:825) */
  int32_t zone;              /* zone id: row/column of the contact table */
  int32_t type_index;        /* 0 particle, 1 wall (d_typeIndex) */
  int32_t all_dofs_constrained;
  int32_t compute_force;     /* d_computeForce */
  int32_t material_type;     /* PDEM_MAT_* */
  int32_t irreversible;      /* RNP: irreversible bond breaking */
  double density;            /* rho */
  double horizon;            /* delta */
  double mesh_size;          /* h = minimum node distance in the particle */
  double bounding_radius;    /* d_geom_p->boundingRadius(): damping reach and mass */
  double Kn;                 /* broken-bond contact stiffness 18 K / (pi delta^5) */
  double Rc;                 /* broken-bond contact radius 0.95 h */
  /* material parameters: RNP   {C, beta, rbar, inv_factor, check_sc_factor}
                          PDState {K, G, s0}
                          PMB {c, s0}   PdElastic {c} */
  double mat[6];
} pdem_particle;

/* Contact data for one zone pair after DEMModel::setupContact()
 * (demModel.cpp:1531-1600): contact_r already multiplied by hMin when the deck
 * gave a factor, Kn already multiplied by Kn_Factor. */
typedef struct pdem_contact_pair {
  double contact_r;
  double Kn;
  double mu;
  double betan;
  double kappa;
} pdem_contact_pair;

/* One displacement or force BC set.  The node selection (particle list / region
 * tests of needToComputeDof) and the spatial amplitude are static, so they are
 * resolved at setup into a node list and a per-node amplitude. */
typedef struct pdem_bc_set {
  int32_t time_fn;           /* PDEM_UFN_* or PDEM_FFN_* */
  int32_t dir_mask;          /* bit d set: dof d (0-based) is prescribed */
  int32_t is_zero;           /* displacement BC: zero-displacement set (applied once) */
  int32_t reserved;
  double time_params[4];     /* d_timeFnParams */
  int64_t n_entries;
  const int64_t *nodes;      /* global node ids, in the reference's traversal order */
  const double *amp;         /* per node umax (U BC) or fmax (F BC) after the spatial fn */
} pdem_bc_set;

typedef struct pdem_desc {
  int32_t dim;               /* 2 or 3 */
  int32_t precision;         /* PDEM_PREC_* */
  int32_t multi_particle;    /* Model.Particle_Sim_Type == Multi_Particle: contact active */
  int32_t influence_fn_type; /* PDEM_INF_*: process-global, first material wins */
  double influence_fn_params[2];
  double dt;
  int64_t n_nodes;
  const double *x_ref;       /* [n][3] */
  const double *u;           /* [n][3] initial displacement (usually 0) */
  const double *v;           /* [n][3] initial velocity (initial condition applied) */
  const double *vol;         /* [n] */
  const uint8_t *fix;        /* [n] bit d = dof d fixed (ParticleULoading::setFixity) */
  const int64_t *pt_id;      /* [n] particle id of each node */
  int32_t n_particles;
  int32_t n_zones;
  const pdem_particle *particles;      /* [n_particles], walls included */
  const pdem_contact_pair *contact;    /* [n_zones][n_zones] */
  double gravity[3];
  double search_factor;               /* Neighbor.Search_Factor */
  int64_t neigh_update_interval;      /* Neighbor.Search_Interval */
  int32_t n_ubc;
  int32_t n_fbc;
  const pdem_bc_set *ubc;
  const pdem_bc_set *fbc;
  int64_t contact_capacity;  /* max contact candidates per node, 0 = automatic */
} pdem_desc;

/* State of the contact-search cadence controller (the reference's
 * d_contNeighUpdateInterval / Counter / SearchRadius, demModel.cpp:1808-1955). */
typedef struct pdem_contact_state {
  int64_t interval;
  int64_t counter;
  double search_radius;
  double max_velocity;
  double max_contact_r;
  int64_t n_rebuilds;
  int64_t overflow;          /* nonzero: a node had more candidates than capacity */
  int64_t capacity;
} pdem_contact_state;

typedef struct pdem_ctx pdem_ctx;

/* The deck values a particle's material constants derive from: the fields
 * of the zone's inp::MaterialDeck (src/inp/decks/materialDeck.h:230-301) and
 * the particle's horizon. */
typedef struct pdem_material_deck {
  int32_t material_type;          /* PDEM_MAT_* (Material.Type) */
  int32_t dim;                    /* 2 or 3 */
  int32_t compute_from_classical; /* d_computeParamsFromElastic */
  int32_t influence_fn_type;      /* PDEM_INF_*: the process-global influence function */
  int32_t n_influence_fn_params;  /* 0: the reference's defaults */
  int32_t n_bond_potential_params;
  double influence_fn_params[2];
  double E, K, G, nu, Gc, KIc;    /* d_matData; negative: not given */
  double horizon;                 /* the particle's horizon delta */
  double check_sc_factor;         /* d_checkScFactor (RNP) */
  double bond_potential_params[3];/* used when compute_from_classical is 0 */
} pdem_material_deck;

/* pdem_particle.mat for one particle, as the reference's material
 * constructors compute their private constants (RnpMaterial, PdState,
 * PmbMaterial, PdElastic: material.h:258-311, 541-620, 674-716, 925-1000,
 * 1034-1076, 1260-1306, 1336-1380, 1580-1700), and K_contact, the bulk
 * modulus of computeMaterialProperties(dim) that gives the particle's Kn
 * (baseParticle.cpp:170-179).  Host code: needs no device and no context.
 * Returns 0, or -1 with the reason in pdem_material_last_error(). */
int pdem_material_constants(const pdem_material_deck *deck, double *mat, double *K_contact);
const char *pdem_material_last_error(void);

/* Build the device model: upload state, build the PD neighbor list on device
 * (replaces DEMModel::updatePeridynamicNeighborlist, demModel.cpp:1659-1694),
 * create the bond-state bits (geometry::Fracture ctor, fracture.cpp:20-49),
 * compute the weighted volume m_x (material::computeStateMx,
 * materialUtil.cpp:211-245) and run the first contact search (the
 * updateContactNeighborlist call inside DEMModel::init, demModel.cpp:179-187). */
int pdem_create(const pdem_desc *desc, pdem_ctx **out);
void pdem_destroy(pdem_ctx *ctx);
const char *pdem_last_error(const pdem_ctx *ctx);  /* ctx may be NULL: last create error */

/* The pre-loop part of DEMModel::integrate (demModel.cpp:276-283):
 * computeExternalDisplacementBC() then computeForces(). */
int pdem_integrate_begin(pdem_ctx *ctx);

/* nsteps x DEMModel::integrateCD (demModel.cpp:349-390): node update, n++,
 * t += dt, computeExternalDisplacementBC(), computeForces(). */
int pdem_integrate_cd(pdem_ctx *ctx, int64_t nsteps);

/* DEMModel::computeForces (demModel.cpp:462-600). */
int pdem_compute_forces(pdem_ctx *ctx);
/* DEMModel::computePeridynamicForces (demModel.cpp:602-789): overwrites f of
 * the nodes of force-computing particles; f of the other nodes stays as it
 * is (the last step's forces). */
int pdem_compute_peridynamic_forces(pdem_ctx *ctx);
/* DEMModel::updateContactNeighborlist (demModel.cpp:1696-1806), including the
 * cadence decision of updateContactNeighborSearchParameters (:1808-1955). */
int pdem_update_contact_neighborlist(pdem_ctx *ctx);
/* DEMModel::computeContactForces (demModel.cpp:820-1066); adds to f. */
int pdem_compute_contact_forces(pdem_ctx *ctx);
/* DEMModel::computeExternalForces (demModel.cpp:791-812); adds to f. */
int pdem_compute_external_forces(pdem_ctx *ctx);
/* DEMModel::computeExternalDisplacementBC (demModel.cpp:814-818). */
int pdem_compute_external_displacement_bc(pdem_ctx *ctx);

/* Capture one integrateCD step in a HIP graph and replay it nsteps times.
 * Same results as pdem_integrate_cd; fewer launches. */
int pdem_integrate_cd_graph(pdem_ctx *ctx, int64_t nsteps);

/* nsteps x DEMModel::integrateVerlet (demModel.cpp:392-460), the
 * "velocity_verlet" time scheme: v += dt/2 f/rho and drift, n++, t += dt,
 * computeExternalDisplacementBC(), computeForces(), v += dt/2 f/rho. */
int pdem_integrate_verlet(pdem_ctx *ctx, int64_t nsteps);
/* The same through a captured step (replaces a captured integrateCD step). */
int pdem_integrate_verlet_graph(pdem_ctx *ctx, int64_t nsteps);

int pdem_synchronize(pdem_ctx *ctx);

int pdem_get_field(pdem_ctx *ctx, int32_t field, void *host);
int pdem_set_field(pdem_ctx *ctx, int32_t field, const void *host);
int pdem_get_time(pdem_ctx *ctx, int64_t *n, double *t);
int pdem_get_neighbor_count(pdem_ctx *ctx, int32_t kind, int64_t *nnz);
/* rowptr [n+1], cols [nnz]; candidate lists are in search order */
int pdem_get_neighbors(pdem_ctx *ctx, int32_t kind, int64_t *rowptr, int64_t *cols);
int pdem_get_contact_state(pdem_ctx *ctx, pdem_contact_state *st);

/* DEMModel::ppCompressiveTest (demModel.cpp:2202-2250) quantities:
 * penetration (x - xRef)[f_dir] of the wall's first node and the reaction
 * sum_i f_i[f_dir] V_i over the wall's nodes; f_dir is 0-based. */
int pdem_pp_compressive_test(pdem_ctx *ctx, int32_t wall_id, int32_t f_dir,
                             double *penetration, double *reaction);
/* DEMModel::ppTwoParticleTest (demModel.cpp:2096-2149) quantities of the
 * first two particles: out5 = {pen_dist, contact_area_radius, max_y,
 * max_dist, max_y_loc} (the reference's key data of the same names; the
 * reference's check_dt cadence stays with the caller). */
int pdem_pp_two_particle_test(pdem_ctx *ctx, double *out5);
/* DEMModel::checkStop (demModel.cpp:2151-2200): the quantity of an output
 * criterion and whether it exceeds param (relative compare).
 * PDEM_STOP_MAX_PARTICLE_DIST ("max_particle_dist", two-particle test):
 * distance of the first two particles' centre nodes; node = -1.
 * PDEM_STOP_MAX_NODE_DIST ("max_node_dist"): the largest |x| over all nodes
 * and the first node holding it. The reference exits the process when stop
 * is set; here the caller decides. */
#define PDEM_STOP_MAX_PARTICLE_DIST 1
#define PDEM_STOP_MAX_NODE_DIST 2
int pdem_check_stop(pdem_ctx *ctx, int32_t criterion, double param, int32_t *stop, double *value,
                    int64_t *node);
/* DEMModel::restart (demModel.cpp:84-103): step counter = step, time =
 * step * dt, and the contact-search controller's restart rule
 * (demModel.cpp:1833-1837). Positions, displacements and velocities from the
 * restart file are set with pdem_set_field (as VtkParticleReader::readNodes
 * does, vtkParticleReader.cpp:41-115); call before pdem_integrate_begin.
 * Bond states and damage are not part of the reference's restart (see
 * pdem_checkpoint_* for an exact one). */
int pdem_restart(pdem_ctx *ctx, int64_t step);
/* number of broken bonds and of all bonds (damage curve) */
int pdem_broken_bonds(pdem_ctx *ctx, int64_t *broken, int64_t *total);

/* Strain_Stress post-processing (DEMModel::output, demModel.cpp:2026-2072:
 * fe::getCurrentQuadPoints and fe::getStrainStress, meshUtil.cpp:176-420) at
 * the quadrature points of every particle's mesh.
 * pdem_fe_setup: the element-node connectivity of all particle meshes,
 * particle after particle in mesh order, as global node ids: linear
 * triangles (dim 2, 3 ids per element) or tetrahedra (dim 3, 4 ids);
 * elem_particle [n_elem] the owning particle; elastic [n_particles][4] =
 * {nu, lambda, mu, is_plane_strain} of computeMaterialProperties(dim);
 * quad_order 1-5 (triangles) or 1-3 (tetrahedra), Quad_Approximation_Order.
 * Point k of element e is quadrature point e * points_per_element + k.
 * pdem_strain_stress: current point positions xq [nq][3] and strain, stress
 * [nq][6] in the order xx yy zz yz xz xy, single precision like the
 * reference's SymMatrix3, at the current displacement (any output may be
 * NULL).  The reference's normal strain sums u_i + dN_i/dx_i
 * (meshUtil.cpp:364-372); the output keeps that. */
int pdem_fe_setup(pdem_ctx *ctx, int32_t quad_order, int64_t n_elem, const int64_t *elem_nodes,
                  const int32_t *elem_particle, const double *elastic);
int pdem_fe_quad_count(pdem_ctx *ctx, int64_t *nq);
int pdem_strain_stress(pdem_ctx *ctx, double *xq, float *strain, float *stress);

/* Timing of the last pdem_integrate_cd / _graph call's kernels: for the
 * dominant (PD bond) kernel, average duration in ms measured with HIP events
 * on the stream it was launched on. */
int pdem_profile_enable(pdem_ctx *ctx, int32_t on);
int pdem_profile_get(pdem_ctx *ctx, double *pd_kernel_ms_avg, int64_t *pd_kernel_launches);

/* Launch-shape autotuner for the peridynamic force pass. No counterpart in
 * the reference (its OpenMP loops have no launch shape). Times `reps` force
 * passes on the current state per candidate (FP64 kernels: workgroup size
 * 64/128/256 of the lane-per-node kernel, and for models whose
 * force-computing particles are all RNP the pair-tile kernel, reported as
 * pd_tpb 512; mixed tile kernels: workgroup 256/1024 x tile 256/512/1024
 * nodes), keeps the fastest and re-captures the step graph. Results do not
 * depend on the choice: every kernel sums each node's bonds in list order.
 * pdem_set_launch_config applies a shape directly (e.g. one the tuner chose
 * in an earlier run). */
int pdem_autotune(pdem_ctx *ctx, int32_t reps, double *best_ms);
int pdem_get_launch_config(pdem_ctx *ctx, int32_t *pd_tpb, int32_t *mixed_tpb, int32_t *tile_nodes);
int pdem_set_launch_config(pdem_ctx *ctx, int32_t pd_tpb, int32_t mixed_tpb, int32_t tile_nodes);

/* Tiles of the mixed-precision bond kernels under the current launch shape:
 * tiles whose neighbor window is staged in LDS and tiles that gather from
 * HBM (0 and 0 for FP64 models). No counterpart in the reference (launch
 * shape introspection for tuning and tests). */
int pdem_get_tile_counts(pdem_ctx *ctx, int32_t *lds_tiles, int32_t *hbm_tiles);

/* Contact kernel selection: at each candidate rebuild, nodes with at most
 * light_max candidates (0..16) go to the lane-per-node contact kernel, the
 * others to the wavefront-per-node kernel. Default 16 for 3D models and for
 * 2D models of 2^20 nodes or more (many nodes in contact), else 0 (few
 * contacts: one launch less per step). Takes effect at the next rebuild;
 * results do not depend on it (both kernels sum a node's active pairs in
 * ascending node id). No counterpart in the reference (kernel selection). */
int pdem_set_contact_light_max(pdem_ctx *ctx, int32_t light_max);
/* Step fusion for multi-particle models: the node update also applies the
 * displacement BC and packs the particle centres, the controller runs as one
 * workgroup and the damping pass finalises the nodes, which saves launches
 * where the step is launch-bound.  on: 0 never; 1 (default) all of it for
 * models of up to 2^20 nodes, and for larger models the node update and
 * controller only, where no displacement-BC set uses the sine or rotation
 * time function (the finalisation stays with the next node update); 2 all
 * of it at any size.  Results are the same bit for bit in every mode.
 * No counterpart in the reference. */
int pdem_set_step_fusion(pdem_ctx *ctx, int32_t on);
/* Fused state-based passes (mixed precision, on by default): a particle
 * whose nodes all fit one LDS tile runs its dilation and force passes in one
 * workgroup (window staged once, dilation terms kept in LDS); with 0 every
 * tile runs the two-pass kernels.  Results are the same bit for bit either
 * way. No counterpart in the reference (kernel selection). */
int pdem_set_state_fusion(pdem_ctx *ctx, int32_t on);

/* ---------------------------------------------------------------------
 * Learned pair-contact closure network.  No counterpart in the reference:
 * a 16 -> 128 -> 128 -> 128 -> 4 perceptron (ReLU) whose products run on
 * BF16 MFMA with FP32 accumulation; FP32 master weights and Adam state.
 * Parameters: row-major W1[16][128], b1[128], W2[128][128], b2, W3, b3,
 * W4[128][4], b4[4] (pdem_mlp_param_count() floats).  Inputs are normalised
 * as (x - in_mean) * in_scale, targets as t * out_scale; the loss is
 * mean over rows of sum_o out_weight[o] (y_o - t_o out_scale_o)^2.
 * All buffers are host memory; n rows of 16 (inputs) or 4 (outputs) floats.
 * ------------------------------------------------------------------- */
#define PDEM_MLP_IN 16
#define PDEM_MLP_HIDDEN 128
#define PDEM_MLP_OUT 4

typedef struct pdem_mlp pdem_mlp;

int pdem_mlp_create(uint64_t seed, pdem_mlp **out);
int pdem_mlp_destroy(pdem_mlp *m);
const char *pdem_mlp_last_error(const pdem_mlp *m);
int64_t pdem_mlp_param_count(void);
int pdem_mlp_get_params(pdem_mlp *m, float *params);
int pdem_mlp_set_params(pdem_mlp *m, const float *params);   /* resets Adam state */
int pdem_mlp_set_normalization(pdem_mlp *m, const float *in_mean, const float *in_scale,
                               const float *out_scale, const float *out_weight);
int pdem_mlp_get_normalization(pdem_mlp *m, float *norm40); /* mean|scale|out_scale|out_weight */
/* y = network(x), de-normalised (y_n / out_scale) */
int pdem_mlp_forward(pdem_mlp *m, const float *x, int64_t n, float *y);
/* gradient of the loss over exactly these n rows (no update) */
int pdem_mlp_gradient(pdem_mlp *m, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* iters Adam steps on minibatches drawn with replacement (seeded), cosine
 * learning-rate decay to 2 %; loss_curve (may be NULL) gets iters losses */
int pdem_mlp_train(pdem_mlp *m, const float *x, const float *t, int64_t n, int32_t iters, int32_t batch,
                   float lr, uint64_t seed, float *loss_curve);

/* ---------------------------------------------------------------------
 * Learned pair-contact closure (2D multi-particle models).  No counterpart in
 * the reference: for particle pairs in light, undamaged contact the node-level
 * contact terms of computeContactForces (demModel.cpp:832-934) are replaced
 * by the network's pair force and moment, spread over the contact patch.
 *  setup:   particle_code[n_particles] is feature 11/12 of the pairs (shape
 *           or material class); may be NULL (all 0).
 *  enable:  pairs with penetration Rc - gap <= pen_max_h * min(h_a, h_b), no
 *           broken bonds (max Z < 1) and predicted local damage <= dz_max use
 *           the network; all others keep the plain kernels.  net may be NULL
 *           (or pen_max_h <= 0): then no pair uses the network and the step
 *           equals the plain step.  Every check_every steps (0: never) the
 *           plain pair force of each closure pair is also evaluated and the
 *           difference accumulated.  The network must outlive the context's
 *           use of it and stay on the context's device.
 *  stats:   pairs per mode since enable (0 no contact possible, 1 closure,
 *           2 plain), checked pairs, mean and max |F_closure - F_plain| and
 *           mean |F_plain|, in units of Kn Rc Vbar_a Vbar_b.
 *  log:     training rows of the current state: 16 features, 4 targets
 *           (plain F.n, F.t, M_a / h_a in those units, local max Z) and 3
 *           baseline values (the normal-contact part of the first three
 *           computed from the contact patches; the network predicts target
 *           minus baseline; baseline may be NULL) for every pair with
 *           gap < Rc + band_h * min(h_a, h_b); *n gets the number of rows
 *           found (rows beyond cap are dropped).
 * ------------------------------------------------------------------- */
int pdem_closure_setup(pdem_ctx *ctx, const float *particle_code);
int pdem_closure_enable(pdem_ctx *ctx, const pdem_mlp *net, double pen_max_h, double dz_max,
                        int32_t check_every);
int pdem_closure_disable(pdem_ctx *ctx);
int pdem_closure_stats(pdem_ctx *ctx, int64_t *pairs_by_mode3, int64_t *checked, double *err_mean,
                       double *err_max, double *plain_mean);
int pdem_pair_log(pdem_ctx *ctx, double band_h, int64_t cap, float *features, float *targets, float *baseline,
                  int64_t *n);

/* pdem_pair_log without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight logging loops, hosts that forbid hipMalloc).  The
 * rows are staged in a library-owned device array of
 * PDEM_PAIR_LOG_NOALLOC_MAX_ROWS rows (6.0 MB per GPU, held until the process
 * ends) instead of buffers allocated per call; cap above that maximum is an
 * error.  Same rows as pdem_pair_log (in the order the kernel appends them,
 * which may differ between calls).  Every call on one GPU shares the array:
 * calls that overlap on one GPU, from several contexts or threads, race. */
#define PDEM_PAIR_LOG_NOALLOC_MAX_ROWS 65536
int pdem_pair_log_noalloc(pdem_ctx *ctx, double band_h, int64_t cap, float *features, float *targets,
                          float *baseline, int64_t *n);

/* Checkpoint / restore of the complete step state (node fields, bond bits,
 * damage, the cadence controller's counters, the candidate and partner lists
 * of the last rebuild, zero-set flags): resuming from a checkpoint continues
 * bit for bit.  The reference's restart (vtkParticleReader.cpp:41-115) reads
 * positions and velocities only.  The buffer is host memory of
 * pdem_checkpoint_size() bytes; load requires a context built from the same
 * descriptor (and the same closure set-up). */
int pdem_checkpoint_size(pdem_ctx *ctx, int64_t *bytes);
int pdem_checkpoint_save(pdem_ctx *ctx, void *buf, int64_t bytes);
int pdem_checkpoint_load(pdem_ctx *ctx, const void *buf, int64_t bytes);

/* number of HIP devices visible to the engine (0 when none); no device work */
int pdem_device_count(void);

/* sizeof of the ABI structures, for binding checks:
 * 0 pdem_desc, 1 pdem_particle, 2 pdem_contact_pair, 3 pdem_bc_set, 4 pdem_contact_state,
 * 5 pdem_material_deck */
int64_t pdem_abi_sizeof(int32_t which);

#ifdef __cplusplus
}
#endif

#endif /* PDEM_H */
