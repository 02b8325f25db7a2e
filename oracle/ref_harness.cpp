/*
 * ref_harness.cpp -- C entry points around the reference's own hot-path
 * sources, compiled (never copied) from the reference tree by oracle/Makefile
 * into oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: the tests use it to
 * pin the oracle's bond laws, parameter derivations and neighbor sets against
 * the reference's own code.
 *
 * Reference sources used as they lie:
 *   src/material/mparticle/material.h      RnpMaterial, PmbMaterial, PdElastic, PdState
 *   src/material/mparticle/influenceFn.cpp influence functions
 *   src/util/function.cpp                  isGreater / isLess
 *   src/nsearch/nsearch.h + nanoflann      kd-tree radius searches
 *   src/fe/triElem.cpp, tetElem.cpp,       linear element quadrature data
 *   baseElem.cpp, src/util/matrix.cpp
 */
#include "fe/tetElem.h"
#include "fe/triElem.h"
#include "material/mparticle/material.h"
#include "nsearch/nsearch.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {
inp::MaterialDeck makeDeck(const char *type, double E, double K, double G, double nu,
                           double Gc, double rho, int inf_type, const double *inf_params,
                           int n_inf) {
  inp::MaterialDeck deck;
  deck.d_materialType = type;
  deck.d_matData.d_E = E;
  deck.d_matData.d_K = K;
  deck.d_matData.d_G = G;
  deck.d_matData.d_nu = nu;
  deck.d_matData.d_Gc = Gc;
  deck.d_density = rho;
  deck.d_computeParamsFromElastic = true;
  deck.d_influenceFnType = (size_t)inf_type;
  for (int i = 0; i < n_inf; i++) deck.d_influenceFnParams.push_back(inf_params[i]);
  // the influence function is a per-translation-unit global that the first
  // material constructs; reset it so every call sees its own type
  influence_fn = nullptr;
  return deck;
}
}  // namespace

extern "C" {

int ref_is_greater(double a, double b) { return util::isGreater(a, b); }
int ref_is_less(double a, double b) { return util::isLess(a, b); }

/* out: C, beta, rbar, inv_factor, check_sc_factor, K_from_material, J(0.5), moment(dim) */
void ref_rnp_params(double E, double Gc, double rho, double horizon, int dim, int inf_type,
                    const double *inf_params, int n_inf, double *out) {
  auto deck = makeDeck("RNPBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
  material::RnpMaterial m(deck, (size_t)dim, horizon);
  // C, beta, rbar are private; recover them from the public bond law:
  // getSc(1) = rbar, getBondEF(r, s, fs, false) second = J 4 s C beta / invF
  out[2] = m.getSc(1.0);
  double beta = 0.5 / (out[2] * out[2]);
  out[1] = beta;
  double invF = dim == 2 ? horizon * horizon * horizon * M_PI
                         : horizon * horizon * horizon * horizon * 4. * M_PI / 3.;
  out[3] = invF;
  bool fs = false;
  auto ef = m.getBondEF(horizon * 0.5, 1.0, fs, false);  // J(0.5) 4 C beta / invF
  double J = m.getInfFn(horizon * 0.5);
  out[0] = ef.second * invF / (J * 4. * beta);
  out[4] = deck.d_checkScFactor;
  out[5] = m.computeMaterialProperties((size_t)dim).d_K;
  out[6] = J;
  out[7] = m.getMoment((size_t)dim);
}

/* RnpMaterial::getBondEF: returns fs after the call; e, f in out[0..1] */
int ref_rnp_bond_ef(double E, double Gc, double rho, double horizon, int dim, int inf_type,
                    const double *inf_params, int n_inf, double r, double s, int fs_in,
                    int break_bonds, double *out) {
  auto deck = makeDeck("RNPBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
  material::RnpMaterial m(deck, (size_t)dim, horizon);
  bool fs = fs_in != 0;
  auto ef = m.getBondEF(r, s, fs, break_bonds != 0);
  out[0] = ef.first;
  out[1] = ef.second;
  return fs ? 1 : 0;
}

/* RNP getS / PdState getS */
double ref_rnp_get_s(const double *dx, const double *du) {
  inp::MaterialDeck deck = makeDeck("RNPBond", 1e9, -1., -1., -1., 1., 1., 0, nullptr, 0);
  material::RnpMaterial m(deck, 3, 1.);
  return m.getS(util::Point(dx[0], dx[1], dx[2]), util::Point(du[0], du[1], du[2]));
}

/* out: K, G, s0, nu, E (PdState after computeParameters), J(0.5) */
void ref_pdstate_params(double E, double K, double G, double nu, double Gc, double rho,
                        double horizon, int dim, int inf_type, const double *inf_params,
                        int n_inf, double *out) {
  auto deck = makeDeck("PDState", E, K, G, nu, Gc, rho, inf_type, inf_params, n_inf);
  material::PdState m(deck, (size_t)dim, horizon);
  auto md = m.computeMaterialProperties((size_t)dim);
  out[0] = md.d_K;
  out[1] = md.d_G;
  out[2] = m.getSc(1.0);
  out[3] = md.d_nu;
  out[4] = md.d_E;
  out[5] = m.getInfFn(0.5 * horizon);
}

/* PdState getBondEF (force part) and getS */
double ref_pdstate_bond_f(double E, double K, double G, double nu, double Gc, double horizon,
                          int dim, double r, double s, double mx, double thetax) {
  auto deck = makeDeck("PDState", E, K, G, nu, Gc, 1., 0, nullptr, 0);
  material::PdState m(deck, (size_t)dim, horizon);
  bool fs = false;
  return m.getBondEF(r, s, fs, mx, thetax).second;
}

double ref_pdstate_get_s(const double *dx, const double *du) {
  inp::MaterialDeck deck = makeDeck("PDState", -1., 1e9, 5e8, -1., 1., 1., 0, nullptr, 0);
  material::PdState m(deck, 3, 1.);
  return m.getS(util::Point(dx[0], dx[1], dx[2]), util::Point(du[0], du[1], du[2]));
}

/* PmbMaterial / PdElastic (bond-based, actual stretch, deformed direction):
 * out: c, s0 (PMB; 0 for PdElastic), K_from_material, J(0.5 horizon).
 * c and s0 are private: s0 = getSc(r) for PMB, c from the unbroken force
 * getBondEF(r, 1, fs, false).second / J(r). */
void ref_bond_params(int law, double E, double Gc, double rho, double horizon, int dim, int inf_type,
                     const double *inf_params, int n_inf, double *out) {
  const double r = 0.5 * horizon;
  bool fs = false;
  if (law == 2) {
    auto deck = makeDeck("PMBBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
    material::PmbMaterial m(deck, (size_t)dim, horizon);
    const double J = m.getInfFn(r);
    out[0] = m.getBondEF(r, 1.0, fs, false).second / J;
    out[1] = m.getSc(r);
    out[2] = m.computeMaterialProperties((size_t)dim).d_K;
    out[3] = J;
  } else {
    auto deck = makeDeck("PDElasticBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
    material::PdElastic m(deck, (size_t)dim, horizon);
    const double J = m.getInfFn(r);
    out[0] = m.getBondEF(r, 1.0, fs, false).second / J;
    out[1] = 0.;
    out[2] = m.computeMaterialProperties((size_t)dim).d_K;
    out[3] = J;
  }
}

/* getBondEF(r, s, fs, true) of PMB (law 2) or PdElastic (law 3): returns fs
 * after the call, e and f in out[0..1]; out[2] = getS(dx, du) of the same
 * class for the given vectors, out[3] = getSc(r) */
int ref_bond_ef(int law, double E, double Gc, double rho, double horizon, int dim, int inf_type,
                const double *inf_params, int n_inf, double r, double s, int fs_in, const double *dx,
                const double *du, double *out) {
  bool fs = fs_in != 0;
  const util::Point a(dx[0], dx[1], dx[2]), b(du[0], du[1], du[2]);
  if (law == 2) {
    auto deck = makeDeck("PMBBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
    material::PmbMaterial m(deck, (size_t)dim, horizon);
    auto ef = m.getBondEF(r, s, fs, true);
    out[0] = ef.first;
    out[1] = ef.second;
    out[2] = m.getS(a, b);
    out[3] = m.getSc(r);
  } else {
    auto deck = makeDeck("PDElasticBond", E, -1., -1., -1., Gc, rho, inf_type, inf_params, n_inf);
    material::PdElastic m(deck, (size_t)dim, horizon);
    auto ef = m.getBondEF(r, s, fs, true);
    out[0] = ef.first;
    out[1] = ef.second;
    out[2] = m.getS(a, b);
    out[3] = m.getSc(r);
  }
  return fs ? 1 : 0;
}

/* Influence-function value J(r) for a type */
double ref_influence(int inf_type, const double *params, int n_params, int dim, double r) {
  std::vector<double> p(params, params + n_params);
  if (inf_type == 0) return material::ConstInfluenceFn(p, (size_t)dim).getInfFn(r);
  if (inf_type == 1) return material::LinearInfluenceFn(p, (size_t)dim).getInfFn(r);
  return material::GaussianInfluenceFn(p, (size_t)dim).getInfFn(r);
}
double ref_influence_moment(int inf_type, const double *params, int n_params, int dim, int i) {
  std::vector<double> p(params, params + n_params);
  if (inf_type == 0) return material::ConstInfluenceFn(p, (size_t)dim).getMoment((size_t)i);
  if (inf_type == 1) return material::LinearInfluenceFn(p, (size_t)dim).getMoment((size_t)i);
  return material::GaussianInfluenceFn(p, (size_t)dim).getMoment((size_t)i);
}

/*
 * Neighbor lists exactly as DEMModel builds them with nanoflann:
 *   include_tag = 1: updatePeridynamicNeighborlist (demModel.cpp:1659-1694),
 *     radius r_per_node[i], same tag, j != i;
 *   include_tag = 0: updateContactNeighborlist search (demModel.cpp:1717-1756),
 *     radius r_per_node[0] for all, other tags, j != i, only nodes with query[i].
 * Results in nanoflann's traversal order (sorted = false).  Two-call protocol:
 * cols == NULL fills rowptr only.
 */
void ref_neighbors(const double *x, int64_t n, const int64_t *tags, const double *r_per_node,
                   const uint8_t *query, int include_tag, int64_t *rowptr, int64_t *cols) {
  std::vector<util::Point> pts(n);
  std::vector<size_t> t(n);
  for (int64_t i = 0; i < n; i++) {
    pts[i] = util::Point(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    t[i] = (size_t)tags[i];
  }
  nsearch::NFlannSearchKd<3> ns(pts, 0, 10);
  ns.setInputCloud();
  rowptr[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    std::vector<size_t> neighs;
    std::vector<double> d2;
    std::vector<int64_t> keep;
    if (query == nullptr || query[i]) {
      if (include_tag) {
        ns.radiusSearchIncludeTag(pts[i], r_per_node[i], neighs, d2, t[i], t);
        for (auto j : neighs)
          if ((int64_t)j != i && t[j] == t[i]) keep.push_back((int64_t)j);
      } else {
        ns.radiusSearchExcludeTag(pts[i], r_per_node[0], neighs, d2, t[i], t);
        for (auto j : neighs)
          if ((int64_t)j != i) keep.push_back((int64_t)j);
      }
    }
    rowptr[i + 1] = rowptr[i] + (int64_t)keep.size();
    if (cols) std::memcpy(cols + rowptr[i], keep.data(), keep.size() * sizeof(int64_t));
  }
}

/* computeMaterialProperties(dim) of a material: out = nu, lambda, mu.
 * law 0 RNPBond, 1 PDState, 2 PMBBond, 3 PDElasticBond. */
void ref_material_properties(int law, double E, double K, double G, double nu, double Gc, double rho,
                             double horizon, int dim, int inf_type, const double *inf_params, int n_inf,
                             double *out) {
  static const char *names[4] = {"RNPBond", "PDState", "PMBBond", "PDElasticBond"};
  auto deck = makeDeck(names[law], E, K, G, nu, Gc, rho, inf_type, inf_params, n_inf);
  inp::MatData md;
  if (law == 0) md = material::RnpMaterial(deck, (size_t)dim, horizon).computeMaterialProperties((size_t)dim);
  if (law == 1) md = material::PdState(deck, (size_t)dim, horizon).computeMaterialProperties((size_t)dim);
  if (law == 2) md = material::PmbMaterial(deck, (size_t)dim, horizon).computeMaterialProperties((size_t)dim);
  if (law == 3) md = material::PdElastic(deck, (size_t)dim, horizon).computeMaterialProperties((size_t)dim);
  out[0] = md.d_nu;
  out[1] = md.d_lambda;
  out[2] = md.d_mu;
}

/* TriElem (dim 2) / TetElem (dim 3) getQuadDatas on one element with
 * vertices nodes [dim + 1][3]: returns the number of points; per point the
 * mapped position pts [3], shape values [dim + 1] and physical shape
 * gradients [dim + 1][dim]. */
int ref_fe_quad_datas(int dim, int order, const double *nodes, double *pts, double *shapes, double *ders) {
  std::vector<util::Point> nds;
  for (int k = 0; k <= dim; k++) nds.emplace_back(nodes[3 * k], nodes[3 * k + 1], nodes[3 * k + 2]);
  std::vector<fe::QuadData> qds;
  if (dim == 2) {
    fe::TriElem e((size_t)order);
    qds = e.getQuadDatas(nds);
  } else {
    fe::TetElem e((size_t)order);
    qds = e.getQuadDatas(nds);
  }
  const int nv = dim + 1;
  for (size_t q = 0; q < qds.size(); q++) {
    pts[3 * q] = qds[q].d_p.d_x;
    pts[3 * q + 1] = qds[q].d_p.d_y;
    pts[3 * q + 2] = qds[q].d_p.d_z;
    for (int i = 0; i < nv; i++) {
      shapes[q * nv + i] = qds[q].d_shapes[i];
      for (int d = 0; d < dim; d++) ders[(q * nv + i) * dim + d] = qds[q].d_derShapes[i][d];
    }
  }
  return (int)qds.size();
}

}  // extern "C"
