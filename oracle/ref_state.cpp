/*
 * ref_state.cpp -- C entry points around more of the reference's own
 * sources, compiled (never copied) from the reference tree by oracle/Makefile
 * into oracle/_ref/libref.so beside ref_harness.cpp.  TEST INFRASTRUCTURE
 * ONLY: the tests use it to pin the oracle's weighted volume and the time
 * functions of its boundary conditions against the reference's code.
 *
 * Reference sources used as they lie:
 *   src/material/materialUtil.cpp   computeStateMxI (an anonymous-namespace
 *                                   function: the file is included into this
 *                                   translation unit so that it can be called)
 *   src/util/transformation.cpp     rotate2D, derRotate2D (rotation BC)
 *   src/util/function.cpp           linearStepFunc, hatFunction (force BC)
 *
 * This unit is compiled with hidden visibility and linked with
 * --gc-sections: materialUtil.cpp's exported wrappers (computeStateMx and
 * friends, which need ModelData, Fracture and the MPI-backed thread count)
 * are never referenced and are dropped, so the library has no undefined
 * symbol.  Only the functions below are exported.
 */
#include "material/materialUtil.cpp"
#include "util/function.h"
#include "util/transformation.h"

#include <cstdint>
#include <vector>

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

namespace {
inp::MaterialDeck state_deck(int inf_type, const double *inf_params, int n_inf) {
  inp::MaterialDeck deck;
  deck.d_materialType = "PDState";
  deck.d_matData.d_K = 2.16e7;
  deck.d_matData.d_G = 1.296e7;
  deck.d_matData.d_Gc = 50.;
  deck.d_density = 1200.;
  deck.d_computeParamsFromElastic = true;
  deck.d_influenceFnType = (size_t)inf_type;
  for (int i = 0; i < n_inf; i++) deck.d_influenceFnParams.push_back(inf_params[i]);
  // the influence function is a global that the first material constructs;
  // reset it so every call sees its own type
  influence_fn = nullptr;
  return deck;
}
}  // namespace

/* Weighted volume m_i of every node by the reference's computeStateMxI
 * (materialUtil.cpp:21-73) over the given neighbor lists (CSR, in the order
 * given), reference positions x [n][3], volumes, one mesh size and horizon.
 * The material (PDState) only supplies the horizon and the influence
 * function J.  Returns 0, or 1 when some m_i would be below 1e-18 (the
 * reference exits there; no call is made then). */
REF_EXPORT int ref_state_mx(int dim, const double *x, int64_t n, const double *vol, const int64_t *rowptr,
                            const int64_t *cols, double mesh_size, double horizon, int inf_type,
                            const double *inf_params, int n_inf, double *out) {
  auto deck = state_deck(inf_type, inf_params, n_inf);
  material::PdState m(deck, (size_t)dim, horizon);
  std::vector<util::Point> nodes((size_t)n);
  for (int64_t i = 0; i < n; i++) nodes[i] = util::Point(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
  std::vector<double> v(vol, vol + n);
  std::vector<std::vector<size_t>> nb((size_t)n);
  std::vector<std::vector<float>> sq((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) nb[i].push_back((size_t)cols[k]);
    sq[i].assign(nb[i].size(), 0.f);
  }
  // screen out the inputs on which the reference would exit
  for (int64_t i = 0; i < n; i++) {
    double m_est = 0.;
    for (size_t j : nb[i]) {
      const double r = (nodes[j] - nodes[i]).length();
      m_est += r * r * v[j];
    }
    if (!(m_est > 1e-17)) return 1;
  }
  for (int64_t i = 0; i < n; i++) out[i] = computeStateMxI((size_t)i, nodes, v, nb, sq, mesh_size, &m);
  return 0;
}

/* util::rotate2D and util::derRotate2D (transformation.cpp:53-65) of (x, y, 0) */
REF_EXPORT void ref_rotate2d(double x, double y, double theta, double *out2, double *der2) {
  const util::Point p(x, y, 0.);
  const auto r = util::rotate2D(p, theta);
  const auto d = util::derRotate2D(p, theta);
  out2[0] = r.d_x;
  out2[1] = r.d_y;
  der2[0] = d.d_x;
  der2[1] = d.d_y;
}

/* util::linearStepFunc (function.cpp:62-86) */
REF_EXPORT double ref_linear_step(double x, double x1, double x2) { return util::linearStepFunc(x, x1, x2); }

/* util::hatFunction (function.cpp:25-44) */
REF_EXPORT double ref_hat(double x, double x_min, double x_max) { return util::hatFunction(x, x_min, x_max); }
