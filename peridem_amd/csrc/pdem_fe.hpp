// Strain / stress post-processing at the quadrature points of the particle
// meshes: DEMModel::output's Strain_Stress block (demModel.cpp:2026-2072)
// with fe::getCurrentQuadPoints and fe::getStrainStress (meshUtil.cpp:176-420),
// linear triangles (TriElem::getQuadDatas, triElem.cpp:58-90) and linear
// tetrahedra (TetElem::getQuadDatas, tetElem.cpp:110-150, util::det / inv,
// matrix.cpp:74-150).  One lane per element; FP64 in the reference's operation
// order, strain and stress kept in FP32 like the reference's SymMatrix3
// (util/matrix.h:270-300: every `+=` rounds to float).
//
// The reference's normal strain sums u_i + dN_i/dx (meshUtil.cpp:364-372, a
// `+` where the symmetric gradient has a product); it is kept as is, so the
// output equals the reference's.  The shear terms are the symmetric gradient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdem {

constexpr int FE_MAX_Q = 7;  // most points of a supported rule (triangle, order 5)

struct FeArgs {
  int64_t n_elem = 0;
  int nq = 0;                       // points per element
  const int64_t* nodes = nullptr;   // [n_elem][DIM + 1] global node ids
  const int* part = nullptr;        // [n_elem] owning particle
  const double* elastic = nullptr;  // [n_particles][4] nu, lambda, mu, plane strain flag
  const double* qpt = nullptr;      // [nq][3] reference-element points
  const double* qshape = nullptr;   // [nq][DIM + 1] shape values at the points
  double* xq = nullptr;             // [n_elem * nq][3]
  float* strain = nullptr;          // [n_elem * nq][6]  xx yy zz yz xz xy
  float* stress = nullptr;          // [n_elem * nq][6]
};

// float += double, as `SymMatrix3(i, j) += expr` rounds
__device__ __forceinline__ void facc(float& a, double t) { a = (float)((double)a + t); }

template <int DIM>
__global__ void __launch_bounds__(256) k_strain_stress(DS s, FeArgs a) {
  const int64_t e = gid();
  if (e >= a.n_elem) return;
  constexpr int NV = DIM + 1;
  int64_t g[NV];
  double X[NV][3];
  double U[NV][3];
  for (int k = 0; k < NV; k++) {
    g[k] = a.nodes[e * NV + k];
    for (int d = 0; d < 3; d++) {
      X[k][d] = d < DIM ? s.xr[d][g[k]] : 0.;
      U[k][d] = d < DIM ? s.u[d][g[k]] : 0.;
    }
  }
  // shape-function gradients (constant over a linear element)
  double dN[NV][3] = {};
  if (DIM == 2) {
    const double J00 = X[1][0] - X[0][0], J01 = X[1][1] - X[0][1];
    const double J10 = X[2][0] - X[0][0], J11 = X[2][1] - X[0][1];
    const double det = J00 * J11 - J01 * J10;
    const double gr[3][2] = {{-1., -1.}, {1., 0.}, {0., 1.}};
    for (int i = 0; i < 3; i++) {
      dN[i][0] = (gr[i][0] * J11 - gr[i][1] * J01) / det;
      dN[i][1] = (-gr[i][0] * J10 + gr[i][1] * J00) / det;
    }
  } else {
    double m[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) m[r][c] = X[r + 1][c] - X[0][c];
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
    const double di = 1. / det;
    double n[3][3];
    n[0][0] = di * (m[1][1] * m[2][2] - m[2][1] * m[1][2]);
    n[0][1] = -di * (m[0][1] * m[2][2] - m[2][1] * m[0][2]);
    n[0][2] = di * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    n[1][0] = -di * (m[1][0] * m[2][2] - m[2][0] * m[1][2]);
    n[1][1] = di * (m[0][0] * m[2][2] - m[2][0] * m[0][2]);
    n[1][2] = -di * (m[0][0] * m[1][2] - m[1][0] * m[0][2]);
    n[2][0] = di * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
    n[2][1] = -di * (m[0][0] * m[2][1] - m[2][0] * m[0][1]);
    n[2][2] = di * (m[0][0] * m[1][1] - m[1][0] * m[0][1]);
    const double gr[4][3] = {{-1., -1., -1.}, {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 3; k++) {
        double r = 0.;
        for (int j = 0; j < 3; j++) r += n[k][j] * gr[i][j];
        dN[i][k] = r;
      }
  }
  const double* em = a.elastic + 4 * a.part[e];
  const double nu = em[0], lambda = em[1], mu = em[2];
  const bool plane_strain = em[3] != 0.;
  for (int q = 0; q < a.nq; q++) {
    const double* N = a.qshape + q * NV;
    // current position of the point: mapped reference point plus sum u_i N_i
    double p[3];
    if (DIM == 2) {
      p[0] = N[0] * X[0][0] + N[1] * X[1][0] + N[2] * X[2][0];
      p[1] = N[0] * X[0][1] + N[1] * X[1][1] + N[2] * X[2][1];
      p[2] = a.qpt[3 * q + 2];
    } else {
      for (int d = 0; d < 3; d++) {
        double t = 0.;
        for (int i = 0; i < NV; i++) t += N[i] * X[i][d];
        p[d] = t;
      }
    }
    for (int i = 0; i < NV; i++)
      for (int d = 0; d < 3; d++) p[d] += U[i][d] * N[i];
    const int64_t o = e * a.nq + q;
    for (int d = 0; d < 3; d++) a.xq[3 * o + d] = p[d];

    float sn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < NV; i++) {
      const double* ui = U[i];
      facc(sn[0], ui[0] + dN[i][0]);
      facc(sn[1], ui[1] + dN[i][1]);
      facc(sn[5], 0.5 * ui[0] * dN[i][1] + 0.5 * ui[1] * dN[i][0]);
      if (DIM > 2) {
        facc(sn[2], ui[2] + dN[i][2]);
        facc(sn[3], 0.5 * ui[1] * dN[i][2] + 0.5 * ui[2] * dN[i][1]);
        facc(sn[4], 0.5 * ui[0] * dN[i][2] + 0.5 * ui[2] * dN[i][0]);
      }
    }
    if (DIM == 2 && plane_strain) sn[2] = (float)(-nu * (double)(sn[0] + sn[1]) / (1. - nu));
    const float tr = sn[0] + sn[1] + sn[2];
    float ss[6];
    ss[0] = (float)(lambda * (double)tr + 2. * mu * (double)sn[0]);
    ss[5] = (float)(2. * mu * (double)sn[5]);
    ss[4] = (float)(2. * mu * (double)sn[4]);
    ss[1] = (float)(lambda * (double)tr + 2. * mu * (double)sn[1]);
    ss[3] = (float)(2. * mu * (double)sn[3]);
    ss[2] = (float)(lambda * (double)tr + 2. * mu * (double)sn[2]);
    if (DIM == 2 && !plane_strain) ss[2] = (float)(nu * (double)(ss[0] + ss[1]));
    for (int k = 0; k < 6; k++) {
      a.strain[6 * o + k] = sn[k];
      a.stress[6 * o + k] = ss[k];
    }
  }
}

}  // namespace pdem
