// pdem_mixed.hpp -- mixed-precision state-based bond kernels (PDEM_PREC_MIXED).
//
// Bond arithmetic in FP32, sums in FP64: the dilation theta and the weighted
// volume stay FP64 (cancellation), the per-node force sum is accumulated in
// FP64, and the node update stays FP64.
//
// Work unit: a tile = up to TILE_NODES consecutive nodes of one particle.
// Horizons never cross particles, so a tile's neighbors lie in a window of
// node ids [ws, we) inside its particle.  When the window fits, the workgroup
// stages it in LDS once (structure of arrays, 48 B per node: FP32 position
// offset from the window's first reference node, FP64 displacement, FP32
// volume, FP32 dilation terms) and every bond reads LDS; otherwise the tile
// gathers from HBM through the caches.
//
// The stretch uses the cancellation-free form
//   |xi + eta| - |xi| = (2 xi.eta + eta.eta) / (|xi + eta| + |xi|)
// so FP32 keeps ~1e-7 relative accuracy at small strains.
#pragma once
#include "pdem_kernels.hpp"

namespace pdem {

constexpr int TILE_NODES = 1024;
constexpr int LDS_NODES = 1280;   // 1280 x 48 B = 60 KB of LDS per workgroup
constexpr int MIXED_TPB = 256;

struct DevTile {
  int ts, te;     // nodes [ts, te)
  int ws, we;     // neighbor window [ws, we)
  int pid;
  int lds;        // window staged in LDS
};

struct MixedArgs {
  const DevTile* tiles;
  int ntiles;
  const uint16_t* col16;  // neighbor ids relative to the particle's first node, sliced ELL
  float* fa;              // per node alpha = 15 G / m           (force pass)
  float* fb;              // per node theta ((3K/m) - alpha / 3)  (force pass)
};

__device__ __forceinline__ bool isGreaterF(float a, float b) {
  float fa = fabsf(a), fb = fabsf(b);
  return (a - b) > ((fa < fb ? fb : fa) * 1e-5f);
}

template <int INF>
__device__ __forceinline__ float infJf(const DS& s, float r_over_delta) {
  if (INF == 0) return (float)s.inf_p0;
  if (INF == 1) return (float)s.inf_p0 + (float)s.inf_p1 * r_over_delta;
  return (float)s.inf_p0 * __expf(-r_over_delta * r_over_delta / (float)s.inf_p1);
}

// Stage the tile's window; returns through LDS pointers.
template <int DIM, bool FORCE>
__device__ __forceinline__ void stage_window(const DS& s, const MixedArgs& m, const DevTile& t,
                                             float* lx, double* lu, float* lv, float* la,
                                             float* lb, double* org) {
  const int nw = t.we - t.ws;
  for (int d = 0; d < DIM; d++) org[d] = s.xr[d][t.ws];
  for (int q = threadIdx.x; q < nw; q += blockDim.x) {
    const int j = t.ws + q;
    for (int d = 0; d < DIM; d++) {
      lx[d * LDS_NODES + q] = (float)(s.xr[d][j] - org[d]);
      lu[d * LDS_NODES + q] = s.u[d][j];
    }
    lv[q] = (float)s.vol[j];
    if (FORCE) {
      la[q] = m.fa[j];
      lb[q] = m.fb[j];
    }
  }
  __syncthreads();
}

// Dilation pass (demModel.cpp:610-679) in mixed precision.
template <int DIM, int INF>
__global__ void __launch_bounds__(MIXED_TPB) k_mixed_dilation(DS s, MixedArgs m) {
  __shared__ float lx[3 * LDS_NODES];
  __shared__ double lu[3 * LDS_NODES];
  __shared__ float lv[LDS_NODES];
  const DevTile t = m.tiles[blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  double org[3] = {0., 0., 0.};
  if (t.lds) stage_window<DIM, false>(s, m, t, lx, lu, lv, nullptr, nullptr, org);
  else
    for (int d = 0; d < DIM; d++) org[d] = s.xr[d][t.ws];
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float thr = (float)(p.mat[2] + 1.0e-10);
  const float inv_h = 1.0f / h, inv_delta = 1.0f / horizon;
  const double K = p.mat[0], G = p.mat[1];
  const int gs = (int)p.glob_start;
  for (int i = t.ts + threadIdx.x; i < t.te; i += blockDim.x) {
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const uint16_t* c = m.col16 + s.sptr[sl] + lane;
    unsigned* bw = s.bits + s.bptr[sl] + lane;
    const int n = s.deg[i];
    const int il = i - t.ws;
    float xi[3];
    double ui[3];
    for (int d = 0; d < DIM; d++) {
      xi[d] = t.lds ? lx[d * LDS_NODES + il] : (float)(s.xr[d][i] - org[d]);
      ui[d] = t.lds ? lu[d * LDS_NODES + il] : s.u[d][i];
    }
    double theta = 0.;
    unsigned word = 0, wnew = 0;
    for (int k = 0; k < n; k++) {
      if ((k & 31) == 0) {
        if (k > 0 && wnew != word) bw[(int64_t)((k >> 5) - 1) * WAVE] = wnew;
        word = bw[(int64_t)(k >> 5) * WAVE];
        wnew = word;
      }
      const int j = gs + (int)c[(int64_t)k * WAVE];
      const int jl = j - t.ws;
      float dx[3], du[3];
      float volj;
      for (int d = 0; d < DIM; d++) {
        const float xj = t.lds ? lx[d * LDS_NODES + jl] : (float)(s.xr[d][j] - org[d]);
        const double uj = t.lds ? lu[d * LDS_NODES + jl] : s.u[d][j];
        dx[d] = xj - xi[d];
        du[d] = (float)(uj - ui[d]);
      }
      volj = t.lds ? lv[jl] : (float)s.vol[j];
      float xx = 0.f, xe = 0.f, ee = 0.f, yy = 0.f;
      for (int d = 0; d < DIM; d++) {
        xx += dx[d] * dx[d];
        xe += dx[d] * du[d];
        ee += du[d] * du[d];
        const float y = dx[d] + du[d];
        yy += y * y;
      }
      const float r = sqrtf(xx);
      const float cl = (2.f * xe + ee) / (sqrtf(yy) + r);
      const float st = cl / r;
      bool fs = (word >> (k & 31)) & 1u;
      if (!fs && isGreaterF(fabsf(st), thr)) {
        fs = true;
        wnew |= 1u << (k & 31);
      }
      if (!fs) {
        if (isGreaterF(r, check_low)) volj *= (check_up - r) * inv_h;
        theta += (double)(r * cl * infJf<INF>(s, r * inv_delta) * volj);
      }
    }
    if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
    const double mi = s.mx[i];
    const double th = 3. * theta / mi;
    s.theta[i] = th;
    const double alpha = 15. * G / mi;
    m.fa[i] = (float)alpha;
    m.fb[i] = (float)(th * ((3. * K / mi) - alpha / 3.));
  }
}

// Force pass (demModel.cpp:682-788, PdState::getBondEF material.h:1415-1429)
// in mixed precision: ef_i + ef_j = J (r (b_i + b_j) + cl (a_i + a_j)).
template <int DIM, int INF>
__global__ void __launch_bounds__(MIXED_TPB) k_mixed_force(DS s, MixedArgs m) {
  __shared__ float lx[3 * LDS_NODES];
  __shared__ double lu[3 * LDS_NODES];
  __shared__ float lv[LDS_NODES];
  __shared__ float la[LDS_NODES];
  __shared__ float lb[LDS_NODES];
  const DevTile t = m.tiles[blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  double org[3] = {0., 0., 0.};
  if (t.lds) stage_window<DIM, true>(s, m, t, lx, lu, lv, la, lb, org);
  else
    for (int d = 0; d < DIM; d++) org[d] = s.xr[d][t.ws];
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float inv_h = 1.0f / h, inv_delta = 1.0f / horizon, inv_s0 = (float)(1.0 / p.mat[2]);
  const float Kn = (float)p.Kn, Rc = (float)p.Rc;
  const int gs = (int)p.glob_start;
  for (int i = t.ts + threadIdx.x; i < t.te; i += blockDim.x) {
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const uint16_t* c = m.col16 + s.sptr[sl] + lane;
    const unsigned* bw = s.bits + s.bptr[sl] + lane;
    const int n = s.deg[i];
    const int il = i - t.ws;
    float xi[3];
    double ui[3];
    for (int d = 0; d < DIM; d++) {
      xi[d] = t.lds ? lx[d * LDS_NODES + il] : (float)(s.xr[d][i] - org[d]);
      ui[d] = t.lds ? lu[d * LDS_NODES + il] : s.u[d][i];
    }
    const float ai = t.lds ? la[il] : m.fa[i];
    const float bi = t.lds ? lb[il] : m.fb[i];
    double force[3] = {0., 0., 0.};
    float Zi = 0.f;
    unsigned word = 0;
    for (int k = 0; k < n; k++) {
      if ((k & 31) == 0) word = bw[(int64_t)(k >> 5) * WAVE];
      const int j = gs + (int)c[(int64_t)k * WAVE];
      const int jl = j - t.ws;
      float dx[3], du[3], y[3];
      for (int d = 0; d < DIM; d++) {
        const float xj = t.lds ? lx[d * LDS_NODES + jl] : (float)(s.xr[d][j] - org[d]);
        const double uj = t.lds ? lu[d * LDS_NODES + jl] : s.u[d][j];
        dx[d] = xj - xi[d];
        du[d] = (float)(uj - ui[d]);
        y[d] = dx[d] + du[d];
      }
      float volj = t.lds ? lv[jl] : (float)s.vol[j];
      float xx = 0.f, xe = 0.f, ee = 0.f, yy = 0.f;
      for (int d = 0; d < DIM; d++) {
        xx += dx[d] * dx[d];
        xe += dx[d] * du[d];
        ee += du[d] * du[d];
        yy += y[d] * y[d];
      }
      const float r = sqrtf(xx);
      const float ly = sqrtf(yy);
      const float cl = (2.f * xe + ee) / (ly + r);
      const float S = cl / r;
      const bool fs = (word >> (k & 31)) & 1u;
      float sf;
      if (!fs) {
        if (isGreaterF(r, check_low)) volj *= (check_up - r) * inv_h;
        const float J = infJf<INF>(s, r * inv_delta);
        const float aj = t.lds ? la[jl] : m.fa[j];
        const float bj = t.lds ? lb[jl] : m.fb[j];
        sf = J * (r * (bi + bj) + cl * (ai + aj)) * volj / ly;
      } else {
        sf = Kn * volj * (ly - Rc) / ly;
        if (sf > 0.f) sf = 0.f;
      }
      for (int d = 0; d < DIM; d++) force[d] += (double)(sf * y[d]);
      const float z = fabsf(S * inv_s0);
      if (isGreaterF(z, Zi)) Zi = z;
    }
    for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
    s.Z[i] = Zi;
  }
}

// binning on a rebuild and the force reset of nodes that take no PD force
template <int DIM>
__global__ void k_bin_zero(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  if (s.ctrl->rebuild) bin_node<DIM>(s, i);
  if (!s.P[s.ptid[i]].compute_force)
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
}

}  // namespace pdem
