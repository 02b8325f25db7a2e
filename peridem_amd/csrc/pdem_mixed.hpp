// pdem_mixed.hpp -- mixed-precision state-based bond kernels (PDEM_PREC_MIXED).
//
// Bond arithmetic in FP32, sums in FP64: the dilation theta and the weighted
// volume stay FP64 (cancellation), bond forces are summed in FP32 over blocks
// of KB = 8 bonds and the blocks in FP64, and the node update stays FP64.
//
// Work unit: a tile = up to TILE_NODES consecutive nodes of one particle.
// Horizons never cross particles, so a tile's neighbors lie in a window of
// node ids [ws, we) inside its particle.  LDS tiles (the window fits) stage
// the window once per workgroup (structure of arrays, 48 B per node: FP32
// position offset from the particle's first reference node (independent of
// the tile shape), FP64 displacement,
// FP32 volume, FP32 dilation terms) and every bond reads LDS; the remaining
// tiles gather from HBM through the caches (a separate instantiation).
//
// Neighbor ids are 16-bit offsets from the particle's first node, stored in
// blocks of KB = 8 per lane: one 16-byte load yields a lane's next 8 ids and
// the block loop is unrolled so each id is a compile-time lane of the block.
//
// FMA contraction is allowed inside these kernels only (the FP64 parity
// kernels are built without it).
//
// The stretch uses the cancellation-free form
//   |xi + eta| - |xi| = (2 xi.eta + eta.eta) / (|xi + eta| + |xi|)
// so FP32 keeps ~1e-7 relative accuracy at small strains.
#pragma once
#include "pdem_kernels.hpp"

namespace pdem {

constexpr int TILE_NODES = 1024;
constexpr int LDS_NODES = 1280;   // 1280 x 48 B = 60 KB of LDS per workgroup
constexpr int MIXED_TPB = 1024;   // default workgroup size (the autotuner may pick 256)
constexpr int KB = 8;             // neighbor ids per 16-byte block

struct DevTile {
  int ts, te;     // nodes [ts, te)
  int ws, we;     // neighbor window [ws, we)
  int pid;
  int lds;        // window staged in LDS
};

struct MixedArgs {
  const DevTile* tiles;   // LDS tiles first, then HBM tiles
  int ntiles_lds, ntiles_hbm;
  // col16[sptr16[slice] + (k / KB) * 64 * KB + lane * KB + k % KB]
  const uint16_t* col16;
  const int64_t* sptr16;
  float* fa;              // per node alpha = 15 G / m           (force pass)
  float* fb;              // per node theta ((3K/m) - alpha / 3)  (force pass)
};

__device__ __forceinline__ bool isGreaterF(float a, float b) {
  float fa = fabsf(a), fb = fabsf(b);
  return (a - b) > ((fa < fb ? fb : fa) * 1e-5f);
}
__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware square root (1 ulp, no denormal rescaling): bond lengths squared
// are far from the FP32 denormal range
__device__ __forceinline__ float sqrtf_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int INF>
__device__ __forceinline__ float infJf(const DS& s, float r_over_delta) {
  if (INF == 0) return (float)s.inf_p0;
  if (INF == 1) return (float)s.inf_p0 + (float)s.inf_p1 * r_over_delta;
  return (float)s.inf_p0 * __expf(-r_over_delta * r_over_delta / (float)s.inf_p1);
}

__device__ __forceinline__ int block_id(const uint4& b, int t) {
  const unsigned w = t < 2 ? b.x : t < 4 ? b.y : t < 6 ? b.z : b.w;
  return (int)((w >> ((t & 1) * 16)) & 0xffffu);
}

// node data access: LDS window or HBM
template <int DIM, bool LDS>
struct Win {
  const float* lx;
  const double* lu;
  const float* lv;
  const float* la;
  const float* lb;
  const DS* s;
  const MixedArgs* m;
  int ws;
  double org[3];
  __device__ __forceinline__ float x(int d, int j) const {
    return LDS ? lx[d * LDS_NODES + (j - ws)] : (float)(s->xr[d][j] - org[d]);
  }
  __device__ __forceinline__ double u(int d, int j) const {
    return LDS ? lu[d * LDS_NODES + (j - ws)] : s->u[d][j];
  }
  __device__ __forceinline__ float v(int j) const { return LDS ? lv[j - ws] : (float)s->vol[j]; }
  __device__ __forceinline__ float a(int j) const { return LDS ? la[j - ws] : m->fa[j]; }
  __device__ __forceinline__ float b(int j) const { return LDS ? lb[j - ws] : m->fb[j]; }
};

template <int DIM, bool FORCE>
__device__ __forceinline__ void stage_window(const DS& s, const MixedArgs& m, const DevTile& t,
                                             float* lx, double* lu, float* lv, float* la,
                                             float* lb, const double* org) {
  const int nw = t.we - t.ws;
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
template <int DIM, int INF, bool LDS, int TPB>
__global__ void __launch_bounds__(TPB) k_mixed_dilation(DS s, MixedArgs m) {
#pragma clang fp contract(fast)
  __shared__ float lx[LDS ? 3 * LDS_NODES : 1];
  __shared__ double lu[LDS ? 3 * LDS_NODES : 1];
  __shared__ float lv[LDS ? LDS_NODES : 1];
  const DevTile t = m.tiles[LDS ? blockIdx.x : m.ntiles_lds + blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  Win<DIM, LDS> w{lx, lu, lv, nullptr, nullptr, &s, &m, t.ws, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][p.glob_start];
  if (LDS) stage_window<DIM, false>(s, m, t, lx, lu, lv, nullptr, nullptr, w.org);
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float thr = (float)(p.mat[2] + 1.0e-10);
  const float inv_h = rcpf(h), inv_delta = rcpf(horizon);
  const double K = p.mat[0], G = p.mat[1];
  const int gs = (int)p.glob_start;
  for (int i = t.ts + threadIdx.x; i < t.te; i += blockDim.x) {
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const uint4* cb = (const uint4*)(m.col16 + m.sptr16[sl] + lane * KB);
    unsigned* bw = s.bits + s.bptr[sl] + lane;
    const int n = s.deg[i];
    float xi[3];
    double ui[3];
    for (int d = 0; d < DIM; d++) {
      xi[d] = w.x(d, i);
      ui[d] = w.u(d, i);
    }
    double theta = 0.;
    unsigned word = 0, wnew = 0;
    uint4 blk = n > 0 ? cb[0] : make_uint4(0, 0, 0, 0);
    for (int kb = 0; kb < n; kb += KB) {
      if ((kb & 31) == 0) {
        if (kb > 0 && wnew != word) bw[(int64_t)((kb >> 5) - 1) * WAVE] = wnew;
        word = bw[(int64_t)(kb >> 5) * WAVE];
        wnew = word;
      }
      const uint4 cur = blk;
      if (kb + KB < n) blk = cb[(int64_t)(kb / KB + 1) * WAVE];
#pragma unroll
      for (int tt = 0; tt < KB; tt++) {
        const int k = kb + tt;
        if (k < n) {
          const int j = gs + block_id(cur, tt);
          float dx[3], du[3];
          for (int d = 0; d < DIM; d++) {
            dx[d] = w.x(d, j) - xi[d];
            du[d] = (float)(w.u(d, j) - ui[d]);
          }
          float volj = w.v(j);
          float xx = 0.f, xe = 0.f, ee = 0.f, yy = 0.f;
          for (int d = 0; d < DIM; d++) {
            xx += dx[d] * dx[d];
            xe += dx[d] * du[d];
            ee += du[d] * du[d];
            const float y = dx[d] + du[d];
            yy += y * y;
          }
          const float r = sqrtf_hw(xx);
          const float cl = (2.f * xe + ee) * rcpf(sqrtf_hw(yy) + r);
          const float st = cl * rcpf(r);
          const unsigned bit = 1u << (k & 31);
          bool fs = (word & bit) != 0u;
          if (!fs && isGreaterF(fabsf(st), thr)) {
            fs = true;
            wnew |= bit;
          }
          if (!fs) {
            if (isGreaterF(r, check_low)) volj *= (check_up - r) * inv_h;
            theta += (double)(r * cl * infJf<INF>(s, r * inv_delta) * volj);
          }
        }
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
template <int DIM, int INF, bool LDS, int TPB>
__global__ void __launch_bounds__(TPB) k_mixed_force(DS s, MixedArgs m) {
#pragma clang fp contract(fast)
  __shared__ float lx[LDS ? 3 * LDS_NODES : 1];
  __shared__ double lu[LDS ? 3 * LDS_NODES : 1];
  __shared__ float lv[LDS ? LDS_NODES : 1];
  __shared__ float la[LDS ? LDS_NODES : 1];
  __shared__ float lb[LDS ? LDS_NODES : 1];
  const DevTile t = m.tiles[LDS ? blockIdx.x : m.ntiles_lds + blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  Win<DIM, LDS> w{lx, lu, lv, la, lb, &s, &m, t.ws, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][p.glob_start];
  if (LDS) stage_window<DIM, true>(s, m, t, lx, lu, lv, la, lb, w.org);
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float inv_h = rcpf(h), inv_delta = rcpf(horizon), inv_s0 = (float)(1.0 / p.mat[2]);
  const float Kn = (float)p.Kn, Rc = (float)p.Rc;
  const int gs = (int)p.glob_start;
  for (int i = t.ts + threadIdx.x; i < t.te; i += blockDim.x) {
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const uint4* cb = (const uint4*)(m.col16 + m.sptr16[sl] + lane * KB);
    const unsigned* bw = s.bits + s.bptr[sl] + lane;
    const int n = s.deg[i];
    float xi[3];
    double ui[3];
    for (int d = 0; d < DIM; d++) {
      xi[d] = w.x(d, i);
      ui[d] = w.u(d, i);
    }
    const float ai = w.a(i), bi = w.b(i);
    double force[3] = {0., 0., 0.};
    float Zi = 0.f;
    unsigned word = 0;
    uint4 blk = n > 0 ? cb[0] : make_uint4(0, 0, 0, 0);
    for (int kb = 0; kb < n; kb += KB) {
      if ((kb & 31) == 0) word = bw[(int64_t)(kb >> 5) * WAVE];
      const uint4 cur = blk;
      if (kb + KB < n) blk = cb[(int64_t)(kb / KB + 1) * WAVE];
      float fblk[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int tt = 0; tt < KB; tt++) {
        const int k = kb + tt;
        if (k < n) {
          const int j = gs + block_id(cur, tt);
          float dx[3], du[3], y[3];
          for (int d = 0; d < DIM; d++) {
            dx[d] = w.x(d, j) - xi[d];
            du[d] = (float)(w.u(d, j) - ui[d]);
            y[d] = dx[d] + du[d];
          }
          float volj = w.v(j);
          float xx = 0.f, xe = 0.f, ee = 0.f, yy = 0.f;
          for (int d = 0; d < DIM; d++) {
            xx += dx[d] * dx[d];
            xe += dx[d] * du[d];
            ee += du[d] * du[d];
            yy += y[d] * y[d];
          }
          const float r = sqrtf_hw(xx);
          const float ly = sqrtf_hw(yy);
          const float inv_ly = rcpf(ly);
          const float cl = (2.f * xe + ee) * rcpf(ly + r);
          const float S = cl * rcpf(r);
          const bool fs = (word >> (k & 31)) & 1u;
          float sf;
          if (!fs) {
            if (isGreaterF(r, check_low)) volj *= (check_up - r) * inv_h;
            const float J = infJf<INF>(s, r * inv_delta);
            sf = J * (r * (bi + w.b(j)) + cl * (ai + w.a(j))) * volj * inv_ly;
          } else {
            sf = Kn * volj * (ly - Rc) * inv_ly;
            if (sf > 0.f) sf = 0.f;
          }
          for (int d = 0; d < DIM; d++) fblk[d] += sf * y[d];
          const float z = fabsf(S * inv_s0);
          if (isGreaterF(z, Zi)) Zi = z;
        }
      }
      for (int d = 0; d < DIM; d++) force[d] += (double)fblk[d];
    }
    for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
    s.Z[i] = Zi;
  }
}

// binning on a rebuild and the force reset of nodes that take no PD force
template <int DIM>
__global__ void k_bin_zero(DS s, int bin) {
  int64_t i = gid();
  if (i >= s.N) return;
  if (bin && s.ctrl->rebuild) bin_node<DIM>(s, i);
  if (!s.P[s.ptid[i]].compute_force)
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
}

}  // namespace pdem
