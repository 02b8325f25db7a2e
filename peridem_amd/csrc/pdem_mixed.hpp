// pdem_mixed.hpp -- mixed-precision state-based bond kernels (PDEM_PREC_MIXED).
//
// Bond arithmetic in FP32, sums in FP64: the dilation theta and the weighted
// volume stay FP64 (cancellation), bond forces are summed in FP32 over blocks
// of KB = 8 bonds and the blocks in FP64, and the node update stays FP64.
//
// Work unit: a tile = up to TILE_NODES consecutive nodes of one particle.
// Horizons never cross particles, so a tile's neighbors lie in a window of
// node ids [ws, we) inside its particle.  LDS tiles (the window fits) stage
// the window once per workgroup (48 B per node: a float4 of the FP32
// position offset from the particle's first reference node (independent of
// the tile shape) and the volume, the FP64 displacement as a double2 and a
// double, a float2 of the dilation terms) and every bond reads LDS; the remaining
// tiles (particles whose numbering spreads neighbors over more ids than the
// window holds) gather from HBM through the caches: a separate
// instantiation, launched on a second stream beside the LDS tiles.
//
// Neighbor ids are 16-bit offsets from the particle's first node, stored in
// blocks of KB = 8 per lane: one 16-byte load yields a lane's next 8 ids and
// the block loop is unrolled so each id is a compile-time lane of the block.
//
// Threads take nodes in slot order: inside every tile the nodes are ordered
// by their number of 8-bond blocks (most first, node id second), so the 64
// lanes of a wavefront run lists of about the same length (a wavefront runs
// as long as its longest list; surface nodes have far fewer bonds than
// interior ones).  The 16-bit id lists are laid out by slot; node data,
// bond bits and results stay indexed by node.
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
constexpr int LDS_NODES = 1280;   // window of up to 1280 nodes, ~61 KB of LDS per workgroup
// Window classes: a launch whose LDS windows all hold at most LDS_SMALL nodes
// uses the small instantiation (~19 KB of LDS per workgroup, so small particles
// do not hold LDS sized for the largest window and more workgroups fit a CU).
constexpr int LDS_SMALL = 384;
constexpr int MIXED_TPB = 1024;   // default workgroup size (the autotuner may pick 256)
constexpr int KB = 8;             // neighbor ids per 16-byte block
// waves per SIMD the bond kernels are compiled for (at most 64 VGPRs): two
// 1024-thread workgroups per CU, so one stages its window while the other computes
constexpr int MIXED_WAVES_PER_SIMD = 8;

struct DevTile {
  int ts, te;     // nodes [ts, te)
  int ws, we;     // neighbor window [ws, we)
  int pid;
  int lds;        // window staged in LDS
};

struct MixedArgs {
  const DevTile* tiles;   // LDS tiles first, then HBM tiles
  int ntiles_lds, ntiles_hbm;
  int small;              // every LDS window holds at most LDS_SMALL nodes
  // thread slot q runs node slot[q]; its k-th neighbor id is
  // col16[sptr16[q / 64] + (k / KB) * 64 * KB + (q % 64) * KB + k % KB]
  const int* slot;
  const uint16_t* col16;
  const int64_t* sptr16;
  float* fa;              // per node alpha = 15 G / m           (force pass)
  float* fb;              // per node theta ((3K/m) - alpha / 3)  (force pass)
};

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware square root and reciprocal square root (1 ulp, no denormal
// rescaling): bond lengths squared are far from the FP32 denormal range.
// The reciprocal square root yields both |v| = |v|^2 rsq(|v|^2) and 1 / |v|
// from one transcendental (a quarter-rate instruction).
__device__ __forceinline__ float sqrtf_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsqf_hw(float x) { return __builtin_amdgcn_rsqf(x); }
// isGreater(a, b) with a, b >= 0 (util::isGreater, function.h, slack 1e-5 of
// the larger operand) reduces to a (1 - 1e-5) > b: one compare against a
// threshold scaled once per particle.
constexpr float REL_KEEP = 1.f - 1e-5f;

// Influence function J(r) in FP32 (util::ConstInfluenceFn, LinearInfluenceFn,
// GaussianInfluenceFn of influenceFn.h, argument r / delta) with its
// constants folded once per particle: J = a, a + b r (b = p1 / delta) or
// a exp(-b |xi|^2) (b = 1 / (p1 delta^2)).
template <int INF>
struct InfF {
  float a, b;
  __device__ InfF(const DS& s, float inv_delta)
      : a((float)s.inf_p0),
        b(INF == 1 ? (float)s.inf_p1 * inv_delta
                   : INF == 2 ? inv_delta * inv_delta / (float)s.inf_p1 : 0.f) {}
  __device__ __forceinline__ float operator()(float r, float xx) const {
    if (INF == 0) return a;
    if (INF == 1) return fmaf(b, r, a);
    return a * __expf(-b * xx);
  }
};

__device__ __forceinline__ int block_id(const uint4& b, int t) {
  const unsigned w = t < 2 ? b.x : t < 4 ? b.y : t < 6 ? b.z : b.w;
  return (int)((w >> ((t & 1) * 16)) & 0xffffu);
}

// node data access by particle-relative id jp (node gs + jp): LDS window or HBM.
template <int DIM, bool LDS>
struct Win {
  const float4* lxv;
  const double2* lu01;
  const double* lu2;
  const float2* lab;
  const DS* s;
  const MixedArgs* m;
  int gs;   // particle's first node
  int off;  // gs - window start: LDS index of jp
  double org[3];
  __device__ __forceinline__ float4 xv(int jp) const {
    if (LDS) return lxv[jp + off];
    const int64_t j = gs + jp;
    return make_float4((float)(s->xr[0][j] - org[0]), (float)(s->xr[1][j] - org[1]),
                       DIM == 3 ? (float)(s->xr[2][j] - org[2]) : 0.f, (float)s->vol[j]);
  }
  __device__ __forceinline__ void u(int jp, double* out) const {
    if (LDS) {
      const double2 q = lu01[jp + off];
      out[0] = q.x;
      out[1] = q.y;
      if (DIM == 3) out[2] = lu2[jp + off];
    } else {
      for (int d = 0; d < DIM; d++) out[d] = s->u[d][gs + jp];
    }
  }
  __device__ __forceinline__ float2 ab(int jp) const {
    return LDS ? lab[jp + off] : make_float2(m->fa[gs + jp], m->fb[gs + jp]);
  }
};

template <int DIM, bool FORCE>
__device__ __forceinline__ void stage_window(const DS& s, const MixedArgs& m, const DevTile& t,
                                             float4* lxv, double2* lu01, double* lu2, float2* lab,
                                             const double* org) {
  const int nw = t.we - t.ws;
  for (int q = threadIdx.x; q < nw; q += blockDim.x) {
    const int j = t.ws + q;
    lu01[q] = make_double2(s.u[0][j], s.u[1][j]);
    if (DIM == 3) lu2[q] = s.u[2][j];
    lxv[q] = make_float4((float)(s.xr[0][j] - org[0]), (float)(s.xr[1][j] - org[1]),
                         DIM == 3 ? (float)(s.xr[2][j] - org[2]) : 0.f, (float)s.vol[j]);
    if (FORCE) lab[q] = make_float2(m.fa[j], m.fb[j]);
  }
  __syncthreads();
}

// Dilation pass (demModel.cpp:610-679) in mixed precision.
template <int DIM, int INF, bool LDS, int TPB, int NODES = LDS_NODES>
__global__ void __launch_bounds__(TPB, MIXED_WAVES_PER_SIMD) k_mixed_dilation(DS s, MixedArgs m) {
#pragma clang fp contract(fast)
  __shared__ float4 lxv[LDS ? NODES : 1];
  __shared__ double2 lu01[LDS ? NODES : 1];
  __shared__ double lu2[LDS && DIM == 3 ? NODES : 1];
  const DevTile t = m.tiles[LDS ? blockIdx.x : m.ntiles_lds + blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  const int gs = (int)p.glob_start;
  Win<DIM, LDS> w{lxv, lu01, lu2, nullptr, &s, &m, gs, gs - t.ws, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][gs];
  if (LDS) stage_window<DIM, false>(s, m, t, lxv, lu01, lu2, nullptr, w.org);
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float low_rel = check_low / REL_KEEP;             // r > low_rel: isGreater(r, check_low)
  const float thr_rel = (float)(p.mat[2] + 1.0e-10) / REL_KEEP;  // |S| > thr_rel: breaks
  const float inv_h = rcpf(h), inv_delta = rcpf(horizon), up_h = check_up * inv_h;
  const InfF<INF> J(s, inv_delta);
  const double K = p.mat[0], G = p.mat[1];
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) {
    const int i = m.slot[q];
    const int ip = i - gs;
    const uint4* cb = (const uint4*)(m.col16 + m.sptr16[q / WAVE] + (q % WAVE) * KB);
    unsigned* bw = s.bits + s.bptr[i / WAVE] + i % WAVE;
    const int n = s.deg[i];
    const float4 xvi = w.xv(ip);
    const float xi[3] = {xvi.x, xvi.y, xvi.z};
    double ui[3];
    w.u(ip, ui);
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
      auto bond = [&](const int tt) {
        const int k = kb + tt;
        if (k < n) {
          const int jp = block_id(cur, tt);
          const float4 xvj = w.xv(jp);
          const float xj[3] = {xvj.x, xvj.y, xvj.z};
          double uj[3];
          w.u(jp, uj);
          float dx[3], du[3];
          for (int d = 0; d < DIM; d++) {
            dx[d] = xj[d] - xi[d];
            du[d] = (float)(uj[d] - ui[d]);
          }
          float volj = xvj.w;
          float xx = dx[0] * dx[0], xe = dx[0] * du[0], ee = du[0] * du[0];
          for (int d = 1; d < DIM; d++) {
            xx += dx[d] * dx[d];
            xe += dx[d] * du[d];
            ee += du[d] * du[d];
          }
          const float t2 = 2.f * xe + ee;  // |xi + eta|^2 - |xi|^2
          const float inv_r = rsqf_hw(xx);
          const float r = xx * inv_r;
          const float cl = t2 * rcpf(sqrtf_hw(xx + t2) + r);
          const unsigned bit = 1u << (k & 31);
          bool fs = (word & bit) != 0u;
          if (!fs && fabsf(cl) * inv_r > thr_rel) {
            fs = true;
            wnew |= bit;
          }
          if (!fs) {
            if (r > low_rel) volj *= up_h - r * inv_h;
            theta += (double)(r * cl * J(r, xx) * volj);
          }
        }
      };
#pragma unroll
      for (int tt = 0; tt < KB; tt++) bond(tt);
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
// The damage Z_i is the plain maximum of |S| / s0 over the list (the
// reference's running compare with a 1e-5 relative slack keeps a value at
// most 1e-5 below it; FP32 bond arithmetic is the larger difference).
template <int DIM, int INF, bool LDS, int TPB, int NODES = LDS_NODES>
__global__ void __launch_bounds__(TPB, MIXED_WAVES_PER_SIMD) k_mixed_force(DS s, MixedArgs m) {
#pragma clang fp contract(fast)
  __shared__ float4 lxv[LDS ? NODES : 1];
  __shared__ double2 lu01[LDS ? NODES : 1];
  __shared__ double lu2[LDS && DIM == 3 ? NODES : 1];
  __shared__ float2 lab[LDS ? NODES : 1];
  const DevTile t = m.tiles[LDS ? blockIdx.x : m.ntiles_lds + blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  const int gs = (int)p.glob_start;
  Win<DIM, LDS> w{lxv, lu01, lu2, lab, &s, &m, gs, gs - t.ws, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][gs];
  if (LDS) stage_window<DIM, true>(s, m, t, lxv, lu01, lu2, lab, w.org);
  const float horizon = (float)p.horizon, h = (float)p.h;
  const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
  const float low_rel = check_low / REL_KEEP;  // r > low_rel: isGreater(r, check_low)
  const float inv_h = rcpf(h), inv_delta = rcpf(horizon), inv_s0 = (float)(1.0 / p.mat[2]);
  const float up_h = check_up * inv_h;
  const InfF<INF> J(s, inv_delta);
  const float Kn = (float)p.Kn, Rc = (float)p.Rc;
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) {
    const int i = m.slot[q];
    const int ip = i - gs;
    const uint4* cb = (const uint4*)(m.col16 + m.sptr16[q / WAVE] + (q % WAVE) * KB);
    const unsigned* bw = s.bits + s.bptr[i / WAVE] + i % WAVE;
    const int n = s.deg[i];
    const float4 xvi = w.xv(ip);
    const float xi[3] = {xvi.x, xvi.y, xvi.z};
    double ui[3];
    w.u(ip, ui);
    const float2 abi = w.ab(ip);
    const float ai = abi.x, bi = abi.y;
    double force[3] = {0., 0., 0.};
    float Zs = 0.f;  // maximum of |S|; Z = Zs / s0
    unsigned word = 0;
    uint4 blk = n > 0 ? cb[0] : make_uint4(0, 0, 0, 0);
    for (int kb = 0; kb < n; kb += KB) {
      if ((kb & 31) == 0) word = bw[(int64_t)(kb >> 5) * WAVE];
      const uint4 cur = blk;
      if (kb + KB < n) blk = cb[(int64_t)(kb / KB + 1) * WAVE];
      float fblk[3] = {0.f, 0.f, 0.f};
      auto bond = [&](const int tt) {
        const int k = kb + tt;
        if (k < n) {
          const int jp = block_id(cur, tt);
          const float4 xvj = w.xv(jp);
          const float xj[3] = {xvj.x, xvj.y, xvj.z};
          double uj[3];
          w.u(jp, uj);
          float dx[3], du[3], y[3];
          for (int d = 0; d < DIM; d++) {
            dx[d] = xj[d] - xi[d];
            du[d] = (float)(uj[d] - ui[d]);
            y[d] = dx[d] + du[d];
          }
          float volj = xvj.w;
          float xx = dx[0] * dx[0], xe = dx[0] * du[0], ee = du[0] * du[0];
          for (int d = 1; d < DIM; d++) {
            xx += dx[d] * dx[d];
            xe += dx[d] * du[d];
            ee += du[d] * du[d];
          }
          const float t2 = 2.f * xe + ee;  // |y|^2 - |xi|^2
          const float yy = xx + t2;
          const float inv_r = rsqf_hw(xx), inv_ly = rsqf_hw(yy);
          const float r = xx * inv_r, ly = yy * inv_ly;
          const float cl = t2 * rcpf(ly + r);
          const bool fs = (word & (1u << (k & 31))) != 0u;
          float sf;
          if (!fs) {
            if (r > low_rel) volj *= up_h - r * inv_h;
            const float Jr = J(r, xx);
            const float2 abj = w.ab(jp);
            sf = Jr * (r * (bi + abj.y) + cl * (ai + abj.x)) * volj * inv_ly;
          } else {
            sf = Kn * volj * (ly - Rc) * inv_ly;
            if (sf > 0.f) sf = 0.f;
          }
          for (int d = 0; d < DIM; d++) fblk[d] += sf * y[d];
          Zs = fmaxf(Zs, fabsf(cl) * inv_r);
        }
      };
#pragma unroll
      for (int tt = 0; tt < KB; tt++) bond(tt);
      for (int d = 0; d < DIM; d++) force[d] += (double)fblk[d];
    }
    for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
    s.Z[i] = Zs * inv_s0;
  }
}

// the force reset of nodes that take no PD force (a list built at set-up)
// and, on a rebuild step, the binning of every node; grid-stride, so that a
// step without a rebuild launches few workgroups
constexpr int BZ_BLOCKS = 1024;
template <int DIM>
__global__ void k_bin_zero(DS s, int bin) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = gid(); k < s.n_noforce; k += stride) {
    const int64_t i = s.noforce_nodes[k];
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
  }
  if (bin && s.ctrl->rebuild)
    for (int64_t i = gid(); i < s.N; i += stride) bin_node<DIM>(s, i);
}

}  // namespace pdem
