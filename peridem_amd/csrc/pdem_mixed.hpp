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
  int lds;        // window staged in LDS; 2: also closed (the tile is its whole particle)
};

struct MixedArgs {
  const DevTile* tiles;   // closed LDS tiles, then open LDS tiles, then HBM tiles
  int ntiles_closed, ntiles_lds, ntiles_hbm;
  int small;              // every LDS window holds at most LDS_SMALL nodes
  // thread slot q runs node slot[q]; its k-th neighbor id is
  // col16[sptr16[q / 64] + (k / KB) * 64 * KB + (q % 64) * KB + k % KB]
  const int* slot;
  const uint16_t* col16;
  const int64_t* sptr16;
  float* fa;              // per node alpha = 15 G / m           (force pass)
  float* fb;              // per node theta ((3K/m) - alpha / 3)  (force pass)
  // particles with open LDS tiles: max |u - u(first node)| over the particle
  // (rounded up to float), for the fixed-point step of their windows
  const int* open_pids;
  int n_open_pids;
  float* pumax;           // [np]
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

// Node data by particle-relative id jp (node gs + jp): the LDS window or HBM.
//
// LDS window records, per window node: a float4 of the FP32 position offset
// from the particle's first reference node and the volume; an int4 of the
// displacement relative to the window's first node in fixed point (three
// int32 in units of a power-of-two step `us` chosen per window so that every
// value stays below 2^29 in magnitude) and, in the force pass, the node's
// b term; the node's alpha term as a float.  A bond's relative displacement
// is the exact integer difference of two records, rounded once to FP32 and
// scaled by us (exact): as accurate as rounding the FP64 difference to FP32,
// plus a quantisation of 2^-29 of the window's largest relative
// displacement.  32 bytes per bond in the dilation pass, 36 in the force pass.
// HBM tiles read the FP64 displacements (us = 1).
template <int DIM, bool LDS>
struct Win {
  const float4* lxv;
  const int4* lq;
  const float* la;
  const DS* s;
  const MixedArgs* m;
  int gs;    // particle's first node
  int off;   // gs - window start: LDS index of jp
  float us;  // fixed-point step of the displacement records (1 for HBM tiles)
  double org[3];
  __device__ __forceinline__ float4 xv(int jp) const {
    if (LDS) return lxv[jp + off];
    const int64_t j = gs + jp;
    return make_float4((float)(s->xr[0][j] - org[0]), (float)(s->xr[1][j] - org[1]),
                       DIM == 3 ? (float)(s->xr[2][j] - org[2]) : 0.f, (float)s->vol[j]);
  }
  // the node's own displacement record (fixed point) or displacement (FP64)
  struct Own {
    int q[3];
    double u[3];
    float b;  // the node's b term (force pass)
  };
  __device__ __forceinline__ Own own(int jp) const {
    Own o{};
    if (LDS) {
      const int4 r = lq[jp + off];
      o.q[0] = r.x;
      o.q[1] = r.y;
      o.q[2] = r.z;
      o.b = __int_as_float(r.w);
    } else {
      for (int d = 0; d < DIM; d++) o.u[d] = s->u[d][gs + jp];
    }
    return o;
  }
  // relative displacement of neighbor jp in units of us (FP32); with B, bj
  // gets the neighbor's b term from the same LDS record (the dilation pass
  // reads only the displacement words: in the fused kernel it runs while b
  // terms are being written)
  template <bool B>
  __device__ __forceinline__ void du(int jp, const Own& o, float* out, float& bj) const {
    if (LDS) {
      if (B) {
        const int4 r = lq[jp + off];
        out[0] = (float)(r.x - o.q[0]);
        out[1] = (float)(r.y - o.q[1]);
        if (DIM == 3) out[2] = (float)(r.z - o.q[2]);
        bj = __int_as_float(r.w);
      } else {
        const int* r = (const int*)(lq + jp + off);
        out[0] = (float)(r[0] - o.q[0]);
        out[1] = (float)(r[1] - o.q[1]);
        if (DIM == 3) out[2] = (float)(r[2] - o.q[2]);
        bj = 0.f;
      }
    } else {
      for (int d = 0; d < DIM; d++) out[d] = (float)(s->u[d][gs + jp] - o.u[d]);
      bj = 0.f;
    }
  }
  __device__ __forceinline__ float2 ab(int jp, float bj) const {
    return LDS ? make_float2(la[jp + off], bj) : make_float2(m->fa[gs + jp], m->fb[gs + jp]);
  }
};

// max |u - u(particle's first node)| over nodes [n0, n1) of the workgroup's
// threads, as a float rounded up (the fixed-point step only has to bound it);
// the maximum does not depend on the order, so every tile of a particle gets
// the same value
template <int DIM>
__device__ __forceinline__ float particle_umax(const DS& s, int gs, int n0, int n1, float* red) {
  double u0[3] = {0., 0., 0.};
  for (int d = 0; d < DIM; d++) u0[d] = s.u[d][gs];
  double mx = 0.;
  for (int j = n0 + (int)threadIdx.x; j < n1; j += blockDim.x)
    for (int d = 0; d < DIM; d++) mx = fmax(mx, fabs(s.u[d][j] - u0[d]));
  float mf = (float)(mx * (1.0 + 0x1p-20));
  for (int o = WAVE / 2; o > 0; o >>= 1) mf = fmaxf(mf, __shfl_xor(mf, o));
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = mf;
  __syncthreads();
  mf = 0.f;
  for (int w = 0; w < (int)(blockDim.x / WAVE); w++) mf = fmaxf(mf, red[w]);
  __syncthreads();
  return mf;
}

// the maximum for the particles with open LDS tiles: one workgroup each
template <int DIM>
__global__ void __launch_bounds__(256) k_mixed_umax(DS s, MixedArgs m) {
  __shared__ float red[256 / WAVE];
  const int pid = m.open_pids[blockIdx.x];
  const DevParticle& p = s.P[pid];
  const float mf = particle_umax<DIM>(s, (int)p.glob_start, (int)p.glob_start, (int)p.glob_end, red);
  if (threadIdx.x == 0) m.pumax[pid] = mf;
}

// Stage a tile's window in LDS: positions and volumes, fixed-point
// displacements relative to the particle's first node (the step us: the
// smallest power of two with max |u - u_first| / us < 2^29 over the whole
// particle, so that the records, and the results, do not depend on the tile
// shape) and, for the force pass of an open tile, the dilation terms of
// every window node.  A closed tile (the whole particle) takes the maximum
// over its window; an open tile reads it from k_mixed_umax.  Returns us.
template <int DIM, bool FORCE>
__device__ __forceinline__ float stage_window(const DS& s, const MixedArgs& m, const DevTile& t, float4* lxv,
                                              int4* lq, float* la, const double* org, float* red) {
  const int nw = t.we - t.ws;
  const int gs = (int)s.P[t.pid].glob_start;
  double u0[3] = {0., 0., 0.};
  for (int d = 0; d < DIM; d++) u0[d] = s.u[d][gs];
  for (int q = threadIdx.x; q < nw; q += blockDim.x) {
    const int j = t.ws + q;
    lxv[q] = make_float4((float)(s.xr[0][j] - org[0]), (float)(s.xr[1][j] - org[1]),
                         DIM == 3 ? (float)(s.xr[2][j] - org[2]) : 0.f, (float)s.vol[j]);
  }
  const float mf = t.lds == 2 ? particle_umax<DIM>(s, gs, t.ws, t.we, red) : m.pumax[t.pid];
  // us = 2^e with mf / us < 2^29 (e >= -126: a window at rest keeps all zeros)
  int e = -126;
  if (mf > 0.f) {
    int ex;
    frexpf(mf, &ex);  // mf < 2^ex
    e = ex - 29 > -126 ? ex - 29 : -126;
  }
  const float us = ldexpf(1.f, e);
  const double inv_us = ldexp(1., -e);
  for (int q = threadIdx.x; q < nw; q += blockDim.x) {
    const int j = t.ws + q;
    int qq[3] = {0, 0, 0};
    for (int d = 0; d < DIM; d++) qq[d] = (int)rint((s.u[d][j] - u0[d]) * inv_us);
    lq[q] = make_int4(qq[0], qq[1], qq[2], FORCE ? __float_as_int(m.fb[j]) : 0);
    if (FORCE) la[q] = m.fa[j];
  }
  __syncthreads();
  return us;
}

// Per-particle constants of the mixed bond loops (FP32, folded once per
// workgroup).
template <int INF>
struct StateConsts {
  float low_rel;   // r > low_rel: isGreater(r, horizon - h/2), the volume correction applies
  float thr_rel;   // |S| > thr_rel: the bond breaks (isGreater(|S|, s0 + 1e-10))
  float inv_h, up_h, inv_s0, Kn, Rc;
  double K, G;
  InfF<INF> J;
  __device__ StateConsts(const DS& s, const DevParticle& p)
      : J(s, rcpf((float)p.horizon)) {
    const float horizon = (float)p.horizon, h = (float)p.h;
    const float check_up = horizon + 0.5f * h, check_low = horizon - 0.5f * h;
    low_rel = check_low / REL_KEEP;
    thr_rel = (float)(p.mat[2] + 1.0e-10) / REL_KEEP;
    inv_h = rcpf(h);
    up_h = check_up * inv_h;
    inv_s0 = (float)(1.0 / p.mat[2]);
    Kn = (float)p.Kn;
    Rc = (float)p.Rc;
    K = p.mat[0];
    G = p.mat[1];
  }
};

// Dilation of the node in thread slot q (demModel.cpp:610-679) in mixed
// precision: breaks bonds, writes theta and the damage Z and returns the
// node's force-pass terms (alpha, theta (3K/m - alpha/3)) as FP32.  The
// damage Z_i (demModel.cpp:770-773) is the plain maximum of |S| / s0 over
// the list, broken bonds included: |S| is the break test's value, taken
// here rather than in the force pass (the reference's running compare with
// a 1e-5 relative slack keeps a value at most 1e-5 below it; FP32 bond
// arithmetic is the larger difference).
template <int DIM, int INF, bool LDS>
__device__ __forceinline__ float2 dilation_node(const DS& s, const MixedArgs& m, const Win<DIM, LDS>& w,
                                                const StateConsts<INF>& c, int q, int& node) {
#pragma clang fp contract(fast)
  const int i = m.slot[q];
  node = i;
  const int ip = i - w.gs;
  const uint4* cb = (const uint4*)(m.col16 + m.sptr16[q / WAVE] + (q % WAVE) * KB);
  unsigned* bw = s.bits + s.bptr[i / WAVE] + i % WAVE;
  const int n = s.deg[i];
  const float4 xvi = w.xv(ip);
  const float xi[3] = {xvi.x, xvi.y, xvi.z};
  const auto own = w.own(ip);
  const float us = w.us;
  // |xi + eta|^2 - |xi|^2 = us^2 ee + 2 us xe in one multiply and one FMA
  // (us is a power of two: the same value as us (us ee + 2 xe) while us^2
  // is a normal float, i.e. for displacements above ~1e-14 of the unit)
  const float us2 = us * us, us2x = 2.f * us;
  double theta = 0.;
  float Zs = 0.f;  // maximum of |S| over all bonds, broken ones included; Z = Zs / s0
  unsigned word = 0, wnew = 0;
  uint4 blk = n > 0 ? cb[0] : make_uint4(0, 0, 0, 0);
  for (int kb = 0; kb < n; kb += KB) {
    if ((kb & 31) == 0) {
      if (kb > 0 && wnew != word) bw[(int64_t)((kb >> 5) - 1) * WAVE] = wnew;
      word = bw[(int64_t)(kb >> 5) * WAVE];
      wnew = word;
    }
    const unsigned wblk = word >> (kb & 31);  // this block's 8 bond bits from bit 0
    unsigned brk = 0;                         // bonds of the block that break now
    const uint4 cur = blk;
    if (kb + KB < n) blk = cb[(int64_t)(kb / KB + 1) * WAVE];
    auto bond = [&](const int tt) {
      if (kb + tt < n) {
        const int jp = block_id(cur, tt);
        const float4 xvj = w.xv(jp);
        const float xj[3] = {xvj.x, xvj.y, xvj.z};
        float dx[3], du[3], bj;
        w.template du<false>(jp, own, du, bj);  // in units of us
        for (int d = 0; d < DIM; d++) dx[d] = xj[d] - xi[d];
        float volj = xvj.w;
        float xx = dx[0] * dx[0], xe = dx[0] * du[0], ee = du[0] * du[0];
        for (int d = 1; d < DIM; d++) {
          xx += dx[d] * dx[d];
          xe += dx[d] * du[d];
          ee += du[d] * du[d];
        }
        const float t2 = fmaf(us2, ee, us2x * xe);  // |xi + eta|^2 - |xi|^2
        const float inv_r = rsqf_hw(xx);
        const float r = xx * inv_r;
        const float cl = t2 * rcpf(sqrtf_hw(xx + t2) + r);
        bool fs = ((wblk >> tt) & 1u) != 0u;
        const float sa = fabsf(cl) * inv_r;  // |S|
        // maximum of non-negative floats: their bit patterns order alike
        Zs = __uint_as_float(max(__float_as_uint(Zs), __float_as_uint(sa)));
        if (!fs && sa > c.thr_rel) {
          fs = true;
          brk |= 1u << tt;
        }
        if (!fs) {
          if (r > c.low_rel) volj *= c.up_h - r * c.inv_h;
          theta += (double)(r * cl * c.J(r, xx) * volj);
        }
      }
    };
#pragma unroll
    for (int tt = 0; tt < KB; tt++) bond(tt);
    wnew |= brk << (kb & 31);
  }
  if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
  const double mi = s.mx[i];
  const double th = 3. * theta / mi;
  s.theta[i] = th;
  s.Z[i] = Zs * c.inv_s0;
  const double alpha = 15. * c.G / mi;
  return make_float2((float)alpha, (float)(th * ((3. * c.K / mi) - alpha / 3.)));
}

// Force of the node in thread slot q (demModel.cpp:682-788,
// PdState::getBondEF material.h:1415-1429) in mixed precision:
// ef_i + ef_j = J (r (b_i + b_j) + cl (a_i + a_j)).
template <int DIM, int INF, bool LDS>
__device__ __forceinline__ void force_node(const DS& s, const MixedArgs& m, const Win<DIM, LDS>& w,
                                           const StateConsts<INF>& c, int q) {
#pragma clang fp contract(fast)
  const int i = m.slot[q];
  const int ip = i - w.gs;
  const uint4* cb = (const uint4*)(m.col16 + m.sptr16[q / WAVE] + (q % WAVE) * KB);
  const unsigned* bw = s.bits + s.bptr[i / WAVE] + i % WAVE;
  const int n = s.deg[i];
  const float4 xvi = w.xv(ip);
  const float xi[3] = {xvi.x, xvi.y, xvi.z};
  const auto own = w.own(ip);
  const float us = w.us;
  const float2 abi = w.ab(ip, own.b);
  const float ai = abi.x, bi = abi.y;
  const float us2 = us * us, us2x = 2.f * us;
  double force[3] = {0., 0., 0.};
  unsigned word = 0;
  uint4 blk = n > 0 ? cb[0] : make_uint4(0, 0, 0, 0);
  for (int kb = 0; kb < n; kb += KB) {
    if ((kb & 31) == 0) word = bw[(int64_t)(kb >> 5) * WAVE];
    const unsigned wblk = word >> (kb & 31);  // this block's 8 bond bits from bit 0
    const uint4 cur = blk;
    if (kb + KB < n) blk = cb[(int64_t)(kb / KB + 1) * WAVE];
    float fblk[3] = {0.f, 0.f, 0.f};
    auto bond = [&](const int tt) {
      if (kb + tt < n) {
        const int jp = block_id(cur, tt);
        const float4 xvj = w.xv(jp);
        const float xj[3] = {xvj.x, xvj.y, xvj.z};
        float dx[3], du[3], y[3], bj;
        w.template du<true>(jp, own, du, bj);  // in units of us
        for (int d = 0; d < DIM; d++) {
          dx[d] = xj[d] - xi[d];
          y[d] = fmaf(us, du[d], dx[d]);
        }
        float volj = xvj.w;
        float xx = dx[0] * dx[0], xe = dx[0] * du[0], ee = du[0] * du[0];
        for (int d = 1; d < DIM; d++) {
          xx += dx[d] * dx[d];
          xe += dx[d] * du[d];
          ee += du[d] * du[d];
        }
        const float t2 = fmaf(us2, ee, us2x * xe);  // |y|^2 - |xi|^2
        const float yy = xx + t2;
        const float inv_r = rsqf_hw(xx), inv_ly = rsqf_hw(yy);
        const float r = xx * inv_r, ly = yy * inv_ly;
        const float cl = t2 * rcpf(ly + r);
        const bool fs = (wblk & (1u << tt)) != 0u;
        float sf;
        if (!fs) {
          if (r > c.low_rel) volj *= c.up_h - r * c.inv_h;
          const float Jr = c.J(r, xx);
          const float2 abj = w.ab(jp, bj);
          sf = Jr * (r * (bi + abj.y) + cl * (ai + abj.x)) * volj * inv_ly;
        } else {
          sf = c.Kn * volj * (ly - c.Rc) * inv_ly;
          if (sf > 0.f) sf = 0.f;
        }
        for (int d = 0; d < DIM; d++) fblk[d] += sf * y[d];
      }
    };
#pragma unroll
    for (int tt = 0; tt < KB; tt++) bond(tt);
    for (int d = 0; d < DIM; d++) force[d] += (double)fblk[d];
  }
  for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
}

// Tile ranges of MixedArgs::tiles: closed LDS tiles (whole particles, fused
// kernel), open LDS tiles, HBM tiles.
__device__ __forceinline__ const DevTile& tile_of(const MixedArgs& m, bool lds) {
  return m.tiles[(lds ? m.ntiles_closed : m.ntiles_closed + m.ntiles_lds) + blockIdx.x];
}

// LDS of the window records (LDS tiles) or a dummy (HBM tiles)
#define PDEM_WINDOW_LDS(LDS_ON, NODES)                 \
  __shared__ float4 lxv[(LDS_ON) ? (NODES) : 1];       \
  __shared__ int4 lq[(LDS_ON) ? (NODES) : 1];          \
  __shared__ float la[(LDS_ON) ? (NODES) : 1];         \
  __shared__ float red[1024 / WAVE];

// Dilation pass of the open tiles (tiles that do not hold a whole particle:
// the force pass of a tile reads the dilation terms of nodes of other tiles).
template <int DIM, int INF, bool LDS, int TPB, int NODES = LDS_NODES>
__global__ void __launch_bounds__(TPB, MIXED_WAVES_PER_SIMD) k_mixed_dilation(DS s, MixedArgs m) {
  PDEM_WINDOW_LDS(LDS, NODES)
  const DevTile t = tile_of(m, LDS);
  const DevParticle& p = s.P[t.pid];
  const int gs = (int)p.glob_start;
  Win<DIM, LDS> w{lxv, lq, la, &s, &m, gs, gs - t.ws, 1.f, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][gs];
  if (LDS) w.us = stage_window<DIM, false>(s, m, t, lxv, lq, la, w.org, red);
  const StateConsts<INF> c(s, p);
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) {
    int i;
    const float2 ab = dilation_node<DIM, INF, LDS>(s, m, w, c, q, i);
    m.fa[i] = ab.x;
    m.fb[i] = ab.y;
  }
}

// Force pass of the open tiles.
template <int DIM, int INF, bool LDS, int TPB, int NODES = LDS_NODES>
__global__ void __launch_bounds__(TPB, MIXED_WAVES_PER_SIMD) k_mixed_force(DS s, MixedArgs m) {
  PDEM_WINDOW_LDS(LDS, NODES)
  const DevTile t = tile_of(m, LDS);
  const DevParticle& p = s.P[t.pid];
  const int gs = (int)p.glob_start;
  Win<DIM, LDS> w{lxv, lq, la, &s, &m, gs, gs - t.ws, 1.f, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][gs];
  if (LDS) w.us = stage_window<DIM, true>(s, m, t, lxv, lq, la, w.org, red);
  const StateConsts<INF> c(s, p);
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) force_node<DIM, INF, LDS>(s, m, w, c, q);
}

// Both passes (demModel.cpp:610-679 and :682-788) of a closed tile, a whole
// particle, so that every neighbor of its nodes is one of its nodes, in one
// workgroup: the window is staged in LDS once, the dilation pass leaves each
// node's force-pass terms in its LDS record, and after a barrier the force
// pass reads them there.  Nothing of the dilation pass goes through HBM but
// theta.
template <int DIM, int INF, int TPB, int NODES = LDS_NODES>
__global__ void __launch_bounds__(TPB, MIXED_WAVES_PER_SIMD) k_mixed_state(DS s, MixedArgs m) {
  PDEM_WINDOW_LDS(true, NODES)
  const DevTile t = m.tiles[blockIdx.x];
  const DevParticle& p = s.P[t.pid];
  const int gs = (int)p.glob_start;
  Win<DIM, true> w{lxv, lq, la, &s, &m, gs, gs - t.ws, 1.f, {0., 0., 0.}};
  for (int d = 0; d < DIM; d++) w.org[d] = s.xr[d][gs];
  w.us = stage_window<DIM, false>(s, m, t, lxv, lq, la, w.org, red);
  const StateConsts<INF> c(s, p);
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) {
    int i;
    const float2 ab = dilation_node<DIM, INF, true>(s, m, w, c, q, i);
    la[i - t.ws] = ab.x;
    lq[i - t.ws].w = __float_as_int(ab.y);
  }
  __syncthreads();
  for (int q = t.ts + threadIdx.x; q < t.te; q += blockDim.x) force_node<DIM, INF, true>(s, m, w, c, q);
}

#undef PDEM_WINDOW_LDS

// the force reset of nodes that take no PD force (a list built at set-up);
// grid-stride
constexpr int BZ_BLOCKS = 1024;
template <int DIM>
__global__ void k_zero_noforce(DS s) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = gid(); k < s.n_noforce; k += stride) {
    const int64_t i = s.noforce_nodes[k];
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
  }
}

// copy of f of the nodes that take no PD force (save != 0: into buf, else
// back from it), around a standalone PD force pass; grid-stride
template <int DIM>
__global__ void k_noforce_copy(DS s, double* buf, int save) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = gid(); k < s.n_noforce; k += stride) {
    const int64_t i = s.noforce_nodes[k];
    for (int d = 0; d < DIM; d++) {
      if (save)
        buf[d * s.n_noforce + k] = s.f[d][i];
      else
        s.f[d][i] = buf[d * s.n_noforce + k];
    }
  }
}

}  // namespace pdem
