// pdem_closure.hpp -- learned pair-contact closure for 2D particle pairs.
//
// For every pair of non-wall particles listed at the last contact rebuild
// (a < b, centres within the damping reach plus twice the search radius) the
// closure kernel works in the pair frame n = (c_b - c_a) / |c_b - c_a|:
//
//   extents   ext_a = max (x_i - c_a).n over a's surface nodes, ext_b likewise
//             along -n; gap = |c_b - c_a| - ext_a - ext_b.  Every node pair
//             (i in a, j in b) is at least `gap` apart, so gap >= Rc means no
//             contact force is possible (mode 0, exact).
//   patches   surface nodes within Rc of the extents; for gap >= 0 every node
//             pair closer than Rc is a patch pair, so the nearest patch pair
//             dmin >= Rc also means no contact force (mode 0, exact).
//   gating    penetration Rc - dmin <= pen_max_h * min(h_a, h_b), gap >= 0,
//             no damage in either particle (max Z < 1) and a predicted local
//             damage below dz_max: closure (mode 1); otherwise the plain
//             node-level contact kernels handle the pair (mode 2).
//   network   16 features -> (F.n, F.t, M_a, local damage) in units of
//             Kn Rc Vbar_a Vbar_b (and h_a for M), BF16 MFMA (pdem_mlp.hpp).
//   transfer  the pair force F on a (-F on b) and the moments M_a, M_b go to
//             the patch nodes in proportion to their overlap weights (the
//             spatial pattern of the plain node-level forces) plus a torque
//             term with zero net force, so each patch receives exactly F and M.
//
// A sampled check every `check_every` steps evaluates the plain node-level
// pair force of every closure pair from the candidate lists and accumulates
// the normalised difference.  k_pair_log writes the same features and the
// plain targets for training (pairs within a band around contact).
//
// Lanes: one wavefront per pair: the 64 lanes split its node loops (b's
// contact patch staged in LDS) and meet through shuffles; every lane ends
// with the pair's features, so the 16 rows of the network's MFMA fragment
// are copies of the pair.  Only pairs whose particles had a candidate node
// pair at the last rebuild are evaluated.
#pragma once
#include "pdem_kernels.hpp"
#include "pdem_mlp.hpp"

namespace pdem {

struct ClosureArgs {
  mlp::Net net;
  int use_net;                // 0: no network (closure band empty: gating only)
  double pen_max_h;           // band: Rc - gap <= pen_max_h * min(h_a, h_b)
  float dz_max;               // predicted local damage above this: plain
  int check_every;            // sampled plain check cadence in steps (0: never)
  const int* ref_node;        // [nn] orientation reference node
  const double* vbar;         // [nn] mean nodal volume
  const float* code;          // [nn] particle code (feature)
  unsigned long long* stats;  // [3] pairs per mode (cumulative), [3] checked pairs, [4] max error bits
  double* err_sum;            // [0] sum |F_cl - F_plain| / Fs, [1] sum |F_plain| / Fs
};

constexpr int CL_NFEAT = 16;
constexpr double CL_V0 = 1.0;  // velocity scale of the features, m/s

// reductions over the 64 lanes of a wavefront (fixed butterfly order)
__device__ __forceinline__ double wave_max(double v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

struct PairGeo {
  int a, b, pa, pb, sa;
  double ca[2], cb[2], n[2], d;
  double ext_a, ext_b, gap, rc, ha, hb;
  float za, zb;  // largest damage Z of each particle (this step)
  DevContact ct;
};

// largest (x_i - c).dir over the surface nodes of non-wall particle q
__device__ __forceinline__ double extent(const DS& s, int q, const double* c, const double* dir, int lane) {
  double m = -1e300;
  for (int k = s.surf_ptr[q] + lane; k < s.surf_ptr[q + 1]; k += WAVE) {
    const int i = s.surf_idx[k];
    m = fmax(m, (s.x[0][i] - c[0]) * dir[0] + (s.x[1][i] - c[1]) * dir[1]);
  }
  return wave_max(m);
}

// largest Z over the nodes of particle P
__device__ __forceinline__ float zmax(const DS& s, const DevParticle& P, int lane) {
  float z = 0.f;
  for (int64_t i = P.glob_start + lane; i < P.glob_end; i += WAVE) z = fmaxf(z, s.Z[i]);
  for (int o = WAVE / 2; o > 0; o >>= 1) z = fmaxf(z, __shfl_xor(z, o));
  return z;
}

// Geometry of the pair (a, dlist[a][sa]) in the pair frame; every lane ends
// with the same values.  Every node pair (i in a, j in b) is at least gap apart.
__device__ __forceinline__ void pair_geometry(const DS& s, int a, int sa, int lane, PairGeo& g) {
  const int nn = s.n_nonwall;
  g.a = a;
  g.sa = sa;
  g.b = s.dlist[(int64_t)g.a * s.dcap + g.sa];
  g.pa = s.nonwall_pid[g.a];
  g.pb = s.nonwall_pid[g.b];
  for (int d = 0; d < 2; d++) {
    g.ca[d] = s.pcx[d * nn + g.a];
    g.cb[d] = s.pcx[d * nn + g.b];
  }
  const double dx = g.cb[0] - g.ca[0], dy = g.cb[1] - g.ca[1];
  g.d = sqrt(dx * dx + dy * dy);
  const bool ok = g.d > 0.;
  g.n[0] = ok ? dx / g.d : 1.;
  g.n[1] = ok ? dy / g.d : 0.;
  const double nb[2] = {-g.n[0], -g.n[1]};
  g.ct = s.CT[s.pzone[g.a] * s.nz + s.pzone[g.b]];
  g.rc = g.ct.contact_r;
  const DevParticle& Pa = s.P[g.pa];
  const DevParticle& Pb = s.P[g.pb];
  g.ha = Pa.h;
  g.hb = Pb.h;
  g.ext_a = extent(s, g.a, g.ca, g.n, lane);
  g.ext_b = extent(s, g.b, g.cb, nb, lane);
  g.gap = ok ? g.d - g.ext_a - g.ext_b : -1e300;  // coincident centres: plain
  g.za = g.zb = 0.f;  // pair_damage fills them where the gating needs them
}

// largest damage Z of both particles of the pair
__device__ __forceinline__ void pair_damage(const DS& s, int lane, PairGeo& g) {
  g.za = zmax(s, s.P[g.pa], lane);
  g.zb = zmax(s, s.P[g.pb], lane);
}

__device__ __forceinline__ double omega(const DS& s, const ClosureArgs& ca, int q, const double* c) {
  const int nn = s.n_nonwall;
  const int r = ca.ref_node[q];
  const double rx = s.x[0][r] - c[0], ry = s.x[1][r] - c[1];
  const double vx = s.v[0][r] - s.pcv[q], vy = s.v[1][r] - s.pcv[nn + q];
  const double r2 = rx * rx + ry * ry;
  return r2 > 0. ? (rx * vy - ry * vx) / r2 : 0.;
}

// Nearest node pair and overlap of the contact patches (surface nodes within
// Rc of the extents).  With gap >= 0 every node pair closer than Rc is a
// patch pair (a node outside a's patch is more than gap + Rc from every node
// of b along n), so dmin >= Rc means no contact force (exact).
// Overlap weight of a patch pair: c_ij = (1 - R_ij / Rc) V_i V_j for
// R_ij < Rc.  W = sum c_ij; Wn, Wt: the weights times the pair direction
// e_ij (i to j) along n and t, Wm = sum c_ij (r_i x e_ij): the normal part of
// the plain contact force on a is -Kn Rc (Wn n + Wt t), its moment about c_a
// -Kn Rc Wm (the network predicts the remainder: friction and corrections);
// the first moments of the weights about each centre (P.sa, P.sb) and second
// moments (P.saa, P.sbb) give the weighted patch centroids and polar moments
// of the force transfer.
struct PatchSums {
  double dmin, W, Wn, Wt, Wm, sa[2], saa, sb[2], sbb;
};

// b's patch nodes staged per wavefront (position and volume); a pair with a
// larger patch is left to the plain method
constexpr int CL_PATCH = 128;
struct PatchStage {
  double x[CL_PATCH], y[CL_PATCH], v[CL_PATCH];
};

// Patch sums of the pair with b's patch staged in `st`; false when b's patch
// does not fit (the caller then treats the pair as plain)
__device__ __forceinline__ bool patch_contact(const DS& s, const PairGeo& g, int lane, PatchStage& st,
                                              PatchSums& P) {
  const double thr_a = g.ext_a - g.rc, thr_b = g.ext_b - g.rc;
  int nb = 0;
  for (int m = s.surf_ptr[g.b] + lane; m - lane < s.surf_ptr[g.b + 1]; m += WAVE) {
    bool in = false;
    double xj0 = 0., xj1 = 0.;
    int j = 0;
    if (m < s.surf_ptr[g.b + 1]) {
      j = s.surf_idx[m];
      xj0 = s.x[0][j];
      xj1 = s.x[1][j];
      in = -((xj0 - g.cb[0]) * g.n[0] + (xj1 - g.cb[1]) * g.n[1]) >= thr_b;
    }
    const unsigned long long msk = __ballot(in);
    const int pos = nb + __popcll(msk & ((1ull << lane) - 1ull));
    if (in && pos < CL_PATCH) {
      st.x[pos] = xj0;
      st.y[pos] = xj1;
      st.v[pos] = s.vol[j];
    }
    nb += __popcll(msk);
  }
  if (nb > CL_PATCH) return false;
  __builtin_amdgcn_wave_barrier();  // LDS writes of the wavefront before its reads
  double dm = 1e300, w = 0., wn = 0., wt = 0., wm = 0., a0 = 0., a1 = 0., aa = 0., b0s = 0., b1s = 0., bb = 0.;
  for (int k = s.surf_ptr[g.a] + lane; k < s.surf_ptr[g.a + 1]; k += WAVE) {
    const int i = s.surf_idx[k];
    const double xi0 = s.x[0][i], xi1 = s.x[1][i];
    const double ri0 = xi0 - g.ca[0], ri1 = xi1 - g.ca[1];
    if (ri0 * g.n[0] + ri1 * g.n[1] < thr_a) continue;
    const double Vi = s.vol[i];
    for (int m = 0; m < nb; m++) {
      const double xj0 = st.x[m], xj1 = st.y[m];
      const double rj0 = xj0 - g.cb[0], rj1 = xj1 - g.cb[1];
      const double dx = xj0 - xi0, dy = xj1 - xi1;
      const double R = sqrt(dx * dx + dy * dy);
      dm = fmin(dm, R);
      if (R < g.rc) {
        const double c = (1. - R / g.rc) * Vi * st.v[m];
        w += c;
        // direction of the node pair (i to j) in the pair frame
        wn += c * (dx * g.n[0] + dy * g.n[1]) / R;
        wt += c * (-dx * g.n[1] + dy * g.n[0]) / R;
        wm += c * (ri0 * dy - ri1 * dx) / R;
        a0 += c * ri0;
        a1 += c * ri1;
        aa += c * (ri0 * ri0 + ri1 * ri1);
        b0s += c * rj0;
        b1s += c * rj1;
        bb += c * (rj0 * rj0 + rj1 * rj1);
      }
    }
  }
  __builtin_amdgcn_wave_barrier();  // the stage is reused by the wavefront's next pair
  P.dmin = -wave_max(-dm);
  P.W = wave_sum(w);
  P.Wn = wave_sum(wn);
  P.Wt = wave_sum(wt);
  P.Wm = wave_sum(wm);
  P.sa[0] = wave_sum(a0);
  P.sa[1] = wave_sum(a1);
  P.saa = wave_sum(aa);
  P.sb[0] = wave_sum(b0s);
  P.sb[1] = wave_sum(b1s);
  P.sbb = wave_sum(bb);
  return true;
}

__device__ __forceinline__ void pair_features(const DS& s, const ClosureArgs& ca, const PairGeo& g, double dmin,
                                              const PatchSums& P, float* x16) {
  const double vv = ca.vbar[g.a] * ca.vbar[g.b];
  const int nn = s.n_nonwall;
  const double t[2] = {-g.n[1], g.n[0]};
  const double dvx = s.pcv[g.b] - s.pcv[g.a], dvy = s.pcv[nn + g.b] - s.pcv[nn + g.a];
  const double Ra = s.prad[g.a], Rb = s.prad[g.b];
  x16[0] = (float)(g.gap / g.rc);
  x16[1] = (float)(fmin(g.ha, g.hb) / g.rc);
  x16[2] = (float)(P.Wt / vv);
  x16[3] = (float)(Ra / g.ha);
  x16[4] = (float)(Rb / g.hb);
  x16[5] = (float)(g.ext_a / Ra);
  x16[6] = (float)(g.ext_b / Rb);
  x16[7] = (float)((dvx * g.n[0] + dvy * g.n[1]) / CL_V0);
  x16[8] = (float)((dvx * t[0] + dvy * t[1]) / CL_V0);
  x16[9] = (float)(omega(s, ca, g.a, g.ca) * Ra / CL_V0);
  x16[10] = (float)(omega(s, ca, g.b, g.cb) * Rb / CL_V0);
  x16[11] = ca.code[g.a];
  x16[12] = ca.code[g.b];
  x16[13] = (float)g.ct.mu;
  x16[14] = (float)(dmin / g.rc);
  x16[15] = (float)(P.Wn / vv);
}

// normal-contact baseline of a pair in network units (F.n, F.t, M_a / h_a):
// the network's targets are the plain values minus this baseline
__device__ __forceinline__ void pair_baseline(const ClosureArgs& ca, const PairGeo& g, const PatchSums& P,
                                              double* b3) {
  const double vv = ca.vbar[g.a] * ca.vbar[g.b];
  b3[0] = -P.Wn / vv;
  b3[1] = -P.Wt / vv;
  b3[2] = -P.Wm / (vv * g.ha);
}

// force unit of a pair: Kn Rc Vbar_a Vbar_b
__device__ __forceinline__ double pair_unit(const ClosureArgs& ca, const PairGeo& g) {
  return g.ct.Kn * g.rc * ca.vbar[g.a] * ca.vbar[g.b];
}

// Plain net contact force and moment (about c_a) on particle a from particle
// b: the node-level terms of computeContactForces (demModel.cpp:868-904)
// summed with the nodal volumes, from the candidate lists; zloc = largest Z
// of a node of a with an active contact with b.
__device__ __forceinline__ void plain_pair(const DS& s, const PairGeo& g, int lane, double* F, double& M,
                                           double& zloc) {
  double f0 = 0., f1 = 0., m = 0., z = 0.;
  const DevParticle& P = s.P[g.pa];
  const DevContact ct = g.ct;
  for (int64_t i = P.glob_start + lane; i < P.glob_end; i += WAVE) {
    const int nc = s.ccount[i];
    if (nc == 0) continue;
    double yi[2] = {s.x[0][i], s.x[1][i]}, vi[2] = {s.v[0][i], s.v[1][i]};
    double fi[2] = {0., 0.};
    bool any = false;
    for (int k = 0; k < nc; k++) {
      const int j = s.cand[cand_idx(s.N, i, k)];
      if (s.ptid[j] != g.pb) continue;
      double yji[2] = {s.x[0][j] - yi[0], s.x[1][j] - yi[1]};
      const double R = len3<2>(yji);
      if (!isLess(R, ct.contact_r)) continue;
      double fn[2], ft[2];
      contact_pair<2>(s, vi, j, ct, R, yji, fn, ft);
      fi[0] += fn[0];
      fi[1] += fn[1];
      fi[0] += ft[0];
      fi[1] += ft[1];
      any = true;
    }
    if (any) {
      const double V = s.vol[i];
      f0 += V * fi[0];
      f1 += V * fi[1];
      m += V * ((yi[0] - g.ca[0]) * fi[1] - (yi[1] - g.ca[1]) * fi[0]);
      z = fmax(z, (double)s.Z[i]);
    }
  }
  F[0] = wave_sum(f0);
  F[1] = wave_sum(f1);
  M = wave_sum(m);
  zloc = wave_max(z);
}

// Record of one side of a closure pair (CL_REC doubles): the pair direction
// (own centre to partner), own and partner patch thresholds, the force per
// unit overlap weight g = F / W, the torque coefficient k and the weighted
// patch centroid (relative to the own centre).  A patch node i receives
// w_i (g + k perp(r_i - xbar)), w_i = sum_j (1 - R_ij / Rc) V_j over the
// partner's patch nodes closer than Rc: the patch then carries exactly F and,
// about the centre, M (the torque term has no net force).
__device__ __forceinline__ void closure_record(double* rec, const double* dir, double thr_own, double thr_partner,
                                               const double* F, double M, double W, const double* s1, double s2,
                                               double rc, double h) {
  double g[2] = {0., 0.}, xb[2] = {0., 0.}, k = 0.;
  if (W > 0.) {
    g[0] = F[0] / W;
    g[1] = F[1] / W;
    xb[0] = s1[0] / W;
    xb[1] = s1[1] / W;
    // torque term only for patches wider than h / 4 (weighted spread), with
    // the extra moment limited to |F| h: a narrow patch cannot carry a couple
    const double I = s2 - W * (xb[0] * xb[0] + xb[1] * xb[1]);
    double dM = M - (xb[0] * F[1] - xb[1] * F[0]);
    const double lim = sqrt(F[0] * F[0] + F[1] * F[1]) * h;
    dM = fmax(-lim, fmin(lim, dM));
    k = I > 0.0625 * W * h * h ? dM / I : 0.;
  }
  rec[0] = dir[0];
  rec[1] = dir[1];
  rec[2] = thr_own;
  rec[3] = thr_partner;
  rec[4] = g[0];
  rec[5] = g[1];
  rec[6] = k;
  rec[7] = xb[0];
  rec[8] = xb[1];
  rec[9] = rc;
}

// slot of particle `target` in q's partner list, or -1 (all lanes)
__device__ __forceinline__ int find_slot(const DS& s, int q, int target, int lane) {
  const int* dl = s.dlist + (int64_t)q * s.dcap;
  const int n = s.dcount[q];
  for (int base = 0; base < n; base += WAVE) {
    const unsigned long long m = __ballot(base + lane < n && dl[base + lane] == target);
    if (m) return base + __ffsll((long long)m) - 1;
  }
  return -1;
}

// Pending record of a pair for the network (CL_PEND doubles): slot indices
// of b in a's list and of a in b's list (-1: none), the 16 features, the
// baseline (3), the force unit, n, c_b - c_a, the patch thresholds
// ext_a - Rc and ext_b - Rc, Rc, h_a, h_b, the patch sums W, sa (2), saa,
// sb (2), sbb, and on check steps the plain pair force (2) and moment.
constexpr int CL_PEND = 42;
constexpr int CLP_IA = 0, CLP_IB = 1, CLP_X = 2, CLP_B3 = 18, CLP_UNIT = 21, CLP_N = 22, CLP_D = 24, CLP_TA = 26,
              CLP_TB = 27, CLP_RC = 28, CLP_HA = 29, CLP_HB = 30, CLP_W = 31, CLP_SA = 32, CLP_SAA = 34,
              CLP_SB = 35, CLP_SBB = 37, CLP_FP = 38, CLP_MP = 40, CLP_CHK = 41;

// Geometry and gating, one wavefront per connected pair (grid-stride; the
// listed pairs without a candidate node pair at the last rebuild have no
// contact force until the next one, as the plain method only evaluates
// candidates: mode 0, their cmode entries were cleared at that rebuild).
// Modes 0 and 2 are final here; a pair in the closure band goes to the
// pending list, where k_cl_net runs the network and decides.
constexpr int CL_WAVES = 4;  // wavefronts per workgroup of the pair kernels
__global__ void __launch_bounds__(CL_WAVES * WAVE, 4) k_closure(DS s, ClosureArgs ca) {
  __shared__ PatchStage stage[CL_WAVES];
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const int nconn = min((int)s.ctrl->nconn, s.cpair_cap);
  const bool check = ca.check_every > 0 && (s.ctrl->n % ca.check_every) == 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int npairs = min((int)s.ctrl->npairs, s.cpair_cap);
    if (npairs > nconn) atomicAdd(&ca.stats[0], (unsigned long long)(npairs - nconn));
  }
  __shared__ unsigned counts[CL_WAVES][3];
  unsigned cnt[3] = {0u, 0u, 0u};  // pairs per final mode (lane 0), summed per workgroup at the end
  for (int k = blockIdx.x * CL_WAVES + wave; k < nconn; k += gridDim.x * CL_WAVES) {
    const int ia = s.conn_list[k];
    const int prev = s.cmode[ia];
    PairGeo g;
    pair_geometry(s, ia / s.dcap, ia % s.dcap, lane, g);
    bool near = g.gap < g.rc;
    PatchSums P{};
    bool fits = true;
    // patches: nearest node pair (exact no-contact test), overlap weights
    if (near && g.gap >= 0.) {
      fits = patch_contact(s, g, lane, stage[wave], P);
      if (fits && P.dmin >= g.rc) near = false;  // no node pair within Rc
    }
    const double dmin = (near && g.gap >= 0.) ? P.dmin : g.gap;
    int mode = 0;
    if (near) {
      mode = 2;
      if (ca.use_net && fits && g.gap >= 0. && g.rc - dmin <= ca.pen_max_h * fmin(g.ha, g.hb)) {
        pair_damage(s, lane, g);
        if (g.za < 1.f && g.zb < 1.f) mode = 1;
      }
    }
    if (mode == 0 && prev == 0) {  // both sides already hold mode 0
      cnt[0]++;
      continue;
    }
    const int sb = find_slot(s, g.b, g.a, lane);
    if (mode != 1) {
      if (lane == 0) {
        s.cmode[ia] = (int8_t)mode;
        if (sb >= 0) s.cmode[(int64_t)g.b * s.dcap + sb] = (int8_t)mode;
        if (mode == 2) {
          atomicOr(&s.pflags[g.a], 1);
          atomicOr(&s.pflags[g.b], 1);
        }
        if (sb < 0) s.ctrl->overflow = 6;
      }
      cnt[mode]++;
      continue;
    }
    double Fp[2] = {0., 0.}, Mp = 0., zl = 0.;
    if (check) plain_pair(s, g, lane, Fp, Mp, zl);
    if (lane == 0) {
      const unsigned row = atomicAdd(&s.ctrl->npend, 1u);
      double* r = s.cl_pend + (int64_t)row * CL_PEND;
      float x16[CL_NFEAT];
      pair_features(s, ca, g, dmin, P, x16);
      double b3[3];
      pair_baseline(ca, g, P, b3);
      r[CLP_IA] = (double)ia;
      r[CLP_IB] = sb >= 0 ? (double)((int64_t)g.b * s.dcap + sb) : -1.;
      for (int f = 0; f < CL_NFEAT; f++) r[CLP_X + f] = (double)x16[f];
      for (int q = 0; q < 3; q++) r[CLP_B3 + q] = b3[q];
      r[CLP_UNIT] = pair_unit(ca, g);
      for (int d = 0; d < 2; d++) {
        r[CLP_N + d] = g.n[d];
        r[CLP_D + d] = g.cb[d] - g.ca[d];
        r[CLP_SA + d] = P.sa[d];
        r[CLP_SB + d] = P.sb[d];
        r[CLP_FP + d] = Fp[d];
      }
      r[CLP_TA] = g.ext_a - g.rc;
      r[CLP_TB] = g.ext_b - g.rc;
      r[CLP_RC] = g.rc;
      r[CLP_HA] = g.ha;
      r[CLP_HB] = g.hb;
      r[CLP_W] = P.W;
      r[CLP_SAA] = P.saa;
      r[CLP_SBB] = P.sbb;
      r[CLP_MP] = Mp;
      r[CLP_CHK] = check ? 1. : 0.;
    }
  }
  if (lane == 0)
    for (int q = 0; q < 3; q++) counts[wave][q] = cnt[q];
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long t = 0;
    for (int w = 0; w < CL_WAVES; w++) t += counts[w][threadIdx.x];
    if (t) atomicAdd(&ca.stats[threadIdx.x], t);
  }
}

// The network for the pending pairs, 16 per wavefront (one row each):
// predicted local damage above dz_max sends a pair to the plain method,
// otherwise the predicted force and moment (remainder over the baseline) go
// to both sides' transfer records (mode 1).
__global__ void __launch_bounds__(CL_WAVES * WAVE) k_cl_net(DS s, ClosureArgs ca) {
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE, pl = lane & 15;
  const int npend = min((int)s.ctrl->npend, s.cpair_cap);
  for (int base = (blockIdx.x * CL_WAVES + wave) * 16; base < npend; base += gridDim.x * CL_WAVES * 16) {
    const int k = base + pl;
    const bool valid = k < npend;
    const double* r = s.cl_pend + (int64_t)(valid ? k : base) * CL_PEND;
    float x16[CL_NFEAT];
    for (int f = 0; f < CL_NFEAT; f++) x16[f] = (float)r[CLP_X + f];
    mlp::s4 h1[8], h2[8], h3[8];
    const mlp::f4 y = mlp::forward(ca.net, mlp::input_frag(ca.net, x16, lane), h1, h2, h3, lane);
    double ys[4];
    for (int q = 0; q < 4; q++) ys[q] = (double)(__shfl(y[q], pl) / ca.net.out_scale[q]);
    if (!valid || lane >= 16) continue;
    const int64_t ia = (int64_t)r[CLP_IA], ib = (int64_t)r[CLP_IB];
    const int a = (int)(ia / s.dcap), b = s.dlist[ia];
    const int mode = ys[3] > (double)ca.dz_max ? 2 : 1;
    s.cmode[ia] = (int8_t)mode;
    if (ib >= 0) s.cmode[ib] = (int8_t)mode;
    if (mode == 2 || ib >= 0) {
      const int bit = mode == 2 ? 1 : 2;
      atomicOr(&s.pflags[a], bit);
      atomicOr(&s.pflags[b], bit);
    }
    if (ib < 0) s.ctrl->overflow = 6;
    atomicAdd(&ca.stats[mode], 1ull);
    if (mode != 1) continue;
    const double unit = r[CLP_UNIT], n[2] = {r[CLP_N], r[CLP_N + 1]};
    const double Fn = (ys[0] + r[CLP_B3]) * unit, Ft = (ys[1] + r[CLP_B3 + 1]) * unit;
    const double Fa[2] = {Fn * n[0] - Ft * n[1], Fn * n[1] + Ft * n[0]};
    const double Ma = (ys[2] + r[CLP_B3 + 2]) * unit * r[CLP_HA];
    if (ib >= 0) {
      const double nb[2] = {-n[0], -n[1]};
      const double Fb[2] = {-Fa[0], -Fa[1]};
      // torque balance: M_b = -M_a + (c_b - c_a) x F_a
      const double Mb = -Ma + (r[CLP_D] * Fa[1] - r[CLP_D + 1] * Fa[0]);
      const double sa[2] = {r[CLP_SA], r[CLP_SA + 1]}, sb[2] = {r[CLP_SB], r[CLP_SB + 1]};
      closure_record(s.cout + ia * CL_REC, n, r[CLP_TA], r[CLP_TB], Fa, Ma, r[CLP_W], sa, r[CLP_SAA], r[CLP_RC],
                     r[CLP_HA]);
      closure_record(s.cout + ib * CL_REC, nb, r[CLP_TB], r[CLP_TA], Fb, Mb, r[CLP_W], sb, r[CLP_SBB], r[CLP_RC],
                     r[CLP_HB]);
    }
    if (r[CLP_CHK] != 0.) {
      const double e = sqrt((Fa[0] - r[CLP_FP]) * (Fa[0] - r[CLP_FP]) +
                            (Fa[1] - r[CLP_FP + 1]) * (Fa[1] - r[CLP_FP + 1])) / unit;
      const double fp = sqrt(r[CLP_FP] * r[CLP_FP] + r[CLP_FP + 1] * r[CLP_FP + 1]) / unit;
      atomicAdd(&ca.err_sum[0], e);
      atomicAdd(&ca.err_sum[1], fp);
      atomicAdd(&ca.stats[3], 1ull);
      atomicMax(&ca.stats[4], (unsigned long long)__double_as_longlong(e));
    }
  }
}

// Training rows: every listed pair with gap < Rc + band_h * min(h_a, h_b):
// features, the plain targets (F.n, F.t, M_a / h_a in force units, local Z)
// and the normal-contact baseline of the patches in the same units (zero when
// no patch pair is within Rc); the network learns targets minus baseline.
// Pairs whose partner patch does not fit the stage are not logged.
__global__ void __launch_bounds__(CL_WAVES * WAVE) k_pair_log(DS s, ClosureArgs ca, double band_h, float* feat,
                                                              float* targ, float* basel, unsigned* count, int cap) {
  __shared__ PatchStage stage[CL_WAVES];
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const int npairs = min((int)s.ctrl->npairs, s.cpair_cap);
  for (int k = blockIdx.x * CL_WAVES + wave; k < npairs; k += gridDim.x * CL_WAVES) {
    PairGeo g;
    pair_geometry(s, s.cpairs[2 * (int64_t)k], s.cpairs[2 * (int64_t)k + 1], lane, g);
    if (!(g.gap < g.rc + band_h * fmin(g.ha, g.hb))) continue;
    PatchSums P{};
    if (g.gap >= 0. && !patch_contact(s, g, lane, stage[wave], P)) continue;
    const double dmin = g.gap >= 0. ? P.dmin : g.gap;
    double Fp[2], Mp, zl;
    plain_pair(s, g, lane, Fp, Mp, zl);
    if (lane == 0) {
      const unsigned row = atomicAdd(count, 1u);
      if ((int)row < cap) {
        float x16[CL_NFEAT];
        pair_features(s, ca, g, dmin, P, x16);
        for (int f = 0; f < CL_NFEAT; f++) feat[(int64_t)row * CL_NFEAT + f] = x16[f];
        const double unit = pair_unit(ca, g);
        double b3[3];
        pair_baseline(ca, g, P, b3);
        targ[(int64_t)row * 4 + 0] = (float)((Fp[0] * g.n[0] + Fp[1] * g.n[1]) / unit);
        targ[(int64_t)row * 4 + 1] = (float)((-Fp[0] * g.n[1] + Fp[1] * g.n[0]) / unit);
        targ[(int64_t)row * 4 + 2] = (float)(Mp / (unit * g.ha));
        if (basel)
          for (int r = 0; r < 3; r++) basel[(int64_t)row * 3 + r] = (float)b3[r];
        targ[(int64_t)row * 4 + 3] = (float)zl;
      }
    }
  }
}

// Pair connections at a rebuild (whatever the step when `force`: set-up):
// each candidate node pair (i in non-wall particle a, j in non-wall particle
// b) stamps the partner slot of max(a, b) in min(a, b)'s list with the
// rebuild count, and the first stamp appends the slot to conn_list.  One lane
// per node; consecutive candidates of one particle share the search.
__global__ void k_cl_connect(DS s, int force) {
  if (gid() == 0) s.ctrl->npend = 0;  // k_closure fills the pending list of this step
  if (!force && !s.ctrl->rebuild) return;
  const int ep = (int)s.ctrl->n_rebuilds;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gid(); i < s.N; i += stride) {
    const int qa = s.P[s.ptid[i]].nonwall_index;
    const int cnt = qa >= 0 ? s.ccount[i] : 0;
    int prev = -1;
    for (int k = 0; k < cnt; k++) {
      const int qb = s.P[s.ptid[s.cand[cand_idx(s.N, i, k)]]].nonwall_index;
      if (qb < 0 || qb == prev) continue;
      prev = qb;
      const int lo = qa < qb ? qa : qb, hi = qa < qb ? qb : qa;
      const int* dl = s.dlist + (int64_t)lo * s.dcap;
      int l = 0, h = s.dcount[lo] - 1;
      while (l <= h) {
        const int mid = (l + h) >> 1;
        const int v = dl[mid];
        if (v == hi) {
          const int64_t slot = (int64_t)lo * s.dcap + mid;
          // the first stamp of this rebuild lists the pair
          if (atomicExch(&s.cconn[slot], ep) != ep) {
            const unsigned r = atomicAdd(&s.ctrl->nconn, 1u);
            if ((int)r < s.cpair_cap) s.conn_list[r] = (int)slot;
          }
          break;
        }
        if (v < hi)
          l = mid + 1;
        else
          h = mid - 1;
      }
    }
  }
}

constexpr int CL_CONNECT_BLOCKS = 1024;  // grid of k_cl_connect (grid-stride over nodes)

// Closure lists from the current candidate and partner lists (set-up and
// logging; afterwards the rebuilds maintain them): wall_near per particle
// (stamped with the rebuild count) and the pair list.
__global__ void k_cl_wall_near(DS s) {
  const int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& P = s.P[s.ptid[i]];
  if (P.nonwall_index < 0) return;
  const int cnt = s.ccount[i];
  for (int k = 0; k < cnt; k++)
    if (s.pmeta[s.ptid[s.cand[cand_idx(s.N, i, k)]]] >> 16) {
      s.wall_near[P.nonwall_index] = (int)s.ctrl->n_rebuilds;
      return;
    }
}
__global__ void k_cl_pairs(DS s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.n_nonwall) return;
  const int cnt = s.dcount[q];
  for (int k = 0; k < cnt; k++)
    if (s.dlist[(int64_t)q * s.dcap + k] > q) {
      const unsigned r = atomicAdd(&s.ctrl->npairs, 1u);
      if ((int)r < s.cpair_cap) {
        s.cpairs[2 * (int64_t)r] = q;
        s.cpairs[2 * (int64_t)r + 1] = k;
      } else {
        s.ctrl->overflow = 5;
      }
    }
}

}  // namespace pdem
