// Bond-based RNP force pass with each bond's scalar evaluated once per node
// pair (FP64 parity mode).
//
// The RNP bond law (RnpMaterial::getBondEF, material.h:342-368) depends on
// the pair only through quantities that are the same from both ends: the
// bond length r, the linearised stretch S = (xi . eta) / |xi|^2 (both vectors
// change sign together), Sc = rbar / sqrt(r), J(r), the break test and
// |S / Sc|.  A workgroup owns a tile of consecutive nodes of one particle:
//   1. every pair (i, j) with i < j inside the tile is evaluated once, one
//      lane per pair (the exp, four of the eight divides and a square root);
//      its force scalar ef (FP64) and |S / Sc| rounded to FP32 (negative when
//      the bond breaks now) go to LDS;
//   2. every node then runs its neighbor list in order, as
//      computePeridynamicForces does (demModel.cpp:682-788): in-tile bonds
//      read the pair record and add ef V_j (dx / r), out-of-tile bonds are
//      evaluated in full.  Sums, the break bits and the damage maximum
//      follow the list order, so the result equals the one-lane-per-node
//      kernel bit for bit.  The damage maximum keeps a value only as float,
//      so the FP32 |S / Sc| is the value it stores; the relative compare that
//      decides whether it is stored is taken on the FP32 value when that
//      cannot change its outcome (margin 2^-22 of the value, eight times the
//      rounding error) and on the FP64 value, recomputed, otherwise.
// Tiles, the pair list of every tile and the pair slot of every in-tile bond
// (uint16) are built once at set-up (PD lists never change).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdem {

constexpr int RT_THREADS = 512;   // workgroup size; tiles hold at most this many nodes
constexpr int RT_MAX_PAIRS = 6144;
constexpr uint16_t RT_OUT = 0xFFFF;

struct RtTile {
  int t0, t1;  // node range [t0, t1), inside one particle
  int npairs;  // in-tile pairs (i < j)
  int work;    // 1: RNP particle with forces; 0: zero the force only
  int64_t pbase;  // first pair of the tile in the pair arrays
};

struct RnpTileArgs {
  const RtTile* tiles = nullptr;
  const uint32_t* pl = nullptr;    // [pairs] tile-local (i | j << 16), i < j, in slot order
  const uint16_t* ps = nullptr;    // sliced ELL like col: pair slot of in-tile bonds, RT_OUT otherwise
  const uint16_t* perm = nullptr;  // [node t0 + q]: tile-local node of pass-2 lane q (degree order)
  int ntiles = 0;
  size_t lds_bytes = 0;
};

template <int DIM, int INF>
__global__ void __launch_bounds__(RT_THREADS) k_pd_rnp_tile(DS s, RnpTileArgs a) {
  extern __shared__ double lds[];
  const RtTile t = a.tiles[blockIdx.x];
  const int n_t = t.t1 - t.t0;
  if (!t.work) {
    for (int q = threadIdx.x; q < n_t; q += blockDim.x) {
      const int64_t i = t.t0 + q;
      if (!s.P[s.ptid[i]].compute_force)
        for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
    }
    return;
  }
  double* pe = lds;                       // [npairs] ef
  double* sx = lds + t.npairs;            // [DIM][n_t] reference coordinates
  double* su = sx + DIM * n_t;            // [DIM][n_t] displacements
  double* sv = su + DIM * n_t;            // [n_t] volumes
  float* pz = (float*)(sv + n_t);         // [npairs] +-|S / Sc| as float
  for (int q = threadIdx.x; q < n_t; q += blockDim.x) {
    const int64_t i = t.t0 + q;
    for (int d = 0; d < DIM; d++) {
      sx[d * n_t + q] = s.xr[d][i];
      su[d * n_t + q] = s.u[d][i];
    }
    sv[q] = s.vol[i];
  }
  const DevParticle& p = s.P[s.ptid[t.t0]];
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const double Cc = p.mat[0], beta = p.mat[1], rbar = p.mat[2], invF = p.mat[3], fsc = p.mat[4];
  const bool irrev = p.irreversible != 0;
  const double Kn = p.Kn, Rc = p.Rc;
  __syncthreads();

  // 1. pair records, one lane per pair
  for (int q = threadIdx.x; q < t.npairs; q += blockDim.x) {
    const uint32_t ij = a.pl[t.pbase + q];
    const int iq = (int)(ij & 0xFFFFu), jq = (int)(ij >> 16);
    double dx[3], du[3];
    for (int d = 0; d < DIM; d++) {
      dx[d] = sx[d * n_t + jq] - sx[d * n_t + iq];
      du[d] = su[d * n_t + jq] - su[d * n_t + iq];
    }
    const double rji = len3<DIM>(dx);
    const double Sji = dot3<DIM>(dx, du) / dot3<DIM>(dx, dx);
    const double Sc = rbar / sqrt(rji);
    const double J = infJ<INF>(s, 0., rji, horizon);
    const bool brk = irrev && isGreater(fabs(Sji), fsc * Sc);
    double ef = 0.;
    if (!brk) ef = J * 4. * Sji * Cc * beta * exp(-beta * rji * Sji * Sji) / invF;
    const double z = fabs(Sji / Sc);
    pe[q] = ef;
    pz[q] = brk ? -(float)z : (float)z;
  }
  __syncthreads();

  // 2. per-node sums in list order; lanes take the tile's nodes in degree
  // order, so a wavefront's lists have about the same length
  for (int qq = threadIdx.x; qq < n_t; qq += blockDim.x) {
    const int q = a.perm[t.t0 + qq];
    const int64_t i = t.t0 + q;
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const int* __restrict__ c = s.col + s.sptr[sl] + lane;
    const uint16_t* __restrict__ ps = a.ps + s.sptr[sl] + lane;
    unsigned* bw = s.bits + s.bptr[sl] + lane;
    const int n = s.deg[i];
    double xi[3], ui[3];
    for (int d = 0; d < DIM; d++) {
      xi[d] = sx[d * n_t + q];
      ui[d] = su[d * n_t + q];
    }
    double force[3] = {0., 0., 0.};
    float Zi = 0.f;
    unsigned word = 0, wnew = 0;
    for (int k = 0; k < n; k++) {
      if ((k & 31) == 0) {
        if (k > 0 && wnew != word) bw[(int64_t)((k >> 5) - 1) * WAVE] = wnew;
        word = bw[(int64_t)(k >> 5) * WAVE];
        wnew = word;
      }
      const int j = c[(int64_t)k * WAVE];
      const int slot = ps[(int64_t)k * WAVE];
      const bool fs0 = (word >> (k & 31)) & 1u;
      double dx[3], xj[3], uj[3];
      double volj, z;
      bool zexact = true;  // z holds the FP64 value (else its FP32 rounding)
      if (slot != RT_OUT) {
        const int jq = j - t.t0;
        for (int d = 0; d < DIM; d++) {
          xj[d] = sx[d * n_t + jq];
          uj[d] = su[d * n_t + jq];
          dx[d] = xj[d] - xi[d];
        }
        volj = sv[jq];
        const float zs = pz[slot];
        z = (double)fabsf(zs);
        zexact = false;
        if (!fs0) {
          const double rji = len3<DIM>(dx);
          if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
          double ef = pe[slot];
          if (zs < 0.f) {  // breaks in this step: no force from it now
            wnew |= 1u << (k & 31);
            ef = 0.;
          }
          const double scalar_f = ef * volj;
          for (int d = 0; d < DIM; d++) force[d] += scalar_f * (dx[d] / rji);
        }
      } else {
        double du[3];
        for (int d = 0; d < DIM; d++) {
          xj[d] = s.xr[d][j];
          uj[d] = s.u[d][j];
          dx[d] = xj[d] - xi[d];
          du[d] = uj[d] - ui[d];
        }
        volj = s.vol[j];
        const double rji = len3<DIM>(dx);
        const double Sji = dot3<DIM>(dx, du) / dot3<DIM>(dx, dx);
        const double Sc = rbar / sqrt(rji);
        z = fabs(Sji / Sc);
        if (!fs0) {
          if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
          const double J = infJ<INF>(s, 0., rji, horizon);
          bool fs = false;
          if (irrev && isGreater(fabs(Sji), fsc * Sc)) fs = true;
          double ef = 0.;
          if (!fs) ef = J * 4. * Sji * Cc * beta * exp(-beta * rji * Sji * Sji) / invF;
          if (fs) wnew |= 1u << (k & 31);
          const double scalar_f = ef * volj;
          for (int d = 0; d < DIM; d++) force[d] += scalar_f * (dx[d] / rji);
        }
      }
      if (fs0) {
        // bond broken before this step: short-range repulsion (demModel.cpp:760-768)
        double y[3];
        for (int d = 0; d < DIM; d++) y[d] = (xj[d] + uj[d]) - (xi[d] + ui[d]);
        const double R = len3<DIM>(y);
        double scalar_f = Kn * volj * (R - Rc) / R;
        if (scalar_f > 0.) scalar_f = 0.;
        for (int d = 0; d < DIM; d++) force[d] += scalar_f * y[d];
      }
      if (zexact) {
        if (isGreater(z, (double)Zi)) Zi = (float)z;
      } else {
        // z is |S / Sc| rounded to float: decide on it unless the compare is
        // within reach of that rounding, then on the recomputed FP64 value
        const double Zd = (double)Zi;
        const double g = (z - Zd) - ((z < Zd ? Zd : z) * COMPARE_EPS);
        if (z != 0. && fabs(g) <= 0x1p-22 * z) {
          double du[3];
          for (int d = 0; d < DIM; d++) du[d] = uj[d] - ui[d];
          const double rji = len3<DIM>(dx);
          const double Sji = dot3<DIM>(dx, du) / dot3<DIM>(dx, dx);
          const double zx = fabs(Sji / (rbar / sqrt(rji)));
          if (isGreater(zx, Zd)) Zi = (float)zx;
        } else if (g > 0.) {
          Zi = (float)z;
        }
      }
    }
    if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
    for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
    s.Z[i] = Zi;
  }
}

}  // namespace pdem
