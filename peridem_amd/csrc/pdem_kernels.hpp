// pdem_kernels.hpp -- gfx950 kernels of the explicit PeriDEM step.
//
// Data layout in HBM (structure of arrays, FP64 state):
//   xr[d], x[d], u[d], v[d], f[d]   one array per component, N doubles each
//   vol, mx, theta, vmag            N doubles;  Z: N floats;  fix: N bytes;  ptid: N ints
//   PD neighbor list: sliced ELL, slices of 64 consecutive nodes (one wavefront):
//     col[sptr[s] + k*64 + lane]    neighbor k of node 64 s + lane, ascending node id
//     bits[bptr[s] + (k/32)*64 + lane] bit k%32 = bond broken
//   so every load of a wavefront's k-th neighbor id or bond word is one
//   contiguous 256-byte (ids) / 256-byte (words) access, and a lane owns its
//   bond words (no atomics on the parity path).
//   contact candidates: blocked ELL, count ccount[i]: candidate k of node i
//   at cand[cand_idx(N, i, k)] = cand[((k / CAND_BLOCK) N + i) CAND_BLOCK +
//   k % CAND_BLOCK], so a node's candidates come in 32-byte runs (a
//   wavefront taking one node's candidates reads 8 runs per 64 candidates)
//   and the runs of consecutive nodes lie side by side (a wavefront taking
//   one candidate rank of consecutive nodes reads 2 KB in one piece).
//
// FP64 kernels are compiled with -ffp-contract=off and evaluate every
// expression in the reference's operation order, so they reproduce the
// reference's x86-64 FP64 arithmetic (exp() may differ in the last ulp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cfloat>
#include <climits>

namespace pdem {

constexpr int CAND_BLOCK = 8;  // contact candidates per run of the blocked layout
__host__ __device__ inline int64_t cand_idx(int64_t N, int64_t i, int k) {
  return ((int64_t)(k / CAND_BLOCK) * N + i) * CAND_BLOCK + (k % CAND_BLOCK);
}

constexpr int WAVE = 64;
constexpr int LIGHT_MAX_C = 16;  // candidate count up to which one lane takes a node
constexpr double COMPARE_EPS = 1e-5;
constexpr int CL_REC = 10;       // doubles per closure force record

struct DevParticle {
  int64_t glob_start, glob_end, center;
  int zone, type_index, compute_force, material_type;
  int irreversible, searches, nonwall_index, pad;
  double rho, horizon, h, Rb, Kn, Rc;
  double mat[6];
};

struct DevContact {
  double contact_r, Kn, mu, betan, kappa;
};

// device-resident step state: time, cadence controller of the contact search
struct Ctrl {
  int64_t n;
  double time;
  double dt;
  double current_dt;
  int64_t interval;
  int64_t counter;
  double radius;
  double max_velocity;
  double max_contact_r;
  double s_factor;
  int64_t deck_interval;
  int64_t n_rebuilds;
  int64_t init_calls;
  int32_t rebuild;
  int32_t overflow;
  int32_t vdirty;
  uint32_t nactive;  // nodes with more than LIGHT_MAX candidates (wavefront per node)
  uint32_t nlight;   // nodes with 1..LIGHT_MAX candidates (lane per node)
  uint32_t npairs;   // closure: particle pairs listed at the last rebuild
  uint32_t nconn;    // closure: listed pairs with a candidate node pair at the last rebuild
  uint32_t npend;    // closure: pairs of this step waiting for the network
  uint32_t pad_;
  unsigned long long vmax_acc;  // bit pattern of max |v| (non-negative doubles order as integers)
  double vmax_last;
  // scratch results
  double pp_pen, pp_reac;
  unsigned long long broken;
  int64_t restart_step;   // DEMModel::restart: the restart step (restart_active)
  int32_t restart_active;
  // the force pass left f without the damping and gravity terms: the next
  // node update adds them (as k_finalize would, in its order), and a host
  // read of f first runs k_finalize over the pending forces
  int32_t f_pending;
  int32_t pad2_;
  // contact skin: an upper bound of how far any node has moved since the
  // last rebuild's query (infinite when unknown), and the squared moves of
  // this step's displacement-BC kernels (bit pattern, atomic maximum)
  double dsum;
  unsigned long long mv_bc_acc;
};

// displacement-BC set parameters and the per-node entries of the sets that
// are applied every step (zero sets excluded), for the fused node update
struct BcSetParams {
  int time_fn, dir_mask;
  double tp[4];
};
struct BcEntry {
  double amp;
  int set;
  int pad;
};

struct DS {
  int64_t N;
  int np, nz;
  double* xr[3];
  double* x[3];
  double* u[3];
  double* v[3];
  double* f[3];
  double *vol, *mx, *theta, *vmag;
  float* Z;
  uint8_t* fix;
  int* ptid;
  const DevParticle* P;
  const DevContact* CT;
  int inf_type;
  double inf_p0, inf_p1;
  int state_global;  // particle 0 is state-based: run the dilation pass
  // PD neighbors, sliced ELL
  const int64_t* sptr;
  const int64_t* bptr;
  int* deg;
  int* col;
  unsigned* bits;
  // contact candidates
  int cap;
  int* ccount;
  int* cand;
  int* active;  // [N] node ids with many candidates, appended at each rebuild
  int* light;   // [N] node ids with few candidates
  // cell grid
  double glo[3];
  double cell;
  int gn[3];
  int csub;  // cells per largest search radius: the query scans +-csub cells
  int64_t ncells;
  // Cell list of the last rebuild: the nodes sorted by cell key (a counting
  // sort: per-cell counts, an exclusive scan over the cells in blocks of
  // CELL_BLOCK, a scatter).  Cell c holds the cell_count[c] entries from
  // cell_boff[c / CELL_BLOCK] + cell_start[c] on; no cell has a capacity.
  int* cell_count;  // [ncells]
  int* cell_start;  // [ncells] exclusive prefix of the counts inside the cell's block
  int* cell_btot;   // [nblocks] count total of each block of cells
  int* cell_boff;   // [nblocks] exclusive prefix of the block totals
  int64_t ncell_blocks;
  int* node_cell;   // [N] cell of each node (this rebuild)
  int* node_slot;   // [N] rank of the node inside its cell
  int* cell_nodes;  // [N] node ids by cell
  float4* cell_xf;  // [N] beside cell_nodes: position - glo in FP32, particle id (bits)
  int* cell_owner;  // [ncells] particle id when one particle fills the cell, -1 mixed, -2 empty
  // Candidate positions for the contact loops of the steps between rebuilds:
  // one 32-byte record per node (x, y, z in FP64 and the particle id's bits
  // in the fourth slot), so a candidate costs one gather instead of four
  // arrays.  Every node update (k_integrate, k_ubc) rewrites, after moving
  // them, the records of the nodes of particles that do not search (walls:
  // any of their nodes may be a candidate) and of the nodes marked at the
  // last rebuild (xa_epoch[i] equal to the low byte of the rebuild count: the
  // query marks each searching node that has candidates, which by the
  // symmetry of the FP64 distance test are the searching nodes that are
  // candidates of another).  On a rebuild step the contact loops read the
  // node arrays.
  double4* xa;        // [N]
  uint8_t* xa_epoch;  // [N]
  // damping
  int n_nonwall;
  const int* nonwall_pid;
  const int64_t* nonwall_center;  // centre node of each non-wall particle
  int nwall, nwords;
  const int* widx;
  const int* wall_gid;
  unsigned* wbits;
  int* wlo;
  int* whi;
  double* fdamp;  // [np][3]
  // per non-wall particle, packed each step for the damping pass
  double* pcx;      // [3][n_nonwall] centre position
  double* pcv;      // [3][n_nonwall] centre velocity
  const double* prad;  // [n_nonwall] bounding radius
  const double* prho;  // [n_nonwall] density
  const int* pzone;    // [n_nonwall] zone
  // damping partners per non-wall particle, rebuilt with the candidate lists:
  // particles whose centres were within damping reach plus twice the search
  // radius, ascending
  int dcap;
  int* dcount;      // [n_nonwall]
  int* dlist;       // [n_nonwall][dcap] non-wall indices
  // per particle: zone | wall << 16, for the contact loops
  const int* pmeta;
  double rc2_max;  // largest contact radius squared: distance prefilter
  // contact skin: a node whose nearest candidate at the last rebuild
  // (cdmin, rounded down) minus twice the bound of node moves since then
  // (Ctrl::dsum) exceeds rc_skip (the largest contact radius and a margin)
  // has no candidate within contact distance, and the contact loops skip it
  double rc_skip;
  int n_ubc_sets;  // displacement-BC sets (launches of k_ubc per step at most)
  float* cdmin;  // [N]
  unsigned long long* blockmv;  // per node-update workgroup: max squared move (bit pattern)
  int light_max;   // nodes with at most this many candidates take one lane
  // learned pair-contact closure (cl_on == 0: none of these is read)
  int cl_on;
  int* cpairs;          // [cpair_cap][2] (a, slot): pair a < dlist[a][slot], listed at rebuilds
  int cpair_cap;
  int8_t* cmode;        // [n_nonwall][dcap] 0: no contact possible, 1: closure, 2: plain
  int* cconn;           // [n_nonwall][dcap] rebuild count at which a candidate node pair joined a and dlist[a][slot]
  int* conn_list;       // [cpair_cap] a * dcap + slot of the connected pairs (a < dlist[a][slot])
  double* cl_pend;      // [cpair_cap][CL_PEND] pairs of this step waiting for the network
  double* cout;         // [n_nonwall][dcap][CL_REC] closure force records
  int* pflags;          // [n_nonwall] this step: bit 0 some pair is plain, bit 1 some pair is closure
  int* wall_near;       // [n_nonwall] rebuild count at which a node had a wall candidate
  const uint8_t* surf;  // [N] surface node flag
  const int* surf_ptr;  // [n_nonwall + 1] surface nodes per non-wall particle
  const int* surf_idx;
  // node-update block maxima of |v|
  unsigned long long* blockmax;
  int nblockmax;
  // nodes of wall particles, ascending (finalisation inside the damping pass)
  const int64_t* wall_nodes;
  int64_t n_wall_nodes;
  const int64_t* noforce_nodes;  // nodes of particles that compute no peridynamic force
  int64_t n_noforce;
  int nf_zero;  // the node update zeroes f of nodes without peridynamic force (mixed bond kernels)
  // per-node displacement-BC entries (CSR over nodes, set order)
  const int* bc_ptr;
  const BcEntry* bc_ent;
  const BcSetParams* bc_sets;
  double gravity[3];
  int has_gravity;
  Ctrl* ctrl;
};

// --------------------------------------------------------------- helpers
__device__ __forceinline__ bool isGreater(double a, double b) {
  double fa = fabs(a), fb = fabs(b);
  return (a - b) > ((fa < fb ? fb : fa) * COMPARE_EPS);
}
__device__ __forceinline__ bool isLess(double a, double b) {
  double fa = fabs(a), fb = fabs(b);
  return (b - a) > ((fa < fb ? fb : fa) * COMPARE_EPS);
}
__device__ __forceinline__ double infFn(const DS& s, double r) {
  if (s.inf_type == 0) return s.inf_p0;
  if (s.inf_type == 1) return s.inf_p0 + s.inf_p1 * r;
  return s.inf_p0 * exp(-r * r / s.inf_p1);
}
template <int DIM>
__device__ __forceinline__ double len3(const double* a) {
  return DIM == 3 ? sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) : sqrt(a[0] * a[0] + a[1] * a[1]);
}
template <int DIM>
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return DIM == 3 ? a[0] * b[0] + a[1] * b[1] + a[2] * b[2] : a[0] * b[0] + a[1] * b[1];
}

__device__ __forceinline__ int64_t gid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// ------------------------------------------------- PD neighbor list build
// j in the same particle, j != i, 0 + dx^2 + dy^2 + dz^2 < horizon^2
// (nanoflann L2_Simple_Adaptor, strict <), scanned in ascending j.
template <int DIM>
__global__ void k_pd_count(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  const double r2 = p.horizon * p.horizon;
  double xi[3];
  for (int d = 0; d < DIM; d++) xi[d] = s.xr[d][i];
  int c = 0;
  for (int64_t j = p.glob_start; j < p.glob_end; j++) {
    if (j == i) continue;
    double d2 = 0.;
    for (int d = 0; d < DIM; d++) {
      double t = xi[d] - s.xr[d][j];
      d2 += t * t;
    }
    if (d2 < r2) c++;
  }
  s.deg[i] = c;
}

template <int DIM>
__global__ void k_pd_fill(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  const double r2 = p.horizon * p.horizon;
  double xi[3];
  for (int d = 0; d < DIM; d++) xi[d] = s.xr[d][i];
  const int64_t sl = i / WAVE, lane = i % WAVE;
  int* out = s.col + s.sptr[sl] + lane;
  int k = 0;
  for (int64_t j = p.glob_start; j < p.glob_end; j++) {
    if (j == i) continue;
    double d2 = 0.;
    for (int d = 0; d < DIM; d++) {
      double t = xi[d] - s.xr[d][j];
      d2 += t * t;
    }
    if (d2 < r2) {
      out[(int64_t)k * WAVE] = (int)j;
      k++;
    }
  }
}

// per node: smallest and largest neighbor id (tile windows of the mixed kernels)
__global__ void k_nbr_range(DS s, int* lo, int* hi) {
  int64_t i = gid();
  if (i >= s.N) return;
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* c = s.col + s.sptr[sl] + lane;
  int a = (int)i, b = (int)i;
  for (int k = 0; k < s.deg[i]; k++) {
    const int j = c[(int64_t)k * WAVE];
    a = j < a ? j : a;
    b = j > b ? j : b;
  }
  lo[i] = a;
  hi[i] = b;
}

// neighbor ids relative to the particle's first node, 16 bit, in blocks of
// 8 per lane (layout of the mixed-precision kernels), in thread-slot order:
// slot q holds node slot[q], and slot slices of 64 are padded to their
// largest degree
__global__ void k_col16(DS s, uint16_t* col16, const int64_t* sptr16, const int* slot) {
  int64_t q = gid();
  if (q >= s.N) return;
  const int64_t i = slot[q];
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int64_t gs = s.P[s.ptid[i]].glob_start;
  const int64_t base = sptr16[q / WAVE] + (q % WAVE) * 8;
  for (int k = 0; k < s.deg[i]; k++) {
    const int64_t o = s.sptr[sl] + (int64_t)k * WAVE + lane;
    col16[base + (int64_t)(k / 8) * WAVE * 8 + (k % 8)] = (uint16_t)(s.col[o] - gs);
  }
}

// m_x = sum_j r^2 J(r/delta) V_j^c  (materialUtil.cpp:21-73)
template <int DIM>
__global__ void k_mx(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* c = s.col + s.sptr[sl] + lane;
  const int n = s.deg[i];
  double xi[3];
  for (int d = 0; d < DIM; d++) xi[d] = s.xr[d][i];
  double m = 0.;
  for (int k = 0; k < n; k++) {
    int j = c[(int64_t)k * WAVE];
    double dx[3];
    for (int d = 0; d < DIM; d++) dx[d] = s.xr[d][j] - xi[d];
    double rji = len3<DIM>(dx);
    if (isGreater(rji, horizon)) continue;
    double volj = s.vol[j];
    if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
    m += rji * rji * infFn(s, rji / horizon) * volj;
  }
  s.mx[i] = m;
}

// ---------------------------------------------- contact search controller
// DEMModel::updateContactNeighborSearchParameters (demModel.cpp:1808-1955):
// decides whether this step rebuilds the candidate lists.  Run by one thread.
__device__ void ctrl_decide(Ctrl* cg) {
  Ctrl h = *cg;
  Ctrl* c = &h;
  if (c->vdirty) {
    c->vmax_last = __longlong_as_double((long long)c->vmax_acc);
    c->vdirty = 0;
  }
  bool upd;
  const int64_t deckI = c->deck_interval;
  if (c->interval == 0 && isLess(c->radius, 1.e-16)) {
    c->interval = deckI;
    c->counter = c->n % c->interval;
    c->radius = c->max_contact_r * c->s_factor;
  }
  if (c->n == 0 && (c->init_calls + 1) <= 2) {
    c->init_calls++;
    if (c->init_calls == 1) {
      upd = true;
    } else {
      c->counter++;
      upd = (c->counter - 1) % c->interval == 0;
    }
  } else {
    if (c->n == 0) c->init_calls++;
    // restart step (demModel.cpp:1833-1837)
    if (c->restart_active && c->n == c->restart_step) c->counter = c->n % c->interval;
    if (c->interval == 1) {
      c->radius = c->max_contact_r;
      c->counter++;
      upd = (c->counter - 1) % c->interval == 0;
    } else {
      int64_t upi = c->interval > 5 ? (int64_t)(0.2 * (double)c->interval) : 1;
      if (c->counter > 0 && c->counter % upi != 0) {
        c->counter++;
        upd = (c->counter - 1) % c->interval == 0;
      } else {
        c->max_velocity = c->vmax_last;
        double safety = c->s_factor > 5 ? c->s_factor : 10;
        double r_contact = c->s_factor * c->max_contact_r;
        double r_vel = c->max_velocity * c->current_dt * (double)deckI * safety;
        if (isGreater(r_vel, r_contact)) {
          c->interval = (int64_t)(c->max_contact_r / (c->max_velocity * c->current_dt));
          c->radius = r_contact;
          c->counter = 0;
          if (c->interval < 1) {
            c->interval = 1;
            c->radius = c->max_contact_r;
          }
        } else {
          c->radius = c->interval < 2 ? c->max_contact_r : r_contact;
        }
        c->counter++;
        upd = (c->counter - 1) % c->interval == 0;
      }
    }
  }
  c->rebuild = upd ? 1 : 0;
  if (upd) {
    c->n_rebuilds++;
    c->nactive = 0u;
    c->nlight = 0u;
    c->npairs = 0u;
    c->nconn = 0u;
  }
  *cg = h;
}

// Controller step.  Workgroup 0 reduces the node-update block maxima of |v|
// and runs the controller; all workgroups pack the non-wall particles' centre
// data (and clear the closure's per-step flags).  The cell table is cleared
// after each rebuild's query (k_contact_force), so no step clears it here.
constexpr int CTRL_BLOCKS = 64;

__global__ void __launch_bounds__(1024) k_ctrl(DS s, int advance) {
  __shared__ unsigned long long red[1024 / WAVE], redm[1024 / WAVE];
  Ctrl* c = s.ctrl;
  const int t = threadIdx.x;
  // the node update before this kernel consumed any pending forces, and the
  // force pass after it writes new ones
  if (blockIdx.x == 0 && t == 0) c->f_pending = 0;
  if (blockIdx.x == 0) {
    if (advance) {
      // the fused node update left the time advance to this kernel
      if (t == 0) {
        c->current_dt = c->dt;
        c->n += 1;
        c->time += c->dt;
        c->vdirty = 1;
      }
      __syncthreads();
    }
    const bool moved = c->vdirty != 0;  // a node update ran since the last controller step
    unsigned long long mvb = 0;           // its largest squared node move (bit pattern)
    if (moved) {
      // eight loads in flight per thread (one workgroup reads every block maximum)
      unsigned long long m = 0, mv = 0;
      const int nb = s.nblockmax, stride = (int)blockDim.x;
      for (int k0 = t; k0 < nb; k0 += 8 * stride) {
        unsigned long long v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int k = k0 + u * stride;
          v[u] = k < nb ? s.blockmax[k] : 0ull;
          w[u] = k < nb ? s.blockmv[k] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          m = v[u] > m ? v[u] : m;
          mv = w[u] > mv ? w[u] : mv;
        }
      }
      for (int off = WAVE / 2; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off);
        m = o > m ? o : m;
        o = __shfl_xor(mv, off);
        mv = o > mv ? o : mv;
      }
      if ((t & (WAVE - 1)) == 0) {
        red[t / WAVE] = m;
        redm[t / WAVE] = mv;
      }
      __syncthreads();
      if (t == 0) {
        unsigned long long mm = 0;
        for (int k = 0; k < (int)(blockDim.x / WAVE); k++) {
          mm = red[k] > mm ? red[k] : mm;
          mvb = redm[k] > mvb ? redm[k] : mvb;
        }
        c->vmax_acc = mm;
      }
    }
    if (t == 0) {
      ctrl_decide(c);
      // contact skin: a rebuild's query sees the positions of now; otherwise
      // every node has moved at most this step's largest node-update move
      // plus, per displacement-BC set, that set's largest move (a relative
      // margin covers the rounding of the bound)
      double step = (double)s.n_ubc_sets * sqrt(__longlong_as_double((long long)c->mv_bc_acc));
      if (moved) step += sqrt(__longlong_as_double((long long)mvb));
      c->dsum = c->rebuild ? 0. : c->dsum + step * (1. + 1e-9);
      c->mv_bc_acc = 0ull;
    }
  }
  if (advance) return;  // centre data packed by the fused node update
  const int nn = s.n_nonwall;
  for (int q = blockIdx.x * blockDim.x + t; q < nn; q += gridDim.x * blockDim.x) {
    const int64_t cn = s.nonwall_center[q];
    for (int d = 0; d < 3; d++) {
      s.pcx[d * nn + q] = s.x[d] ? s.x[d][cn] : 0.;
      s.pcv[d * nn + q] = s.v[d] ? s.v[d][cn] : 0.;
    }
    if (s.cl_on) s.pflags[q] = 0;
  }
}

// empty the cell table (all cells, also at set-up)
__device__ __forceinline__ void clear_cells(const DS& s, int64_t first, int64_t stride) {
  for (int64_t k = first; k < s.ncells; k += stride) {
    s.cell_count[k] = 0;
    s.cell_owner[k] = -2;
  }
}
__global__ void k_cells_init(DS s) { clear_cells(s, gid(), (int64_t)gridDim.x * blockDim.x); }

// ParticleULoading::apply (particleULoading.cpp:115-230) for one node of a set
// (TRIG false: only the time functions 0-2, without the sine and rotation
// code, whose FP64 sin/cos raise a kernel's register count)
// Returns the squared distance the node moved.
template <int DIM, bool TRIG = true>
__device__ __forceinline__ double apply_ubc(const DS& s, int time_fn, int dir_mask, const double* tp, double umax,
                                            int64_t i, double time) {
  double du = 0., dv = 0.;
  if (time_fn == 0)
    du = umax;
  else if (time_fn == 1) {
    du = umax * time;
    dv = umax;
  } else if (time_fn == 2) {
    du = umax * time + tp[1] * time * time;
    dv = umax + tp[1] * time;
  } else if (TRIG && time_fn == 3) {
    double a = M_PI * tp[1];
    du = umax * sin(a * time);
    dv = umax * a * cos(a * time);
  }
  double ui[3] = {0., 0., 0.}, vi[3] = {0., 0., 0.};
  for (int d = 0; d < 3; d++)
    if (dir_mask >> d & 1) {
      ui[d] = du;
      vi[d] = dv;
    }
  if (TRIG && time_fn == 4) {
    double dx0 = s.xr[0][i] - tp[1], dx1 = s.xr[1][i] - tp[2];
    double dx2 = (DIM == 3 ? s.xr[2][i] : 0.) - tp[3];
    double th = tp[0] * time;
    double rx0 = dx0 * cos(th) - dx1 * sin(th), rx1 = dx0 * sin(th) + dx1 * cos(th);
    double d0 = -dx0 * sin(th) - dx1 * cos(th), d1 = dx0 * cos(th) - dx1 * sin(th);
    ui[0] += rx0 - dx0;
    ui[1] += rx1 - dx1;
    ui[2] += 0.0 - dx2;
    vi[0] += tp[0] * d0;
    vi[1] += tp[0] * d1;
    vi[2] += tp[0] * 0.0;
  }
  double m2 = 0.;
  for (int d = 0; d < DIM; d++)
    if (dir_mask >> d & 1) {
      s.u[d][i] = ui[d];
      s.v[d][i] = vi[d];
      const double xn = ui[d] + s.xr[d][i];
      const double dd = xn - s.x[d][i];
      m2 += dd * dd;
      s.x[d][i] = xn;
    }
  return m2;
}

// the candidate-position record of node i from the node arrays
template <int DIM>
__device__ __forceinline__ void xa_store(const DS& s, int64_t i, int pid) {
  s.xa[i] = make_double4(s.x[0][i], s.x[1][i], DIM == 3 ? s.x[2][i] : 0., __longlong_as_double((long long)pid));
}
// after a node update: rewrite the record of a wall node or of a node marked
// at the last rebuild
template <int DIM>
__device__ __forceinline__ void xa_refresh(const DS& s, int64_t i, int pid, bool searches) {
  if (s.xa && (!searches || s.xa_epoch[i] == (uint8_t)s.ctrl->n_rebuilds)) xa_store<DIM>(s, i, pid);
}
// every node's record (after positions are set from the host or restored)
template <int DIM>
__global__ void k_xa_all(DS s) {
  const int64_t i = gid();
  if (i < s.N) xa_store<DIM>(s, i, s.ptid[i]);
}

// positions set from the host or restored: the contact skin bound is
// unknown until the next rebuild
__global__ void k_skin_invalidate(DS s) { s.ctrl->dsum = INFINITY; }

// ------------------------------------------------------------ node update
// integrateCD (demModel.cpp:349-390): semi-implicit Euler on free dofs;
// with `half`, the first half kick and drift of integrateVerlet
// (demModel.cpp:392-425): v += 0.5 (dt / rho) f.
// With `fused` (multi-particle models), the same launch also applies the
// displacement-BC sets of the step (computeExternalDisplacementBC, per node
// in set order, at the advanced time t + dt) and packs the non-wall
// particles' centre data for the damping pass; the time advance is then left
// to the controller kernel that follows (k_ctrl with `advance`).  TRIG false:
// no set uses the sine or rotation time function.
template <int DIM, bool fused, bool TRIG = true>
__global__ void __launch_bounds__(256) k_integrate(DS s, int multi, int half) {
  int64_t i = gid();
  Ctrl* ct = s.ctrl;
  const double dt = ct->dt;
  const bool pend = ct->f_pending != 0;  // k_finalize's terms still to add (same order)
  double vm = 0.;
  double mv2 = 0.;  // squared distance the node moves in this launch (a bound for BC nodes)
  if (i < s.N) {
    const int pid = s.ptid[i];
    const DevParticle& p = s.P[pid];
    const uint8_t fx = s.fix[i];
    // the node's candidate-position record is rewritten below (read before
    // any store: a store could alias the flag)
    const bool rec = multi && s.xa && (!p.searches || s.xa_epoch[i] == (uint8_t)ct->n_rebuilds);
    // the fused update's BC entries and centre flag, also read before any store
    int e0 = 0, e1 = 0;
    bool centre = false;
    double t_new = 0.;
    if (fused) {
      e0 = s.bc_ptr[i];
      e1 = s.bc_ptr[i + 1];
      centre = p.type_index == 0 && p.center == i;
      t_new = ct->time + dt;
    }
    double xs[3] = {0., 0., 0.};
    bool xs_ok = false;
    if (p.compute_force) {
      const double rho = p.rho;
      const double a = half ? 0.5 * (dt / rho) : dt / rho;
      // every input is loaded before the first store: the stores go through
      // double pointers that may alias the arrays read, so a load after a
      // store would wait for the stores to complete
      double vd[3] = {0., 0., 0.}, fd[3] = {0., 0., 0.}, ud[3] = {0., 0., 0.};
      for (int d = 0; d < DIM; d++) {
        vd[d] = s.v[d][i];
        xs[d] = s.x[d][i];
        if (!(fx >> d & 1)) {
          fd[d] = s.f[d][i];
          ud[d] = s.u[d][i];
          if (pend) {
            if (p.type_index == 0) fd[d] += s.fdamp[3 * pid + d];
            if (s.has_gravity) fd[d] += p.rho * s.gravity[d];
          }
        }
      }
      for (int d = 0; d < DIM; d++)
        if (!(fx >> d & 1)) {
          vd[d] += a * fd[d];
          s.v[d][i] = vd[d];
          s.u[d][i] = ud[d] + dt * vd[d];
          const double xn = xs[d] + dt * vd[d];
          const double dd = xn - xs[d];
          mv2 += dd * dd;
          xs[d] = xn;
          s.x[d][i] = xs[d];
        }
      xs_ok = true;
      vm = len3<DIM>(vd);
      s.vmag[i] = vm;
    } else if (s.nf_zero) {
      // the step's force pass starts from zero here (its bond kernels skip
      // these nodes); nothing reads their f before it
      for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
    }
    if (fused) {
      if (e1 > e0) {
        xs_ok = false;  // moved by the displacement BC below
        // the node's move is at most the sum of its k + 1 partial moves:
        // squared, at most (k + 1) times the sum of their squares
        double m2 = mv2;
        for (int e = e0; e < e1; e++) {
          const BcEntry be = s.bc_ent[e];
          const BcSetParams& bs = s.bc_sets[be.set];
          m2 += apply_ubc<DIM, TRIG>(s, bs.time_fn, bs.dir_mask, bs.tp, be.amp, i, t_new);
        }
        mv2 = (double)(e1 - e0 + 1) * m2;
      }
      if (centre) {
        const int nn = s.n_nonwall, q = p.nonwall_index;
        for (int d = 0; d < 3; d++) {
          s.pcx[d * nn + q] = d < DIM ? s.x[d][i] : 0.;
          s.pcv[d * nn + q] = d < DIM ? s.v[d][i] : 0.;
        }
        if (s.cl_on) s.pflags[q] = 0;
      }
    }
    if (rec) {
      if (!xs_ok)
        for (int d = 0; d < DIM; d++) xs[d] = s.x[d][i];
      s.xa[i] = make_double4(xs[0], xs[1], xs[2], __longlong_as_double((long long)pid));
    }
  }
  if (!multi) {
    if (i == 0) {
      ct->current_dt = dt;
      ct->n += 1;
      ct->time += dt;
    }
    return;
  }
  // block maxima of |v| and of the squared move (non-negative doubles order
  // like their bit patterns), reduced over blocks by the controller
  __shared__ unsigned long long wmax[256 / WAVE], wmv[256 / WAVE];
  unsigned long long b = (unsigned long long)__double_as_longlong(vm);
  unsigned long long bm = (unsigned long long)__double_as_longlong(mv2);
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
    o = __shfl_xor(bm, off);
    bm = o > bm ? o : bm;
  }
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    wmax[w] = b;
    wmv[w] = bm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0, mm = 0;
    for (int k = 0; k < (int)(blockDim.x / WAVE); k++) {
      m = wmax[k] > m ? wmax[k] : m;
      mm = wmv[k] > mm ? wmv[k] : mm;
    }
    s.blockmax[blockIdx.x] = m;
    s.blockmv[blockIdx.x] = mm;
    if (!fused && blockIdx.x == 0) {
      ct->current_dt = dt;
      ct->n += 1;
      ct->time += dt;
      ct->vdirty = 1;
    }
  }
}

// second half kick of integrateVerlet (demModel.cpp:437-459) with the new
// forces: v += 0.5 (dt / rho) f on free dofs, |v|.
template <int DIM>
__global__ void __launch_bounds__(256) k_kick(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  if (!p.compute_force) return;
  const double a = 0.5 * (s.ctrl->dt / p.rho);
  const uint8_t fx = s.fix[i];
  double vv[3] = {0., 0., 0.};
  for (int d = 0; d < DIM; d++) {
    double vd = s.v[d][i];
    if (!(fx >> d & 1)) {
      vd += a * s.f[d][i];
      s.v[d][i] = vd;
    }
    vv[d] = vd;
  }
  s.vmag[i] = len3<DIM>(vv);
}

// ---------------------------------------------------------- boundary data
struct DevBc {
  int time_fn, dir_mask;
  double tp[4];
  int64_t n;
  const int64_t* nodes;
  const double* amp;
};

// (the largest squared move goes to Ctrl::mv_bc_acc for the contact skin: a
// node in several sets moves at most the sum of their moves, which the
// controller bounds by the number of sets times the largest)
template <int DIM>
__global__ void __launch_bounds__(256) k_ubc(DS s, DevBc bc) {
  __shared__ unsigned long long wm[256 / WAVE];
  int64_t k = gid();
  double m2 = 0.;
  if (k < bc.n) {
    const int64_t i = bc.nodes[k];
    m2 = apply_ubc<DIM>(s, bc.time_fn, bc.dir_mask, bc.tp, bc.amp[k], i, s.ctrl->time);
    if (s.xa) {
      const int pid = s.ptid[i];
      xa_refresh<DIM>(s, i, pid, s.P[pid].searches != 0);
    }
  }
  unsigned long long b = (unsigned long long)__double_as_longlong(m2);
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) wm[threadIdx.x / WAVE] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int w = 0; w < (int)(blockDim.x / WAVE); w++) m = wm[w] > m ? wm[w] : m;
    if (m > 0) atomicMax(&s.ctrl->mv_bc_acc, m);
  }
}

// ParticleFLoading::apply (particleFLoading.cpp:85-173)
template <int DIM>
__global__ void k_fbc(DS s, DevBc bc) {
  int64_t k = gid();
  if (k >= bc.n) return;
  const double time = s.ctrl->time;
  double fmax = bc.amp[k];
  if (bc.time_fn == 1)
    fmax *= time;
  else if (bc.time_fn == 2) {
    double x1 = bc.tp[1], x2 = bc.tp[2];
    double period = floor(time / (x1 + x2));
    fmax *= isLess(time, period * (x1 + x2) + x1) ? time - period * x2 : (period + 1.) * x1;
  } else if (bc.time_fn == 3) {
    if (isGreater(time, bc.tp[1]))
      fmax *= bc.tp[3] * time;
    else
      fmax *= bc.tp[2] * time;
  } else if (bc.time_fn == 4) {
    double a = M_PI * bc.tp[1];
    fmax *= sin(a * time);
  }
  fmax *= bc.tp[0];
  const int64_t i = bc.nodes[k];
  for (int d = 0; d < DIM; d++) {
    double fi = (bc.dir_mask >> d & 1) ? fmax : 0.;
    s.f[d][i] += fi;
  }
}

// -------------------------------------------------------- PD bond forces
template <int INF>
__device__ __forceinline__ double infJ(const DS& s, double r_over_delta, double rji, double horizon) {
  if (INF == 0) return s.inf_p0;
  if (INF == 1) return s.inf_p0 + s.inf_p1 * (rji / horizon);
  const double r = rji / horizon;
  return s.inf_p0 * exp(-r * r / s.inf_p1);
}

// Bond-based RNP: computePeridynamicForces force pass (demModel.cpp:682-788)
// with RnpMaterial::getBondEF (material.h:342-368).  One lane per node; the
// force sum runs over the neighbor list in order, as the reference's does.
// The next neighbor's data is loaded while the current bond is evaluated.
template <int DIM, int INF>
__global__ void __launch_bounds__(256) k_pd_rnp(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  if (!p.compute_force) {
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
    return;
  }
  if (p.material_type != 0) return;
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const double Cc = p.mat[0], beta = p.mat[1], rbar = p.mat[2], invF = p.mat[3], fsc = p.mat[4];
  const bool irrev = p.irreversible != 0;
  const double Kn = p.Kn, Rc = p.Rc;
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* __restrict__ c = s.col + s.sptr[sl] + lane;
  unsigned* bw = s.bits + s.bptr[sl] + lane;
  const int n = s.deg[i];
  double xi[3], ui[3];
  for (int d = 0; d < DIM; d++) {
    xi[d] = s.xr[d][i];
    ui[d] = s.u[d][i];
  }
  double force[3] = {0., 0., 0.};
  float Zi = 0.f;
  unsigned word = 0, wnew = 0;
  double xn[3], un[3], vn = 0.;
  if (n > 0) {
    const int j0 = c[0];
    for (int d = 0; d < DIM; d++) {
      xn[d] = s.xr[d][j0];
      un[d] = s.u[d][j0];
    }
    vn = s.vol[j0];
  }
  for (int k = 0; k < n; k++) {
    if ((k & 31) == 0) {
      if (k > 0 && wnew != word) bw[(int64_t)((k >> 5) - 1) * WAVE] = wnew;
      word = bw[(int64_t)(k >> 5) * WAVE];
      wnew = word;
    }
    double xj[3], uj[3];
    for (int d = 0; d < DIM; d++) {
      xj[d] = xn[d];
      uj[d] = un[d];
    }
    double volj = vn;
    if (k + 1 < n) {
      const int jn = c[(int64_t)(k + 1) * WAVE];
      for (int d = 0; d < DIM; d++) {
        xn[d] = s.xr[d][jn];
        un[d] = s.u[d][jn];
      }
      vn = s.vol[jn];
    }
    bool fs = (word >> (k & 31)) & 1u;
    double dx[3], du[3];
    for (int d = 0; d < DIM; d++) {
      dx[d] = xj[d] - xi[d];
      du[d] = uj[d] - ui[d];
    }
    const double rji = len3<DIM>(dx);
    const double Sji = dot3<DIM>(dx, du) / dot3<DIM>(dx, dx);
    const double Sc = rbar / sqrt(rji);
    if (!fs) {
      if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
      const double J = infJ<INF>(s, 0., rji, horizon);
      if (irrev && isGreater(fabs(Sji), fsc * Sc)) fs = true;
      double ef = 0.;
      if (!fs) ef = J * 4. * Sji * Cc * beta * exp(-beta * rji * Sji * Sji) / invF;
      if (fs) wnew |= 1u << (k & 31);
      const double scalar_f = ef * volj;
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * (dx[d] / rji);
    } else {
      double y[3];
      for (int d = 0; d < DIM; d++) y[d] = (xj[d] + uj[d]) - (xi[d] + ui[d]);
      const double R = len3<DIM>(y);
      double scalar_f = Kn * volj * (R - Rc) / R;
      if (scalar_f > 0.) scalar_f = 0.;
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * y[d];
    }
    if (isGreater(fabs(Sji / Sc), (double)Zi)) Zi = (float)fabs(Sji / Sc);
  }
  if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
  for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
  s.Z[i] = Zi;
}

// Bond-based PMB (LAW 2) and PdElastic (LAW 3) force pass
// (demModel.cpp:682-788 with PmbMaterial / PdElastic, material.h:665-1324):
// actual stretch S = (|xi + eta| - |xi|) / |xi|, force J c S along the
// deformed bond (xi + eta) / |xi + eta|; PMB bonds break once
// |S| > s0 + 1e-10 (relative compare), PdElastic bonds never break; broken
// bonds give the short-range repulsion; damage max |S / Sc| with Sc = s0
// (PMB) or the largest double (PdElastic).
template <int DIM, int INF, int LAW>
__global__ void __launch_bounds__(256) k_pd_bond(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  if (!p.compute_force) {
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
    return;
  }
  if (p.material_type != LAW) return;
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const double cc = p.mat[0], s0 = p.mat[1];
  const double Sc = LAW == 2 ? s0 : DBL_MAX;
  const double Kn = p.Kn, Rc = p.Rc;
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* __restrict__ c = s.col + s.sptr[sl] + lane;
  unsigned* bw = s.bits + s.bptr[sl] + lane;
  const int n = s.deg[i];
  double xi[3], ui[3];
  for (int d = 0; d < DIM; d++) {
    xi[d] = s.xr[d][i];
    ui[d] = s.u[d][i];
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
    double dx[3], du[3], dy[3];
    for (int d = 0; d < DIM; d++) {
      dx[d] = s.xr[d][j] - xi[d];
      du[d] = s.u[d][j] - ui[d];
      dy[d] = dx[d] + du[d];
    }
    double volj = s.vol[j];
    bool fs = (word >> (k & 31)) & 1u;
    const double rji = len3<DIM>(dx);
    const double Sji = (len3<DIM>(dy) - len3<DIM>(dx)) / len3<DIM>(dx);
    if (!fs) {
      if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
      const double J = infJ<INF>(s, 0., rji, horizon);
      if (LAW == 2 && isGreater(fabs(Sji), s0 + 1.0e-10)) fs = true;
      const double ef = fs ? 0. : J * cc * Sji;
      if (fs) wnew |= 1u << (k & 31);
      const double scalar_f = ef * volj;
      const double ly = len3<DIM>(dy);
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * (dy[d] / ly);
    } else {
      double y[3];
      for (int d = 0; d < DIM; d++) y[d] = (s.xr[d][j] + s.u[d][j]) - (xi[d] + ui[d]);
      const double R = len3<DIM>(y);
      double scalar_f = Kn * volj * (R - Rc) / R;
      if (scalar_f > 0.) scalar_f = 0.;
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * y[d];
    }
    if (isGreater(fabs(Sji / Sc), (double)Zi)) Zi = (float)fabs(Sji / Sc);
  }
  if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
  for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
  s.Z[i] = Zi;
}

// State-based dilation pass (demModel.cpp:610-679): bond breaking on the
// actual stretch, theta_i = 3 sum r (|y|-r) J V^c / m_i.
template <int DIM, int INF>
__global__ void __launch_bounds__(256) k_pd_dilation(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  if (!p.compute_force || p.material_type != 1) return;
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const double sc = p.mat[2];
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* c = s.col + s.sptr[sl] + lane;
  unsigned* bw = s.bits + s.bptr[sl] + lane;
  const int n = s.deg[i];
  double xi[3], ui[3];
  for (int d = 0; d < DIM; d++) {
    xi[d] = s.xr[d][i];
    ui[d] = s.u[d][i];
  }
  const double m = s.mx[i];
  double theta = 0.;
  unsigned word = 0, wnew = 0;
  for (int k = 0; k < n; k++) {
    if ((k & 31) == 0) {
      if (k > 0 && wnew != word) bw[(int64_t)((k >> 5) - 1) * WAVE] = wnew;
      word = bw[(int64_t)(k >> 5) * WAVE];
      wnew = word;
    }
    const int j = c[(int64_t)k * WAVE];
    double dx[3], dy[3];
    for (int d = 0; d < DIM; d++) {
      dx[d] = s.xr[d][j] - xi[d];
      dy[d] = (dx[d] + s.u[d][j]) - ui[d];
    }
    const double rji = len3<DIM>(dx);
    const double change_length = len3<DIM>(dy) - rji;
    const double st = change_length / rji;
    bool fs = (word >> (k & 31)) & 1u;
    if (!fs && isGreater(fabs(st), sc + 1.0e-10)) {
      fs = true;
      wnew |= 1u << (k & 31);
    }
    if (!fs) {
      double volj = s.vol[j];
      if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
      theta += rji * change_length * infJ<INF>(s, 0., rji, horizon) * volj;
    }
  }
  if (n > 0 && wnew != word) bw[(int64_t)((n - 1) >> 5) * WAVE] = wnew;
  s.theta[i] = 3. * theta / m;
}

// State-based force pass with PdState::getBondEF for both ends (material.h:1415-1429).
template <int DIM, int INF>
__global__ void __launch_bounds__(256) k_pd_state(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const DevParticle& p = s.P[s.ptid[i]];
  if (!p.compute_force) {
    for (int d = 0; d < DIM; d++) s.f[d][i] = 0.;
    return;
  }
  if (p.material_type != 1) return;
  const double horizon = p.horizon, h = p.h;
  const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
  const double K = p.mat[0], G = p.mat[1], s0 = p.mat[2];
  const double Kn = p.Kn, Rc = p.Rc;
  const int64_t sl = i / WAVE, lane = i % WAVE;
  const int* c = s.col + s.sptr[sl] + lane;
  const unsigned* bw = s.bits + s.bptr[sl] + lane;
  const int n = s.deg[i];
  double xi[3], ui[3];
  for (int d = 0; d < DIM; d++) {
    xi[d] = s.xr[d][i];
    ui[d] = s.u[d][i];
  }
  const double mi = s.mx[i], thetai = s.theta[i];
  const double alpha_i = 15. * G / mi;
  const double factor_i = (3. * K / mi) - alpha_i / 3.;
  double force[3] = {0., 0., 0.};
  float Zi = 0.f;
  unsigned word = 0;
  for (int k = 0; k < n; k++) {
    if ((k & 31) == 0) word = bw[(int64_t)(k >> 5) * WAVE];
    const int j = c[(int64_t)k * WAVE];
    const bool fs = (word >> (k & 31)) & 1u;
    double dx[3], du[3], dy[3];
    for (int d = 0; d < DIM; d++) {
      dx[d] = s.xr[d][j] - xi[d];
      du[d] = s.u[d][j] - ui[d];
      dy[d] = dx[d] + du[d];
    }
    double volj = s.vol[j];
    const double rji = len3<DIM>(dx);
    const double ly = len3<DIM>(dy);
    const double Sji = (ly - len3<DIM>(dx)) / len3<DIM>(dx);
    if (!fs) {
      if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
      const double J = infJ<INF>(s, 0., rji, horizon);
      const double mj = s.mx[j], thetaj = s.theta[j];
      const double cl = Sji * rji;
      const double ef_i = J * (rji * thetai * factor_i + cl * alpha_i);
      const double alpha_j = 15. * G / mj;
      const double factor_j = (3. * K / mj) - alpha_j / 3.;
      const double ef_j = J * (rji * thetaj * factor_j + cl * alpha_j);
      const double scalar_f = (ef_i + ef_j) * volj;
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * (dy[d] / ly);
    } else {
      double y[3];
      for (int d = 0; d < DIM; d++) y[d] = (s.xr[d][j] + s.u[d][j]) - (xi[d] + ui[d]);
      const double R = len3<DIM>(y);
      double scalar_f = Kn * volj * (R - Rc) / R;
      if (scalar_f > 0.) scalar_f = 0.;
      for (int d = 0; d < DIM; d++) force[d] += scalar_f * y[d];
    }
    if (isGreater(fabs(Sji / s0), (double)Zi)) Zi = (float)fabs(Sji / s0);
  }
  for (int d = 0; d < DIM; d++) s.f[d][i] = force[d];
  s.Z[i] = Zi;
}

// ---------------------------------------------------------- cell list
// Uniform grid with cell >= the largest search radius / csub (at most 2^25
// cells: CELL_MAX_BLOCKS blocks of CELL_BLOCK).  At a rebuild
// the nodes are sorted by cell key with a counting sort on the grid:
// k_cells_count (per-cell counts and each node's rank in its cell),
// k_cells_scan (exclusive scan of the counts inside blocks of CELL_BLOCK
// cells, block totals), k_cells_scatter (block offsets, then every node to
// its place).  The order inside a cell does not matter: candidate sets are
// sets and contact sums run in ascending node id.
constexpr int CELL_BLOCK = 4096;  // cells per scan block (1024 threads x 4)
constexpr int CELL_TPB = 1024;
constexpr int CELL_MAX_BLOCKS = (1 << 25) / CELL_BLOCK;

__device__ __forceinline__ int cellc(double a, double lo, double cell, int n) {
  double k = floor((a - lo) / cell);
  int ki = k < 0. ? 0 : (k > (double)(n - 1) ? n - 1 : (int)k);
  return ki;
}

template <int DIM>
__device__ __forceinline__ void count_node(const DS& s, int64_t i) {
  int cx = cellc(s.x[0][i], s.glo[0], s.cell, s.gn[0]);
  int cy = cellc(s.x[1][i], s.glo[1], s.cell, s.gn[1]);
  int cz = DIM == 3 ? cellc(s.x[2][i], s.glo[2], s.cell, s.gn[2]) : 0;
  int c = (cz * s.gn[1] + cy) * s.gn[0] + cx;
  const int pid = s.ptid[i];
  // the owner only moves -2 (empty) -> a particle id -> -1 (several): a cell
  // already read as this particle's or as mixed needs no atomic
  const int seen = s.cell_owner[c];
  if (seen != pid && seen != -1) {
    const int old = atomicCAS(&s.cell_owner[c], -2, pid);
    if (old != -2 && old != pid) s.cell_owner[c] = -1;
  }
  // ranks: one atomic per distinct cell among the wavefront's active lanes
  // (the lanes of a wavefront bin neighboring nodes, mostly into a few cells)
  const int lane = (int)(threadIdx.x & (WAVE - 1));
  unsigned long long todo = __ballot(1);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cl = __shfl(c, leader);
    const unsigned long long same = __ballot(c == cl);
    int base = 0;
    if (lane == leader) base = atomicAdd(&s.cell_count[cl], (int)__popcll(same));
    base = __shfl(base, leader);
    if (c == cl) slot = base + (int)__popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  s.node_cell[i] = c;
  s.node_slot[i] = slot;
}

// grid-stride over the nodes (the launch stays small on steps without a rebuild)
template <int DIM>
__global__ void k_cells_count(DS s) {
  if (!s.ctrl->rebuild) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gid(); i < s.N; i += stride) count_node<DIM>(s, i);
}

// inclusive scan of v over the workgroup (CELL_TPB threads); tot gets the sum
__device__ __forceinline__ int block_scan_incl(int v, int* lds, int& tot) {
  const int lane = (int)(threadIdx.x & (WAVE - 1)), w = (int)(threadIdx.x / WAVE);
  for (int o = 1; o < WAVE; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == WAVE - 1) lds[w] = v;
  __syncthreads();
  if (w == 0) {
    int x = lane < (int)(blockDim.x / WAVE) ? lds[lane] : 0;
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(x, o);
      if (lane >= o) x += t;
    }
    if (lane < (int)(blockDim.x / WAVE)) lds[lane] = x;
  }
  __syncthreads();
  tot = lds[blockDim.x / WAVE - 1];
  const int r = v + (w > 0 ? lds[w - 1] : 0);
  __syncthreads();
  return r;
}

// one workgroup per block of CELL_BLOCK cells: exclusive prefix of the counts
// inside the block and the block's total
__global__ void __launch_bounds__(CELL_TPB) k_cells_scan(DS s) {
  __shared__ int lds[CELL_TPB / WAVE];
  if (!s.ctrl->rebuild) return;
  const int64_t c0 = (int64_t)blockIdx.x * CELL_BLOCK + 4 * threadIdx.x;
  int v[4];
  for (int k = 0; k < 4; k++) v[k] = c0 + k < s.ncells ? s.cell_count[c0 + k] : 0;
  const int mine = v[0] + v[1] + v[2] + v[3];
  int tot;
  const int incl = block_scan_incl(mine, lds, tot);
  int run = incl - mine;
  for (int k = 0; k < 4; k++) {
    if (c0 + k < s.ncells) s.cell_start[c0 + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) s.cell_btot[blockIdx.x] = tot;
}

// block offsets (every workgroup scans the block totals; workgroup 0 stores
// them for the query), then each node to its place in the sorted list
template <int DIM>
__global__ void __launch_bounds__(CELL_TPB) k_cells_scatter(DS s) {
  __shared__ int lds[CELL_TPB / WAVE];
  __shared__ int boff[CELL_MAX_BLOCKS];
  if (!s.ctrl->rebuild) return;
  const int nb = (int)s.ncell_blocks;
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += CELL_TPB) {
    const int b = b0 + (int)threadIdx.x;
    const int v = b < nb ? s.cell_btot[b] : 0;
    int tot;
    const int incl = block_scan_incl(v, lds, tot);
    if (b < nb) {
      boff[b] = carry + incl - v;
      if (blockIdx.x == 0) s.cell_boff[b] = carry + incl - v;
    }
    carry += tot;
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gid(); i < s.N; i += stride) {
    const int c = s.node_cell[i];
    const int64_t e = (int64_t)boff[c / CELL_BLOCK] + s.cell_start[c] + s.node_slot[i];
    s.cell_nodes[e] = (int)i;
    s.cell_xf[e] = make_float4((float)(s.x[0][i] - s.glo[0]), (float)(s.x[1][i] - s.glo[1]),
                               DIM == 3 ? (float)(s.x[2][i] - s.glo[2]) : 0.f, __int_as_float(s.ptid[i]));
  }
}

// radiusSearchExcludeTag for one searching node (demModel.cpp:1717-1756):
// other particles, j != i, 0 + dx^2 + dy^2 + dz^2 < r^2.
// Cell entries are first tested on their FP32 copy (one 16-byte load: the
// position relative to the grid origin and the particle id): an entry is
// dropped when its FP32 distance is at least Rf = r + 2^-19 (max|a| + r),
// a = x_i - glo.  Rounding a and b = x_j - glo to FP32 moves each component
// by at most 2^-24 |a| and 2^-24 (|a| + r) for |x_j - x_i| < r, and the FP32
// difference and sum add a relative 2^-22 at most, so an entry within r of
// node i always passes, with a factor of four to spare.  Entries that pass
// take the FP64 test below, which alone decides the candidate set.
template <int DIM>
__device__ int query_node(const DS& s, int64_t i, int pi) {
  const double r = s.ctrl->radius, r2 = r * r;
  double xi[3] = {0., 0., 0.};
  for (int d = 0; d < DIM; d++) xi[d] = s.x[d][i];
  float af[3] = {0.f, 0.f, 0.f};
  float amax = 0.f;
  for (int d = 0; d < DIM; d++) {
    af[d] = (float)(xi[d] - s.glo[d]);
    amax = fmaxf(amax, fabsf(af[d]));
  }
  const float rf = (float)r;
  const float Rf = rf + 0x1p-19f * (amax + rf);
  const float R2f = Rf * Rf;
  int cc0[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < DIM; d++) {
    cc0[d] = cellc(xi[d], s.glo[d], s.cell, s.gn[d]);
    lo[d] = cc0[d] - s.csub > 0 ? cc0[d] - s.csub : 0;
    hi[d] = cc0[d] + s.csub < s.gn[d] - 1 ? cc0[d] + s.csub : s.gn[d] - 1;
  }
  int cnt = 0;
  double near2 = r2;  // nearest candidate (squared distance)
  const int cap = s.cap;
  for (int z = lo[2]; z <= hi[2]; z++)
    for (int y = lo[1]; y <= hi[1]; y++)
      for (int x = lo[0]; x <= hi[0]; x++) {
        // skip cells whose box lies beyond the search radius (clamped edge
        // cells hold nodes outside the grid: never skipped)
        const int c3[3] = {x, y, z};
        double dmin2 = 0.;
        bool edge = false;
        for (int d = 0; d < DIM; d++) {
          if (c3[d] == 0 || c3[d] == s.gn[d] - 1) edge = true;
          const double b0 = s.glo[d] + c3[d] * s.cell, b1 = b0 + s.cell;
          const double t = xi[d] < b0 ? b0 - xi[d] : (xi[d] > b1 ? xi[d] - b1 : 0.);
          dmin2 += t * t;
        }
        if (!edge && dmin2 > r2 * 1.0000001) continue;
        const int cc = (z * s.gn[1] + y) * s.gn[0] + x;
        if (s.cell_owner[cc] == pi) continue;  // only the node's own particle
        const int m = s.cell_count[cc];
        const int64_t e0 = (int64_t)s.cell_boff[cc / CELL_BLOCK] + s.cell_start[cc];
        const int* cn = s.cell_nodes + e0;
        const float4* cx = s.cell_xf + e0;
        for (int q = 0; q < m; q++) {
          const float4 e = cx[q];
          if (__float_as_int(e.w) == pi) continue;  // the node's own particle (and i itself)
          const float t0 = af[0] - e.x, t1 = af[1] - e.y, t2 = DIM == 3 ? af[2] - e.z : 0.f;
          const float d2f = t0 * t0 + t1 * t1 + t2 * t2;
          if (!(d2f < R2f)) continue;
          const int j = cn[q];
          double d2 = 0.;
          for (int d = 0; d < DIM; d++) {
            double t = xi[d] - s.x[d][j];
            d2 += t * t;
          }
          if (d2 < r2) {
            if (cnt < cap) s.cand[cand_idx(s.N, i, cnt)] = j;
            cnt++;
            near2 = d2 < near2 ? d2 : near2;
          }
        }
      }
  if (cnt > cap) {
    s.ctrl->overflow = 1;
    cnt = cap;
  }
  s.ccount[i] = cnt;
  if (cnt > 0) s.cdmin[i] = __double2float_rd(sqrt(near2));
  return cnt;
}

// contact skin test: node i's candidates of the last rebuild were all at
// least cdmin[i] away then, and every node has moved at most Ctrl::dsum
// since, so none is within contact distance when cdmin[i] - 2 dsum exceeds
// the largest contact radius (with a margin): the contact loops would find
// no active pair for the node and skip it
__device__ __forceinline__ bool skin_far(const DS& s, int64_t i, double two_dsum) {
  return (double)s.cdmin[i] - two_dsum > s.rc_skip;
}

// one active contact pair's two force terms on node i (demModel.cpp:868-904):
// the normal term sf en and the friction term (mu sf) et; the reference adds
// them to the node's force one after the other
template <int DIM>
__device__ __forceinline__ void contact_pair(const DS& s, const double* vi, int j,
                                             const DevContact& ct, double Rji, const double* yji,
                                             double* fn, double* ft) {
  const double volj = s.vol[j];
  double vji[3], en[3], et[3];
  for (int d = 0; d < DIM; d++) {
    vji[d] = s.v[d][j] - vi[d];
    en[d] = yji[d] / Rji;
  }
  const double vn_mag = dot3<DIM>(vji, en);
  for (int d = 0; d < DIM; d++) et[d] = vji[d] - vn_mag * en[d];
  const double etl = len3<DIM>(et);
  if (isGreater(etl, 0.)) {
    for (int d = 0; d < DIM; d++) et[d] = et[d] / etl;
  } else {
    for (int d = 0; d < DIM; d++) et[d] = 0.;
  }
  double scalar_f = ct.Kn * (Rji - ct.contact_r) * volj;
  if (scalar_f > 0.) scalar_f = 0.;
  const double mf = ct.mu * scalar_f;
  for (int d = 0; d < DIM; d++) {
    fn[d] = scalar_f * en[d];
    ft[d] = mf * et[d];
  }
}

// Closure gating seen from the node-level contact loops: a node is visited
// only if its particle has a plain pair or a wall candidate, and a candidate
// in another non-wall particle only if that pair is plain this step.
__device__ __forceinline__ bool node_needs_contact(const DS& s, int nwi) {
  return !s.cl_on || nwi < 0 || (s.pflags[nwi] & 1) != 0 || s.wall_near[nwi] == (int)s.ctrl->n_rebuilds;
}
__device__ __forceinline__ bool pair_plain(const DS& s, int a, int b) {
  const int* dl = s.dlist + (int64_t)a * s.dcap;
  int lo = 0, hi = s.dcount[a] - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = dl[mid];
    if (v == b) return s.cmode[(int64_t)a * s.dcap + mid] == 2;
    if (v < b)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return true;  // not a listed pair: keep the node-level path
}

// ------------------------------------------------------- contact forces
// Candidate query on a rebuild: one lane per searching node; nodes that end
// with candidates (and compute forces) join the light or the active list
// (the lists are consumed per node, so their order is free).

// the candidate query of node i; returns the contact list the node joins:
// 1 light (lane per node), 2 active (wavefront per node), 0 none
template <int DIM>
__device__ __forceinline__ int query_one(const DS& s, int64_t i) {
  const int pid = s.ptid[i];
  const DevParticle& pi = s.P[pid];
  if (!pi.searches) return 0;
  const uint8_t ep = (uint8_t)s.ctrl->n_rebuilds;
  const int cnt = query_node<DIM>(s, i, pid);
  if (s.cl_on && pi.nonwall_index >= 0)
    for (int k = 0; k < cnt; k++)
      if (s.pmeta[s.ptid[s.cand[cand_idx(s.N, i, k)]]] >> 16) {
        s.wall_near[pi.nonwall_index] = (int)s.ctrl->n_rebuilds;
        break;
      }
  // node i is a candidate of each of its searching candidates (the FP64 test
  // is symmetric): its record is kept current until the next rebuild.  The
  // byte store comes last: a store through uint8_t may alias any earlier
  // load, which would otherwise have to be repeated after it.
  if (cnt > 0) s.xa_epoch[i] = ep;
  const bool want = cnt > 0 && pi.compute_force;
  return want ? (cnt <= s.light_max ? 1 : 2) : 0;
}

// candidate j seen from node position yi: yji = x_j - yi, and j's particle id
// when the position came from j's record (between rebuilds), else -1
template <int DIM>
__device__ __forceinline__ void cand_pos(const DS& s, bool rec, int j, const double* yi, double* yji, int& pj) {
  if (rec) {
    const double4 a = s.xa[j];
    yji[0] = a.x - yi[0];
    yji[1] = a.y - yi[1];
    if (DIM == 3) yji[2] = a.z - yi[2];
    pj = (int)__double_as_longlong(a.w);
  } else {
    for (int d = 0; d < DIM; d++) yji[d] = s.x[d][j] - yi[d];
    pj = -1;
  }
}

// computeContactForces node loop for nodes with few candidates: one lane per
// node.  The active pairs' ids (all of them, kept sorted in registers) fix the
// summation order: ascending node id, the order of the oracle's lists.
constexpr int NACT = LIGHT_MAX_C;  // every candidate of a light node fits

template <int DIM>
__global__ void __launch_bounds__(256) k_contact_light(DS s) {
  const int nl = (int)s.ctrl->nlight;
  const bool rec = !s.ctrl->rebuild;
  const double two_dsum = 2. * s.ctrl->dsum;
  for (int a = (int)gid(); a < nl; a += gridDim.x * blockDim.x) {
    const int i = s.light[a];
    if (skin_far(s, i, two_dsum)) continue;
    const int n = s.ccount[i];
    const int pid = s.ptid[i];
    const int mi = s.pmeta[pid];
    const int zi = mi & 0xffff;
    const bool wi = (mi >> 16) != 0;
    const int nwi = s.P[pid].nonwall_index;
    if (!node_needs_contact(s, nwi)) continue;
    double yi[3], vi[3];
    for (int d = 0; d < DIM; d++) {
      yi[d] = s.x[d][i];
      vi[d] = s.v[d][i];
    }
    int act[NACT];
#pragma unroll
    for (int q = 0; q < NACT; q++) act[q] = INT_MAX;
    int nact = 0;
    // distance prefilter over all candidates before any particle data is
    // loaded: independent loads, several candidates in flight
    unsigned pass = 0;
#pragma unroll 4
    for (int k = 0; k < n; k++) {
      const int j = s.cand[cand_idx(s.N, i, k)];
      double yji[3];
      int pj;
      cand_pos<DIM>(s, rec, j, yi, yji, pj);
      if (dot3<DIM>(yji, yji) < s.rc2_max) pass |= 1u << k;
    }
    while (pass) {
      const int k = __builtin_ctz(pass);
      pass &= pass - 1u;
      const int j = s.cand[cand_idx(s.N, i, k)];
      double yji[3];
      int pjid;
      cand_pos<DIM>(s, rec, j, yi, yji, pjid);
      const double d2 = dot3<DIM>(yji, yji);
      if (pjid < 0) pjid = s.ptid[j];
      const int mj = s.pmeta[pjid];
      const bool wj = (mj >> 16) != 0;
      if (j == i || pjid == pid || (wi && wj)) continue;
      const double rc = s.CT[zi * s.nz + (mj & 0xffff)].contact_r;
      // d^2 >= rc^2 implies R >= rc (1 - 2^-53): isLess(R, rc) is false
      if (d2 >= rc * rc) continue;
      if (!isLess(len3<DIM>(yji), rc)) continue;
      if (s.cl_on && !wj && !pair_plain(s, nwi, s.P[pjid].nonwall_index)) continue;
      if (wj && nwi >= 0) {
        const int w = s.widx[j];
        atomicOr(&s.wbits[(int64_t)nwi * s.nwords + (w >> 5)], 1u << (w & 31));
        atomicMin(&s.wlo[nwi], w >> 5);
        atomicMax(&s.whi[nwi], w >> 5);
      }
      nact++;
      int v = j;
#pragma unroll
      for (int q = 0; q < NACT; q++) {
        const int lo = v < act[q] ? v : act[q];
        const int hi = v < act[q] ? act[q] : v;
        act[q] = lo;
        v = hi;
      }
    }
    if (nact == 0) continue;
    double force[3] = {0., 0., 0.};
    for (int r = 0; r < nact; r++) {
      int j = act[0];
#pragma unroll
      for (int q = 0; q < NACT; q++)
        if (q == r) j = act[q];
      double yji[3], fn[3], ft[3];
      int pj;
      cand_pos<DIM>(s, rec, j, yi, yji, pj);
      const int mj = s.pmeta[pj >= 0 ? pj : s.ptid[j]];
      const DevContact ct = s.CT[zi * s.nz + (mj & 0xffff)];
      contact_pair<DIM>(s, vi, j, ct, len3<DIM>(yji), yji, fn, ft);
      for (int d = 0; d < DIM; d++) force[d] += fn[d];
      for (int d = 0; d < DIM; d++) force[d] += ft[d];
    }
    for (int d = 0; d < DIM; d++) s.f[d][i] += force[d];
  }
}

// computeContactForces node loop (demModel.cpp:832-934), one wavefront per
// node of the active list: lanes test candidates in parallel, the active
// pairs' terms go to LDS, and one lane adds them in ascending node order (the
// order of the oracle's sorted candidate lists).  Wall nodes within contact
// distance are recorded for the damping pass (demModel.cpp:998-1022).
constexpr int CWAVES = 4;      // wavefronts per workgroup
constexpr int MAXACT = 64;     // active pairs per node held in LDS

// (at most 128 VGPRs: 4 wavefronts per SIMD; the kernel waits on chains of
// dependent gathers, and 3 wavefronts per SIMD took 28 % longer)
template <int DIM>
__global__ void __launch_bounds__(CWAVES * WAVE, 4) k_contact_force(DS s) {
  __shared__ int sj[CWAVES][MAXACT];
  __shared__ double st[CWAVES][MAXACT][6];
  // this rebuild's query is done: leave the cell table empty for the next
  if (s.ctrl->rebuild) clear_cells(s, gid(), (int64_t)gridDim.x * blockDim.x);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int nwaves = gridDim.x * CWAVES;
  const int nact_nodes = (int)s.ctrl->nactive;
  const bool rec = !s.ctrl->rebuild;
  const double two_dsum = 2. * s.ctrl->dsum;
  // the skin test for the wavefront's next 64 nodes at once (a lane each),
  // then the nodes that pass it one after the other, in the same order
  for (int a0 = blockIdx.x * CWAVES + wv; a0 < nact_nodes; a0 += WAVE * nwaves) {
    const int al = a0 + lane * nwaves;
    int il = 0;
    bool near = false;
    if (al < nact_nodes) {
      il = s.active[al];
      near = !skin_far(s, il, two_dsum);
    }
    unsigned long long todo = __ballot(near);
    while (todo) {
      const int l = __builtin_ctzll(todo);
      todo &= todo - 1ull;
      const int i = __shfl(il, l);
      const int n = s.ccount[i];
      const int pid = s.ptid[i];
      const int mi = s.pmeta[pid];
      const int zi = mi & 0xffff;
      const bool wi = (mi >> 16) != 0;
      const int nwi = s.P[pid].nonwall_index;
      if (!node_needs_contact(s, nwi)) continue;
      double yi[3], vi[3];
      for (int d = 0; d < DIM; d++) {
        yi[d] = s.x[d][i];
        vi[d] = s.v[d][i];
      }
      int cnt = 0;
      for (int base = 0; base < n; base += WAVE) {
        const int k = base + lane;
        bool on = false;
        int j = 0;
        double term[6] = {0., 0., 0., 0., 0., 0.};
        if (k < n) {
          j = s.cand[cand_idx(s.N, i, k)];
          // distance prefilter (largest contact radius) before any particle data is loaded
          double yji[3];
          int pj;
          cand_pos<DIM>(s, rec, j, yi, yji, pj);
          const double d2 = dot3<DIM>(yji, yji);
          const int pjid = d2 < s.rc2_max ? (pj >= 0 ? pj : s.ptid[j]) : pid;
          const int mj = s.pmeta[pjid];
          const bool wj = (mj >> 16) != 0;
          if (!(j == i || pjid == pid || (wi && wj))) {
            const DevContact ct = s.CT[zi * s.nz + (mj & 0xffff)];
            // d^2 >= rc^2 implies R >= rc (1 - 2^-53): isLess(R, rc) is false
            if (d2 < ct.contact_r * ct.contact_r) {
              const double Rji = len3<DIM>(yji);
              if (isLess(Rji, ct.contact_r) &&
                  !(s.cl_on && !wj && !pair_plain(s, nwi, s.P[pjid].nonwall_index))) {
                on = true;
                if (wj && nwi >= 0) {
                  const int w = s.widx[j];
                  atomicOr(&s.wbits[(int64_t)nwi * s.nwords + (w >> 5)], 1u << (w & 31));
                  atomicMin(&s.wlo[nwi], w >> 5);
                  atomicMax(&s.whi[nwi], w >> 5);
                }
                contact_pair<DIM>(s, vi, j, ct, Rji, yji, term, term + 3);
              }
            }
          }
        }
        const unsigned long long m = __ballot(on);
        if (on) {
          const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
          if (pos < MAXACT) {
            sj[wv][pos] = j;
            for (int d = 0; d < 6; d++) st[wv][pos][d] = term[d];
          }
        }
        cnt += __popcll(m);
      }
      // LDS is private to this wavefront's slot and LDS instructions of one
      // wavefront complete in order: no workgroup barrier is needed (the waves
      // of a workgroup run different trip counts of this loop)
      __builtin_amdgcn_wave_barrier();
      if (cnt > MAXACT) {
        if (lane == 0) s.ctrl->overflow = 3;
        cnt = MAXACT;
      }
      if (cnt > 1) {
        // ascending node order: each lane moves one active pair to its rank
        // (a node's candidates are distinct nodes); the wavefront issues all
        // its reads of the slot before the writes
        int kj = INT_MAX;
        double kt[6] = {0., 0., 0., 0., 0., 0.};
        if (lane < cnt) {
          kj = sj[wv][lane];
          for (int d = 0; d < 6; d++) kt[d] = st[wv][lane][d];
        }
        int rank = 0;
        for (int q = 0; q < cnt; q++) rank += sj[wv][q] < kj ? 1 : 0;
        __builtin_amdgcn_wave_barrier();
        if (lane < cnt) {
          sj[wv][rank] = kj;
          for (int d = 0; d < 6; d++) st[wv][rank][d] = kt[d];
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (lane == 0 && cnt > 0) {
        double force[3] = {0., 0., 0.};
        for (int q = 0; q < cnt; q++) {
          for (int d = 0; d < DIM; d++) force[d] += st[wv][q][d];
          for (int d = 0; d < DIM; d++) force[d] += st[wv][q][3 + d];
        }
        for (int d = 0; d < DIM; d++) s.f[d][i] += force[d];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// Damping partners on a rebuild: one wavefront per non-wall particle, lanes
// over the other particles, ballot-ordered append (ascending index).
template <int DIM>
__device__ void damp_pairs_one(const DS& s, int q, int lane) {
  const int nn = s.n_nonwall;
  const double Ri = s.prad[q];
  const double margin = 2.0 * s.ctrl->radius;
  const int zi = s.pzone[q];
  double xc[3];
  for (int d = 0; d < DIM; d++) xc[d] = s.pcx[d * nn + q];
  int cnt = 0;
  for (int base = 0; base < nn; base += WAVE) {
    const int qq = base + lane;
    bool on = false;
    if (qq < nn && qq != q) {
      double d2 = 0.;
      for (int d = 0; d < DIM; d++) {
        const double t = s.pcx[d * nn + qq] - xc[d];
        d2 += t * t;
      }
      const double reach = Ri + s.prad[qq] + 1.01 * s.CT[zi * s.nz + s.pzone[qq]].contact_r + margin;
      on = d2 < reach * reach;
    }
    const unsigned long long m = __ballot(on);
    if (on) {
      const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < s.dcap) {
        s.dlist[(int64_t)q * s.dcap + pos] = qq;
        // a pair the closure kernel does not visit this interval has no contact
        if (s.cl_on) s.cmode[(int64_t)q * s.dcap + pos] = 0;
        if (s.cl_on && qq > q) {
          const unsigned k = atomicAdd(&s.ctrl->npairs, 1u);
          if ((int)k < s.cpair_cap) {
            s.cpairs[2 * (int64_t)k] = q;
            s.cpairs[2 * (int64_t)k + 1] = pos;
          } else {
            s.ctrl->overflow = 5;
          }
        }
      }
    }
    cnt += __popcll(m);
  }
  if (lane == 0) {
    if (cnt > s.dcap) {
      s.ctrl->overflow = 4;
      cnt = s.dcap;
    }
    s.dcount[q] = cnt;
  }
}

// Rebuild steps only: the candidate query (radiusSearchExcludeTag per
// searching node, demModel.cpp:1717-1756) on workgroups [0, qblocks) and the
// damping partner lists (a wavefront per particle) on the others; the two are
// independent.  The cell table is emptied afterwards by the contact-force
// kernel (or k_cells_init outside the step).
template <int DIM>
__global__ void __launch_bounds__(256) k_rebuild_lists(DS s, int qblocks) {
  if (!s.ctrl->rebuild) return;
  if ((int)blockIdx.x < qblocks) {
    // the workgroup's nodes join the light and active lists with one atomic
    // per list and workgroup (ballot counts per wavefront, offsets in LDS)
    __shared__ unsigned wcnt[2][256 / WAVE];
    __shared__ unsigned wbase[2];
    const int lane = (int)(threadIdx.x & (WAVE - 1)), w = (int)(threadIdx.x / WAVE);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < s.N; b0 += (int64_t)qblocks * blockDim.x) {
      const int64_t i = b0 + threadIdx.x;
      const int list = i < s.N ? query_one<DIM>(s, i) : 0;
      const unsigned long long ml = __ballot(list == 1), ma = __ballot(list == 2);
      if (lane == 0) {
        wcnt[0][w] = (unsigned)__popcll(ml);
        wcnt[1][w] = (unsigned)__popcll(ma);
      }
      __syncthreads();
      if (threadIdx.x < 2) {
        unsigned t = 0;
        for (int v = 0; v < (int)(blockDim.x / WAVE); v++) t += wcnt[threadIdx.x][v];
        wbase[threadIdx.x] = t ? atomicAdd(threadIdx.x == 0 ? &s.ctrl->nlight : &s.ctrl->nactive, t) : 0u;
      }
      __syncthreads();
      unsigned ol = wbase[0] + (unsigned)__popcll(ml & below), oa = wbase[1] + (unsigned)__popcll(ma & below);
      for (int v = 0; v < w; v++) {
        ol += wcnt[0][v];
        oa += wcnt[1][v];
      }
      if (list == 1) s.light[ol] = (int)i;
      if (list == 2) s.active[oa] = (int)i;
      __syncthreads();  // wcnt and wbase are rewritten by the next round
    }
    return;
  }
  const int lane = threadIdx.x & (WAVE - 1);
  const int nw = ((int)gridDim.x - qblocks) * (int)(blockDim.x / WAVE);
  for (int q = ((int)blockIdx.x - qblocks) * (int)(blockDim.x / WAVE) + (int)(threadIdx.x / WAVE);
       q < s.n_nonwall; q += nw)
    damp_pairs_one<DIM>(s, q, lane);
}

// Damping (demModel.cpp:937-1052): one wavefront per non-wall particle.
// Particle-particle terms are evaluated by all lanes and accumulated in
// ascending particle order; condensed wall nodes in ascending node order.
// With `fin`, the wavefront then finalises its particle's nodes as
// k_finalize does (damping force, closure forces, gravity), and the blocks
// after the last particle add gravity to the wall nodes, so no separate pass
// over the nodes is needed.
__device__ __forceinline__ void closure_node_force(const DS& s, int a, int64_t i, double* fi);

// the damping sums of non-wall particle q, one wavefront (lanes 0..63);
// lane 0 stores the result in fdamp and in out[0..DIM)
template <int DIM>
__device__ __forceinline__ void damping_wave(const DS& s, int q, double* out) {
  const int nn = s.n_nonwall;
  const int lane = threadIdx.x;
  const int pid = s.nonwall_pid[q];
  const double Ri = s.prad[q];
  const double vol_pi = M_PI * Ri * Ri;
  const double rhoi = s.prho[q];
  const int zi = s.pzone[q];
  double xc[3], vc[3];
  for (int d = 0; d < DIM; d++) {
    xc[d] = s.pcx[d * nn + q];
    vc[d] = s.pcv[d * nn + q];
  }
  double force[3] = {0., 0., 0.};
  const int np_ = s.dcount[q];
  const int* dl = s.dlist + (int64_t)q * s.dcap;
  for (int base = 0; base < np_; base += WAVE) {
    const int qq = base + lane < np_ ? dl[base + lane] : -1;
    double term[3] = {0., 0., 0.};
    bool on = false;
    if (qq >= 0) {
      const double Rj = s.prad[qq];
      double xcj[3];
      for (int d = 0; d < DIM; d++) xcj[d] = s.pcx[d * nn + qq] - xc[d];
      const double dist = len3<DIM>(xcj);
      const DevContact ct = s.CT[zi * s.nz + s.pzone[qq]];
      if (isLess(dist, Rj + Ri + 1.01 * ct.contact_r)) {
        on = true;
        const double vol_pj = M_PI * Rj * Rj;
        const double m1 = rhoi * vol_pi, m2 = s.prho[qq] * vol_pj;
        const double meq = 2. * m1 * m2 / (m1 + m2);
        const double beta_n = ct.betan * sqrt(ct.kappa * ct.contact_r * meq);
        double hat[3] = {0., 0., 0.};
        if (isGreater(dist, 0.))
          for (int d = 0; d < DIM; d++) hat[d] = xcj[d] / dist;
        double vcj[3];
        for (int d = 0; d < DIM; d++) vcj[d] = s.pcv[d * nn + qq] - vc[d];
        double vc_mag = dot3<DIM>(vcj, hat);
        if (vc_mag > 0.) vc_mag = 0.;
        for (int d = 0; d < DIM; d++) term[d] = ((beta_n * vc_mag) * hat[d]) / vol_pi;
      }
    }
    unsigned long long mask = __ballot(on);
    while (mask) {
      const int b = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      for (int d = 0; d < DIM; d++) force[d] += __shfl(term[d], b);
    }
  }
  // wall nodes within contact distance, ascending; bits are cleared as read
  if (lane == 0) {
    const int lo = s.wlo[q], hi = s.whi[q];
    for (int w = lo; w <= hi; w++) {
      unsigned word = s.wbits[(int64_t)q * s.nwords + w];
      if (!word) continue;
      s.wbits[(int64_t)q * s.nwords + w] = 0u;
      while (word) {
        const int b = __ffs(word) - 1;
        word &= word - 1;
        const int j = s.wall_gid[w * 32 + b];
        const double meq = rhoi * vol_pi;
        const DevContact ct = s.CT[zi * s.nz + (s.pmeta[s.ptid[j]] & 0xffff)];
        const double beta_n = ct.betan * sqrt(ct.kappa * ct.contact_r * meq);
        double xcj[3], hat[3] = {0., 0., 0.}, vcj[3];
        for (int d = 0; d < DIM; d++) xcj[d] = s.x[d][j] - xc[d];
        const double l = len3<DIM>(xcj);
        if (isGreater(l, 0.))
          for (int d = 0; d < DIM; d++) hat[d] = xcj[d] / l;
        for (int d = 0; d < DIM; d++) vcj[d] = s.v[d][j] - vc[d];
        double vc_mag = dot3<DIM>(vcj, hat);
        if (vc_mag > 0.) vc_mag = 0.;
        for (int d = 0; d < DIM; d++) force[d] += ((beta_n * vc_mag) * hat[d]) / vol_pi;
      }
    }
    s.wlo[q] = 0x7fffffff;
    s.whi[q] = -1;
    for (int d = 0; d < 3; d++) s.fdamp[3 * pid + d] = d < DIM ? force[d] : 0.;
    for (int d = 0; d < DIM; d++) out[d] = force[d];
  }
}

template <int DIM>
__global__ void __launch_bounds__(256) k_damping(DS s, int fin) {
  const int q = blockIdx.x;
  const int nn = s.n_nonwall;
  // fin 2: the node update of the next step finalises the forces
  if (fin == 2 && q == 0 && threadIdx.x == 0) s.ctrl->f_pending = 1;
  if (q >= nn) {
    // finalisation of wall nodes: gravity only (k_finalize)
    if (fin == 1 && s.has_gravity)
      for (int64_t k = (int64_t)(q - nn) * blockDim.x + threadIdx.x; k < s.n_wall_nodes;
           k += (int64_t)(gridDim.x - nn) * blockDim.x) {
        const int64_t i = s.wall_nodes[k];
        const double rho = s.P[s.ptid[i]].rho;
        for (int d = 0; d < DIM; d++) s.f[d][i] += rho * s.gravity[d];
      }
    return;
  }
  __shared__ double fsh[3];
  if (threadIdx.x < WAVE) damping_wave<DIM>(s, q, fsh);
  if (fin != 1) return;
  __syncthreads();
  const DevParticle& p = s.P[s.nonwall_pid[q]];
  const bool cl = DIM == 2 && s.cl_on && (s.pflags[q] & 2);
  for (int64_t i = p.glob_start + threadIdx.x; i < p.glob_end; i += blockDim.x) {
    for (int d = 0; d < DIM; d++) s.f[d][i] += fsh[d];
    if (cl && s.surf[i]) {
      double fi[2] = {0., 0.};
      closure_node_force(s, q, i, fi);
      s.f[0][i] += fi[0];
      s.f[1][i] += fi[1];
    }
    if (s.has_gravity)
      for (int d = 0; d < DIM; d++) s.f[d][i] += p.rho * s.gravity[d];
  }
}

// closure forces on a surface node (2D; called from the force finalisation):
// for every closure pair of its particle whose patch holds the node, the
// node's overlap weight w_i = sum_j (1 - R_ij / Rc) V_j over the partner's
// patch nodes closer than Rc times (g + k perp(r_i - xbar)).
__device__ __forceinline__ void closure_node_force(const DS& s, int a, int64_t i, double* fi) {
  const int nn = s.n_nonwall;
  const double xi0 = s.x[0][i], xi1 = s.x[1][i];
  const double r0 = xi0 - s.pcx[a], r1 = xi1 - s.pcx[nn + a];
  const int cnt = s.dcount[a];
  for (int k = 0; k < cnt; k++) {
    const int64_t slot = (int64_t)a * s.dcap + k;
    if (s.cmode[slot] != 1) continue;
    const double* rec = s.cout + slot * CL_REC;
    if (r0 * rec[0] + r1 * rec[1] < rec[2]) continue;
    const int b = s.dlist[slot];
    const double cb0 = s.pcx[b], cb1 = s.pcx[nn + b], rc = rec[9];
    double w = 0.;
    for (int m = s.surf_ptr[b]; m < s.surf_ptr[b + 1]; m++) {
      const int j = s.surf_idx[m];
      const double xj0 = s.x[0][j], xj1 = s.x[1][j];
      if ((cb0 - xj0) * rec[0] + (cb1 - xj1) * rec[1] < rec[3]) continue;
      const double dx = xj0 - xi0, dy = xj1 - xi1;
      const double R = sqrt(dx * dx + dy * dy);
      if (R < rc) w += (1. - R / rc) * s.vol[j];
    }
    if (w == 0.) continue;
    const double q0 = r0 - rec[7], q1 = r1 - rec[8];
    fi[0] += w * (rec[4] - rec[6] * q1);
    fi[1] += w * (rec[5] + rec[6] * q0);
  }
}

// f += damping force of the node's particle (non-wall), then the learned
// closure's pair forces on surface nodes (closure mode only), then f += rho g
// (computeExternalForces, demModel.cpp:791-812).
template <int DIM>
__global__ void k_finalize(DS s, int with_damping, int pending_only) {
  int64_t i = gid();
  if (i >= s.N || (pending_only && !s.ctrl->f_pending)) return;
  const int pid = s.ptid[i];
  const DevParticle& p = s.P[pid];
  if (with_damping && p.type_index == 0)
    for (int d = 0; d < DIM; d++) s.f[d][i] += s.fdamp[3 * pid + d];
  if (DIM == 2 && s.cl_on && p.type_index == 0 && s.surf[i] && (s.pflags[p.nonwall_index] & 2)) {
    double fi[2] = {0., 0.};
    closure_node_force(s, p.nonwall_index, i, fi);
    s.f[0][i] += fi[0];
    s.f[1][i] += fi[1];
  }
  if (s.has_gravity)
    for (int d = 0; d < DIM; d++) s.f[d][i] += p.rho * s.gravity[d];
}

// f holds its final forces (after k_finalize over pending forces, or forces
// set without the damping and gravity terms pending)
__global__ void k_clear_pending(DS s) {
  if (gid() == 0) s.ctrl->f_pending = 0;
}

template <int DIM>
__global__ void k_apply_damping(DS s) {
  int64_t i = gid();
  if (i >= s.N) return;
  const int pid = s.ptid[i];
  if (s.P[pid].type_index == 0)
    for (int d = 0; d < DIM; d++) s.f[d][i] += s.fdamp[3 * pid + d];
}

// -------------------------------------------------------- observables
template <int DIM>
__global__ void k_pp_compressive(DS s, int wall, int fdir) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const DevParticle& w = s.P[wall];
  const int64_t i0 = w.glob_start;
  s.ctrl->pp_pen = s.x[fdir][i0] - s.xr[fdir][i0];
  double r = 0.;
  for (int64_t i = w.glob_start; i < w.glob_end; i++) r += s.f[fdir][i] * s.vol[i];
  s.ctrl->pp_reac = r;
}

// DEMModel::ppTwoParticleTest (demModel.cpp:2096-2149), one thread: centre
// distance against the first particle's bounding radius (used for both, as
// the reference does) plus the pair's contact radius, the contact-area radius,
// and the largest y of the second particle's nodes (relative compare, node
// order).  out: pen_dist, contact_area_radius, max_y, max_dist, max_y_loc.
__global__ void k_pp_two_particle(DS s, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const DevParticle& p0 = s.P[0];
  const DevParticle& p1 = s.P[1];
  double xc0[3] = {0., 0., 0.}, xc1[3] = {0., 0., 0.};
  for (int d = 0; d < 3; d++) {
    xc0[d] = s.x[d] ? s.x[d][p0.center] : 0.;
    xc1[d] = s.x[d] ? s.x[d][p1.center] : 0.;
  }
  const double r = p0.Rb;
  const double r_e = r + s.CT[p0.zone * s.nz + p1.zone].contact_r;
  const double dist = sqrt((xc1[0] - xc0[0]) * (xc1[0] - xc0[0]) + (xc1[1] - xc0[1]) * (xc1[1] - xc0[1]) +
                           (xc1[2] - xc0[2]) * (xc1[2] - xc0[2]));
  double pen = dist - r_e - r;
  double area_r = 0.;
  if (isLess(pen, 0.))
    area_r = sqrt(pow(r_e, 2.) - pow(r_e + pen, 2.));
  else if (isGreater(pen, 0.)) {
    pen = 0.;
    area_r = 0.;
  }
  const double max_dist = xc1[1] + p1.Rb;
  double max_y_loc = s.x[1][p1.glob_start];
  for (int64_t i = p1.glob_start; i < p1.glob_end; i++)
    if (isLess(max_y_loc, s.x[1][i])) max_y_loc = s.x[1][i];
  double max_y = 0.;
  if (isLess(max_y, max_y_loc)) max_y = max_y_loc;
  out[0] = pen;
  out[1] = area_r;
  out[2] = max_y;
  out[3] = max_dist;
  out[4] = max_y_loc;
}

// DEMModel::checkStop (demModel.cpp:2151-2200) quantities.
// max_node_dist: |x| of every node (util::Point::length), its maximum
// (util::methods::maxLengthAndMaxLengthIndex: the first node holding it).
__device__ __forceinline__ double node_length(const DS& s, int64_t i) {
  const double x = s.x[0][i], y = s.x[1] ? s.x[1][i] : 0., z = s.x[2] ? s.x[2][i] : 0.;
  return sqrt(x * x + y * y + z * z);
}
__global__ void k_stop_maxlen(DS s, unsigned long long* out) {
  __shared__ unsigned long long red[256 / WAVE];
  const int64_t i = gid();
  unsigned long long b = i < s.N ? (unsigned long long)__double_as_longlong(node_length(s, i)) : 0ull;
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int k = 0; k < (int)(blockDim.x / WAVE); k++) m = red[k] > m ? red[k] : m;
    atomicMax(out, m);
  }
}
__global__ void k_stop_argmax(DS s, const unsigned long long* maxbits, unsigned long long* idx) {
  const int64_t i = gid();
  if (i < s.N && (unsigned long long)__double_as_longlong(node_length(s, i)) == *maxbits)
    atomicMin(idx, (unsigned long long)i);
}
// max_particle_dist: distance of the first two particles' centre nodes
__global__ void k_stop_center_dist(DS s, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t a = s.P[0].center, b = s.P[1].center;
  double d2 = 0.;
  for (int d = 0; d < 3; d++) {
    const double t = s.x[d] ? s.x[d][a] - s.x[d][b] : 0.;
    d2 += t * t;
  }
  *out = sqrt(d2);
}

__global__ void k_count_broken(DS s) {
  int64_t i = gid();
  unsigned long long c = 0;
  if (i < s.N) {
    const int64_t sl = i / WAVE, lane = i % WAVE;
    const int n = s.deg[i];
    const unsigned* bw = s.bits + s.bptr[sl] + lane;
    for (int w = 0; w < (n + 31) / 32; w++) c += __popc(bw[(int64_t)w * WAVE]);
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & (WAVE - 1)) == 0 && c) atomicAdd(&s.ctrl->broken, c);
}

}  // namespace pdem
