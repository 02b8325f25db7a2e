// pdem_engine.hip -- host side of the MI355X PeriDEM step engine and the
// C-ABI declared in include/pdem.h.
//
// One context = one model on the current HIP device, all state resident in
// HBM.  Every step runs as a fixed sequence of kernels on one stream with no
// host decision inside it: the contact-search cadence controller and the
// simulation clock live in device memory (struct Ctrl), so a step can be
// captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pdem.h"
#include "pdem_host.hpp"
#include "pdem_kernels.hpp"
#include "pdem_mixed.hpp"
#include "pdem_closure.hpp"
#include "pdem_fe.hpp"
#include "pdem_rnp_tile.hpp"

using namespace pdem;

namespace {
thread_local std::string g_create_err;
}  // namespace

// Staging of pdem_pair_log_noalloc: a row counter (first 256 bytes) and up to
// PDEM_PAIR_LOG_NOALLOC_MAX_ROWS rows of features, targets and baseline.
// One copy per GPU (6.0 MB), held until the process ends; every call on that
// GPU shares it.
constexpr int64_t NOALLOC_ROWS = PDEM_PAIR_LOG_NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_BYTES = 256 + (size_t)NOALLOC_ROWS * (CL_NFEAT + 4 + 3) * sizeof(float);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
// Device address of `scratch` per device.  The array's name in host code is a
// host-side stand-in, not a device address, so kernels and copies take the
// address looked up here (with the device current) and never the name.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

void* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError{"no device address for the staging array"};
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

__global__ void k_zero_u32(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
}  // namespace

// multi-particle models up to this many nodes use the fused step kernels
constexpr int64_t LIGHT_2D_MIN_NODES = int64_t(1) << 20;  // 2D models from this size use the light-node contact kernel
constexpr int64_t FUSE_MAX_NODES = (int64_t)1 << 20;

struct pdem_ctx {
  int device = 0;  // HIP device current at pdem_create
  void* scratch_dev = nullptr;  // device address of `scratch` on `device`
  int dim = 2, precision = 0, multi = 0;
  int64_t N = 0;
  int np = 0, nz = 0;
  hipStream_t stream = nullptr;
  // side stream of the mixed bond passes: tiles that gather from HBM run on it
  // beside the LDS tiles (forked from and joined back into `stream`)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // second side stream: the cell list of a rebuild step, built beside the
  // bond passes (multi-particle models)
  hipStream_t side2 = nullptr;
  hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
  DS ds{};
  std::vector<void*> allocs;
  std::vector<pdem_particle> hp;
  std::vector<DevBc> ubc, fbc;
  std::vector<int> ubc_zero;
  bool bc_trig = false;  // a displacement-BC set uses the sine or rotation time function
  std::vector<char> zero_done;
  // device arrays that carry state from step to step (checkpoint / restore)
  std::vector<std::pair<void*, size_t>> persist;
  int64_t nslices = 0, nnz_padded = 0, nbitwords = 0, nnz = 0;
  std::vector<int64_t> h_sptr, h_bptr;
  std::vector<int> h_deg;
  bool state_any = false, rnp_any = false, pmb_any = false, elastic_any = false;
  int64_t cap = 0;
  MixedArgs mixed{};
  int ntiles = 0, ntiles_global = 0;
  DevTile* d_tiles = nullptr;
  size_t d_tiles_cap = 0;
  int pd_tpb = 256;        // FP64 bond kernels' workgroup size (autotuned)
  int mixed_tpb = 1024;    // mixed tile kernels' workgroup size (autotuned)
  int tile_nodes = TILE_NODES;
  std::vector<int> nbr_lo, nbr_hi;  // per-node neighbor id range (tile windows)
  int* d_slot = nullptr;            // mixed thread slot -> node
  int64_t* d_sp16 = nullptr;        // mixed 16-bit id slices, by slot
  uint16_t* d_col16 = nullptr;
  size_t col16_cap = 0;
  // learned pair-contact closure
  bool cl_ready = false;        // closure data allocated (pdem_closure_setup)
  ClosureArgs ca{};
  const pdem_mlp* cl_net = nullptr;
  bool graph_ready = false;
  bool graph_verlet = false;  // the captured step is integrateVerlet
  // bond-based RNP pair-tile kernel (FP64, all force-computing particles RNP)
  RnpTileArgs rt{};
  bool rnp_tile = false;  // use it (set up by default, autotuned)
  bool nf_zeroed = false;  // inside a step whose node update zeroed f of the non-force nodes
  int step_fusion = 1;  // fused step kernels: 0 never, 1 models up to FUSE_MAX_NODES, 2 any size
  bool state_fusion = true;  // closed state-based tiles run both passes in one kernel
  double* d_nfsave = nullptr;  // f of the non-force nodes around a standalone PD pass
  bool light_used = false;  // the lane-per-node contact list may hold nodes
  double* pp_out = nullptr;  // two-particle test results
  // strain / stress post-processing (pdem_fe_setup)
  FeArgs fe{};
  std::vector<void*> fe_bufs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used, ev_free;
  std::vector<int> ev_group;  // per used pair: the timed region it belongs to (regions on two streams: the longer)
  int ev_next_group = 0;
  double prof_ms = 0.;
  int64_t prof_n = 0;
  std::string err;

  template <class T>
  T* dalloc(size_t n, bool zero = true) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    allocs.push_back(p);
    // zeroed on the engine stream and waited for: the stream is non-blocking,
    // so a null-stream memset would not be ordered with the stream's kernels
    if (zero) {
      HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    return (T*)p;
  }
  void dfree(void* p) {
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) allocs.erase(it);
    HIPCHK(hipFree(p));
  }
  template <class T>
  T* upload(const T* h, size_t n) {
    T* d = dalloc<T>(n, false);
    if (n) HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
  // Teardown order: every launch on both streams completes first (a replayed
  // graph and its kernel arguments, events recorded by the profiling path and
  // the buffers they read stay valid until then), then the graph and the
  // events go, then the memory, then the streams.
  ~pdem_ctx() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    if (side2) (void)hipStreamSynchronize(side2);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : ev_used) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    for (auto& e : ev_free) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_fork2) (void)hipEventDestroy(ev_fork2);
    if (ev_join2) (void)hipEventDestroy(ev_join2);
    if (cl_net) cl_net->users.fetch_sub(1);
    cl_net = nullptr;
    for (void* p : allocs) (void)hipFree(p);
    if (side) (void)hipStreamDestroy(side);
    if (side2) (void)hipStreamDestroy(side2);
    if (stream) (void)hipStreamDestroy(stream);
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
  }
};

namespace {

// grid of the per-pair closure kernels: one wavefront per listed pair (at
// most cpair_cap), grid-stride beyond 16384 wavefronts
unsigned closure_blocks(pdem_ctx* c) {
  const int64_t waves = std::min<int64_t>(std::max(c->ds.cpair_cap, 1), 16384);
  return (unsigned)((waves + CL_WAVES - 1) / CL_WAVES);
}
// grid of the network kernel: 16 pending pairs per wavefront, grid-stride
// beyond 1024 wavefronts
unsigned net_blocks(pdem_ctx* c) {
  const int64_t waves = std::min<int64_t>((std::max(c->ds.cpair_cap, 1) + 15) / 16, 1024);
  return (unsigned)((waves + CL_WAVES - 1) / CL_WAVES);
}

template <int DIM>
struct Launch {
  // HIP events around the bond-force kernels when profiling: a timed region
  // is one or two event pairs (a kernel on the engine stream and the tiles
  // run beside it on the side stream); its time is the longer pair
  static std::pair<hipEvent_t, hipEvent_t> ev_begin(pdem_ctx* c, hipStream_t st) {
    std::pair<hipEvent_t, hipEvent_t> ev{};
    if (!c->profiling) return ev;
    if (c->ev_free.empty()) {
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      c->ev_free.push_back({a, b});
    }
    ev = c->ev_free.back();
    c->ev_free.pop_back();
    HIPCHK(hipEventRecord(ev.first, st));
    return ev;
  }
  static void ev_end(pdem_ctx* c, std::pair<hipEvent_t, hipEvent_t> ev, hipStream_t st, int group) {
    if (!c->profiling) return;
    HIPCHK(hipEventRecord(ev.second, st));
    c->ev_used.push_back(ev);
    c->ev_group.push_back(group);
  }
  // the side stream starts after everything enqueued so far on the engine stream
  static void fork(pdem_ctx* c) {
    HIPCHK(hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  }
  static void mixed_umax(pdem_ctx* c, hipStream_t st) {
    if (c->mixed.n_open_pids > 0)
      hipLaunchKernelGGL(k_mixed_umax<DIM>, dim3(c->mixed.n_open_pids), dim3(256), 0, st, c->ds, c->mixed);
  }
  template <int INF, int T>
  static void mixed_dilation(pdem_ctx* c) {
    const MixedArgs& ma = c->mixed;
    mixed_umax(c, c->stream);
    if (ma.ntiles_hbm) {
      fork(c);
      hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, false, T>), dim3(ma.ntiles_hbm), dim3(T), 0, c->side,
                         c->ds, ma);
      HIPCHK(hipEventRecord(c->ev_join, c->side));
    }
    if (ma.ntiles_lds && ma.small)
      hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, true, T, LDS_SMALL>), dim3(ma.ntiles_lds), dim3(T), 0, c->stream,
                         c->ds, ma);
    else if (ma.ntiles_lds)
      hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, true, T>), dim3(ma.ntiles_lds), dim3(T), 0, c->stream,
                         c->ds, ma);
    if (ma.ntiles_hbm) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  template <int INF, int T>
  static void mixed_force(pdem_ctx* c) {
    const MixedArgs& ma = c->mixed;
    if (ma.ntiles_hbm) {
      fork(c);
      hipLaunchKernelGGL((k_mixed_force<DIM, INF, false, T>), dim3(ma.ntiles_hbm), dim3(T), 0, c->side,
                         c->ds, ma);
      HIPCHK(hipEventRecord(c->ev_join, c->side));
    }
    if (ma.ntiles_lds && ma.small)
      hipLaunchKernelGGL((k_mixed_force<DIM, INF, true, T, LDS_SMALL>), dim3(ma.ntiles_lds), dim3(T), 0, c->stream,
                         c->ds, ma);
    else if (ma.ntiles_lds)
      hipLaunchKernelGGL((k_mixed_force<DIM, INF, true, T>), dim3(ma.ntiles_lds), dim3(T), 0, c->stream,
                         c->ds, ma);
    if (ma.ntiles_hbm) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  // The state-based passes of the mixed kernels: closed tiles (whole
  // particles) in the fused kernel on the engine stream; open LDS tiles and
  // HBM tiles in the two-pass kernels (dilation of all, then force of all,
  // since their force pass reads other tiles' dilation terms) on the side
  // stream beside them.  Without closed tiles the open LDS tiles keep the
  // engine stream and the HBM tiles the side stream.
  template <int INF, int T>
  static void mixed_state(pdem_ctx* c) {
    const MixedArgs& ma = c->mixed;
    const int grp = c->ev_next_group++;
    if (ma.ntiles_closed == 0) {
      auto ev = ev_begin(c, c->stream);
      mixed_dilation<INF, T>(c);
      mixed_force<INF, T>(c);
      ev_end(c, ev, c->stream, grp);
      return;
    }
    const bool two_pass = ma.ntiles_lds || ma.ntiles_hbm;
    // the fused kernel is enqueued first (right after the fork point), so a
    // host that enqueues behind the GPU does not delay it by the side-stream
    // launches
    if (two_pass) fork(c);
    auto ev = ev_begin(c, c->stream);
    if (ma.small)
      hipLaunchKernelGGL((k_mixed_state<DIM, INF, T, LDS_SMALL>), dim3(ma.ntiles_closed), dim3(T), 0, c->stream,
                         c->ds, ma);
    else
      hipLaunchKernelGGL((k_mixed_state<DIM, INF, T>), dim3(ma.ntiles_closed), dim3(T), 0, c->stream, c->ds, ma);
    ev_end(c, ev, c->stream, grp);
    if (two_pass) {
      auto es = ev_begin(c, c->side);
      mixed_umax(c, c->side);
      if (ma.ntiles_lds && ma.small)
        hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, true, T, LDS_SMALL>), dim3(ma.ntiles_lds), dim3(T), 0,
                           c->side, c->ds, ma);
      else if (ma.ntiles_lds)
        hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, true, T>), dim3(ma.ntiles_lds), dim3(T), 0, c->side, c->ds,
                           ma);
      if (ma.ntiles_hbm)
        hipLaunchKernelGGL((k_mixed_dilation<DIM, INF, false, T>), dim3(ma.ntiles_hbm), dim3(T), 0, c->side,
                           c->ds, ma);
      if (ma.ntiles_lds && ma.small)
        hipLaunchKernelGGL((k_mixed_force<DIM, INF, true, T, LDS_SMALL>), dim3(ma.ntiles_lds), dim3(T), 0, c->side,
                           c->ds, ma);
      else if (ma.ntiles_lds)
        hipLaunchKernelGGL((k_mixed_force<DIM, INF, true, T>), dim3(ma.ntiles_lds), dim3(T), 0, c->side, c->ds, ma);
      if (ma.ntiles_hbm)
        hipLaunchKernelGGL((k_mixed_force<DIM, INF, false, T>), dim3(ma.ntiles_hbm), dim3(T), 0, c->side, c->ds,
                           ma);
      ev_end(c, es, c->side, grp);
      HIPCHK(hipEventRecord(c->ev_join, c->side));
      HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
  }
  template <int INF>
  static void pd_inf(pdem_ctx* c) {
    hipStream_t st = c->stream;
    const unsigned nb = blocks(c->N);
    if (c->precision == PDEM_PREC_MIXED && c->ntiles > 0) {
      if (c->ds.n_noforce > 0 && !c->nf_zeroed)
        hipLaunchKernelGGL(k_zero_noforce<DIM>, dim3(std::min<unsigned>(nb, BZ_BLOCKS)), dim3(TPB), 0, st, c->ds);
      if (c->rnp_any) hipLaunchKernelGGL((k_pd_rnp<DIM, INF>), dim3(nb), dim3(TPB), 0, st, c->ds);
      if (c->pmb_any) hipLaunchKernelGGL((k_pd_bond<DIM, INF, 2>), dim3(nb), dim3(TPB), 0, st, c->ds);
      if (c->elastic_any) hipLaunchKernelGGL((k_pd_bond<DIM, INF, 3>), dim3(nb), dim3(TPB), 0, st, c->ds);
      // timed inside: both state-based passes (dilation and force), all tiles
      if (c->mixed_tpb == 256)
        mixed_state<INF, 256>(c);
      else
        mixed_state<INF, 1024>(c);
      return;
    }
    const unsigned pb = (unsigned)c->pd_tpb, pnb = blocks(c->N, c->pd_tpb);
    if (c->ds.state_global) hipLaunchKernelGGL((k_pd_dilation<DIM, INF>), dim3(pnb), dim3(pb), 0, st, c->ds);
    auto ev = ev_begin(c, st);
    if (c->rnp_tile)
      hipLaunchKernelGGL((k_pd_rnp_tile<DIM, INF>), dim3(c->rt.ntiles), dim3(RT_THREADS), c->rt.lds_bytes, st,
                         c->ds, c->rt);
    else if (c->rnp_any || !(c->state_any || c->pmb_any || c->elastic_any))
      hipLaunchKernelGGL((k_pd_rnp<DIM, INF>), dim3(pnb), dim3(pb), 0, st, c->ds);
    if (c->pmb_any) hipLaunchKernelGGL((k_pd_bond<DIM, INF, 2>), dim3(pnb), dim3(pb), 0, st, c->ds);
    if (c->elastic_any) hipLaunchKernelGGL((k_pd_bond<DIM, INF, 3>), dim3(pnb), dim3(pb), 0, st, c->ds);
    if (c->state_any) hipLaunchKernelGGL((k_pd_state<DIM, INF>), dim3(pnb), dim3(pb), 0, st, c->ds);
    ev_end(c, ev, st, c->ev_next_group++);
  }
  static void pd_forces(pdem_ctx* c) {
    if (c->ds.inf_type == 0)
      pd_inf<0>(c);
    else if (c->ds.inf_type == 1)
      pd_inf<1>(c);
    else
      pd_inf<2>(c);
  }
  // the cell list of a rebuild step (counting sort by cell key); every
  // kernel returns at once on other steps
  static void cells_build(pdem_ctx* c, hipStream_t st) {
    const unsigned nb = (unsigned)std::min<int64_t>(blocks(c->N), 2048);
    hipLaunchKernelGGL(k_cells_count<DIM>, dim3(nb), dim3(TPB), 0, st, c->ds);
    hipLaunchKernelGGL(k_cells_scan, dim3((unsigned)c->ds.ncell_blocks), dim3(CELL_TPB), 0, st, c->ds);
    hipLaunchKernelGGL(k_cells_scatter<DIM>, dim3((unsigned)std::min<int64_t>(blocks(c->N, CELL_TPB), 1024)),
                       dim3(CELL_TPB), 0, st, c->ds);
  }
  // PD forces and, on a rebuild step, the cell list on the second side stream
  // beside them (joined before the candidate query).  The fork point is
  // recorded first and the PD kernels are enqueued before the cell kernels:
  // the same dependencies, and a captured graph keeps the PD chain on the
  // engine's queue.
  static void pd_forces_and_cells(pdem_ctx* c) {
    if (!c->multi) {
      pd_forces(c);
      return;
    }
    HIPCHK(hipEventRecord(c->ev_fork2, c->stream));
    pd_forces(c);
    HIPCHK(hipStreamWaitEvent(c->side2, c->ev_fork2, 0));
    cells_build(c, c->side2);
    HIPCHK(hipEventRecord(c->ev_join2, c->side2));
    HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join2, 0));
  }
  // standalone search: controller, cell list, candidate query
  static void contact_search(pdem_ctx* c) {
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_ctrl, dim3(CTRL_BLOCKS), dim3(1024), 0, st, c->ds, 0);
    cells_build(c, st);
    rebuild_lists(c);
    hipLaunchKernelGGL(k_cells_init, dim3((unsigned)std::min<int64_t>(blocks(c->ds.ncells), 4096)), dim3(TPB), 0,
                       st, c->ds);
  }
  static void contact_force_kernel(pdem_ctx* c) {
    // lane-per-node list (grid-stride); empty unless a rebuild ran with light_max > 0
    if (c->light_used)
      hipLaunchKernelGGL(k_contact_light<DIM>, dim3((unsigned)std::min<int64_t>(blocks(c->N), 1024)), dim3(TPB), 0,
                         c->stream, c->ds);
    const int64_t waves = std::min<int64_t>(c->N, 8192);
    hipLaunchKernelGGL(k_contact_force<DIM>, dim3((unsigned)((waves + CWAVES - 1) / CWAVES)),
                       dim3(CWAVES * WAVE), 0, c->stream, c->ds);
  }
  // fin 1: the damping pass finalises the nodes; 2: it leaves the
  // finalisation to the next node update (f_pending)
  static void damping(pdem_ctx* c, int fin = 0) {
    if (c->ds.n_nonwall > 0) {
      // with fin 1: blocks after the particles add gravity to the wall nodes
      const int64_t wb =
          fin == 1 && c->ds.has_gravity ? std::min<int64_t>((c->ds.n_wall_nodes + 255) / 256, 256) : 0;
      hipLaunchKernelGGL(k_damping<DIM>, dim3((unsigned)(c->ds.n_nonwall + wb)), dim3(fin == 1 ? 256 : 64), 0,
                         c->stream, c->ds, fin);
    }
  }
  // Large models (unfused step) leave damping and gravity to the next node
  // update, which adds them to the forces it reads (one pass over f less per
  // step); a host read of f first finalises them (finalize_pending).  Not
  // with force-BC sets (their terms follow the finalisation), the closure
  // (its transfer runs in the finalisation) or Verlet steps (the second half
  // kick reads f).
  static bool defer_finalize(pdem_ctx* c, bool verlet) {
    return c->multi && !verlet && !fuse_finalize(c) && c->ds.n_nonwall > 0 && !c->ds.cl_on && c->fbc.empty();
  }
  // Fused step (node update with BC and centre data, one-workgroup
  // controller, damping pass finalising the nodes): fewer launches for
  // models small enough that the step is launch-bound.  Larger models fuse
  // the node update and the controller only where no BC set needs the sine
  // or rotation time function (whose FP64 sin/cos cost the fused update half
  // its occupancy), and keep the finalisation deferred to the next node
  // update (a finalising damping pass would add a pass over f).
  static bool large(pdem_ctx* c) { return c->step_fusion == 1 && c->N > FUSE_MAX_NODES; }
  static bool fuse_step(pdem_ctx* c) {
    if (!c->multi || c->step_fusion == 0 || (large(c) && c->bc_trig)) return false;
    for (size_t k = 0; k < c->ubc.size(); k++)
      if (c->ubc_zero[k] && !c->zero_done[k]) return false;  // must apply in set order first
    return true;
  }
  static bool fuse_finalize(pdem_ctx* c) { return fuse_step(c) && !large(c) && c->ds.n_nonwall > 0; }
  // rebuild steps: candidate query (a lane per node: the nodes' scans vary
  // in length, and more wavefronts in flight hide their latency; the launch
  // of other steps, which returns at once, costs under 1 us more than a
  // capped grid) and damping partner lists in one launch
  static void rebuild_lists(pdem_ctx* c) {
    const int qb = (int)blocks(c->N);
    const int db = (int)std::min<int64_t>((std::max(c->ds.n_nonwall, 1) + 3) / 4, 1024);
    hipLaunchKernelGGL(k_rebuild_lists<DIM>, dim3((unsigned)(qb + db)), dim3(TPB), 0, c->stream, c->ds, qb);
  }
  // closure pairs: connections (rebuild steps), modes, network, patch
  // records (2D, closure mode only)
  static void closure(pdem_ctx* c) {
    if (DIM == 2 && c->ds.cl_on) {
      hipLaunchKernelGGL(k_cl_connect, dim3((unsigned)std::min<int64_t>(blocks(c->N), CL_CONNECT_BLOCKS)), dim3(TPB),
                         0, c->stream, c->ds, 0);
      hipLaunchKernelGGL(k_closure, dim3(closure_blocks(c)), dim3(CL_WAVES * WAVE), 0, c->stream, c->ds, c->ca);
      if (c->ca.use_net)
        hipLaunchKernelGGL(k_cl_net, dim3(net_blocks(c)), dim3(CL_WAVES * WAVE), 0, c->stream, c->ds, c->ca);
    }
  }
  static void contact_forces(pdem_ctx* c, bool apply_damping) {
    hipStream_t st = c->stream;
    const unsigned nb = blocks(c->N);
    closure(c);
    contact_force_kernel(c);
    damping(c);
    if (apply_damping) hipLaunchKernelGGL(k_apply_damping<DIM>, dim3(nb), dim3(TPB), 0, st, c->ds);
  }
  static void external_forces(pdem_ctx* c, bool with_damping, bool finalized = false) {
    hipStream_t st = c->stream;
    if (!finalized && (c->ds.has_gravity || with_damping))
      hipLaunchKernelGGL(k_finalize<DIM>, dim3(blocks(c->N)), dim3(TPB), 0, st, c->ds, with_damping ? 1 : 0, 0);
    for (auto& b : c->fbc)
      if (b.n > 0) hipLaunchKernelGGL(k_fbc<DIM>, dim3(blocks(b.n)), dim3(TPB), 0, st, c->ds, b);
  }
  static void displacement_bc(pdem_ctx* c) {
    for (size_t s = 0; s < c->ubc.size(); s++) {
      if (c->ubc_zero[s]) {
        if (c->zero_done[s]) continue;
        c->zero_done[s] = 1;
      }
      const auto& b = c->ubc[s];
      if (b.n > 0) hipLaunchKernelGGL(k_ubc<DIM>, dim3(blocks(b.n)), dim3(TPB), 0, c->stream, c->ds, b);
    }
  }
  // forces after the controller has run: PD (binning on a rebuild), contact
  // (query on a rebuild, then forces), damping, external forces
  static void forces_after_ctrl(pdem_ctx* c, bool verlet = false) {
    pd_forces_and_cells(c);
    const bool fin = fuse_finalize(c);
    const bool defer = defer_finalize(c, verlet);
    if (c->multi) {
      rebuild_lists(c);
      closure(c);
      contact_force_kernel(c);
      damping(c, fin ? 1 : (defer ? 2 : 0));
    }
    external_forces(c, c->multi != 0, fin || defer);
  }
  static void compute_forces(pdem_ctx* c) {
    if (c->multi) hipLaunchKernelGGL(k_ctrl, dim3(CTRL_BLOCKS), dim3(1024), 0, c->stream, c->ds, 0);
    forces_after_ctrl(c);
  }
  // forces left pending by a deferred step, finalised in place
  static void finalize_pending(pdem_ctx* c) {
    if (!c->multi) return;
    hipLaunchKernelGGL(k_finalize<DIM>, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds, 1, 1);
    hipLaunchKernelGGL(k_clear_pending, dim3(1), dim3(64), 0, c->stream, c->ds);
  }
  // integrateCD, or integrateVerlet when `verlet` (half kick and drift,
  // forces, second half kick)
  static void integrate_step(pdem_ctx* c, bool verlet = false) {
    const bool fused = fuse_step(c);
    if (fused && c->bc_trig)
      hipLaunchKernelGGL((k_integrate<DIM, true, true>), dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds,
                         c->multi, verlet ? 1 : 0);
    else if (fused)
      hipLaunchKernelGGL((k_integrate<DIM, true, false>), dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds,
                         c->multi, verlet ? 1 : 0);
    else
      hipLaunchKernelGGL((k_integrate<DIM, false>), dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds, c->multi,
                         verlet ? 1 : 0);
    if (fused) {
      hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(1024), 0, c->stream, c->ds, 1);
    } else {
      displacement_bc(c);
      if (c->multi) hipLaunchKernelGGL(k_ctrl, dim3(CTRL_BLOCKS), dim3(1024), 0, c->stream, c->ds, 0);
    }
    c->nf_zeroed = c->ds.nf_zero != 0;  // the node update above zeroed f of the non-force nodes
    forces_after_ctrl(c, verlet);
    c->nf_zeroed = false;
    if (verlet) hipLaunchKernelGGL(k_kick<DIM>, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds);
  }
};

template <class F2, class F3>
void dispatch(pdem_ctx* c, F2 f2, F3 f3) {
  if (c->dim == 2)
    f2();
  else
    f3();
}

int set_err(pdem_ctx* c, const std::string& m) {
  if (c) c->err = m;
  g_create_err = m;
  return 1;
}

void check_launch(pdem_ctx* c) { HIPCHK(hipGetLastError()); }

// forces a deferred step left pending, finalised before a host reads f or
// a call adds to it
void finalize_pending(pdem_ctx* c) {
  dispatch(c, [&] { Launch<2>::finalize_pending(c); }, [&] { Launch<3>::finalize_pending(c); });
}
// f set without pending terms
void clear_pending(pdem_ctx* c) {
  if (c->multi) hipLaunchKernelGGL(k_clear_pending, dim3(1), dim3(64), 0, c->stream, c->ds);
}

// every node's candidate-position record from the node arrays (after the
// positions were set from the host or restored)
void refresh_xa(pdem_ctx* c) {
  if (c->ds.ctrl) hipLaunchKernelGGL(k_skin_invalidate, dim3(1), dim3(1), 0, c->stream, c->ds);
  if (!c->ds.xa) return;
  dispatch(
      c, [&] { hipLaunchKernelGGL(k_xa_all<2>, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds); },
      [&] { hipLaunchKernelGGL(k_xa_all<3>, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds); });
}

Ctrl read_ctrl(pdem_ctx* c) {
  Ctrl h;
  HIPCHK(hipMemcpyAsync(&h, c->ds.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return h;
}

void check_overflow(pdem_ctx* c) {
  Ctrl h = read_ctrl(c);
  if (h.overflow == 1)
    throw HipError{"contact candidate capacity exceeded (" + std::to_string(c->cap) +
                   " per node); create the model with a larger contact_capacity"};
  if (h.overflow == 4)
    throw HipError{"more than 256 damping partners within reach of one particle"};
  if (h.overflow == 3)
    throw HipError{"more than 64 active contact pairs at one node"};
  if (h.overflow == 5)
    throw HipError{"closure pair list capacity exceeded (" + std::to_string(c->ds.cpair_cap) + " pairs)"};
  if (h.overflow == 6) throw HipError{"closure: damping partner lists are not symmetric"};
}

// Pair tiles of the RNP kernel: each force-computing particle is cut into
// chunks of at most RT_THREADS consecutive nodes whose in-tile pairs fit
// RT_MAX_PAIRS; other particles get force-free chunks (zero and bin only).
// Per tile: its pairs (i < j, both inside) in slot order, node i's upper
// neighbors being a contiguous range of its ascending list; per bond: the
// pair slot of in-tile bonds (the partner's slot for lower neighbors).
void build_rnp_tiles(pdem_ctx* c) {
  const int64_t N = c->N;
  if (N > INT_MAX / 4) return;
  std::vector<int> col((size_t)c->nnz_padded);
  HIPCHK(hipMemcpyAsync(col.data(), c->ds.col, col.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  auto at = [&](int64_t i, int k) { return c->h_sptr[i / WAVE] + (int64_t)k * WAVE + i % WAVE; };
  std::vector<uint16_t> ps((size_t)c->nnz_padded, RT_OUT);
  std::vector<uint32_t> pl;
  std::vector<RtTile> tiles;
  // upper in-tile range [kb, ke) of node i in a tile ending at t1
  auto range = [&](int64_t i, int64_t t1, int& kb, int& ke) {
    const int n = c->h_deg[i];
    kb = 0;
    while (kb < n && col[at(i, kb)] <= i) kb++;
    ke = kb;
    while (ke < n && col[at(i, ke)] < t1) ke++;
  };
  auto pairs = [&](int64_t t0, int64_t t1) {
    int64_t np = 0;
    for (int64_t i = t0; i < t1; i++) {
      int kb, ke;
      range(i, t1, kb, ke);
      np += ke - kb;
    }
    return np;
  };
  std::vector<std::pair<int64_t, int64_t>> work;
  for (const auto& q : c->hp) {
    const bool w = q.compute_force && q.material_type == PDEM_MAT_RNP;
    const int64_t n = q.glob_end - q.glob_start;
    const int64_t nt = (n + RT_THREADS - 1) / RT_THREADS;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int64_t t0 = q.glob_start; t0 < q.glob_end; t0 += chunk) {
      const int64_t t1 = std::min<int64_t>(t0 + chunk, q.glob_end);
      if (!w) {
        tiles.push_back(RtTile{(int)t0, (int)t1, 0, 0, 0});
        continue;
      }
      // halve until the pairs fit
      std::vector<std::pair<int64_t, int64_t>> st{{t0, t1}};
      while (!st.empty()) {
        auto r = st.back();
        st.pop_back();
        if (r.second - r.first > 1 && pairs(r.first, r.second) > RT_MAX_PAIRS) {
          const int64_t m = (r.first + r.second) / 2;
          st.push_back({m, r.second});
          st.push_back({r.first, m});
          continue;
        }
        work.push_back(r);
      }
    }
  }
  size_t lds = 0;
  std::vector<int> kb_(RT_THREADS), ke_(RT_THREADS), base_(RT_THREADS);
  for (const auto& r : work) {
    const int64_t t0 = r.first, t1 = r.second;
    const int64_t pbase = (int64_t)pl.size();
    int base = 0;
    for (int64_t i = t0; i < t1; i++) {
      int kb, ke;
      range(i, t1, kb, ke);
      kb_[i - t0] = kb;
      ke_[i - t0] = ke;
      base_[i - t0] = base;
      for (int k = kb; k < ke; k++) {
        ps[at(i, k)] = (uint16_t)(base + k - kb);
        pl.push_back((uint32_t)(i - t0) | ((uint32_t)(col[at(i, k)] - t0) << 16));
      }
      base += ke - kb;
    }
    if (base > RT_MAX_PAIRS) return;  // a single node with too many neighbors: keep the lane kernel
    // lower in-tile neighbors take the partner's slot
    for (int64_t i = t0; i < t1; i++) {
      const int n = c->h_deg[i];
      for (int k = 0; k < n; k++) {
        const int64_t j = col[at(i, k)];
        if (j < t0 || j >= i) continue;
        const int jb = kb_[j - t0], je = ke_[j - t0];
        int lo = jb, hi = je;  // ascending: binary search for i in j's upper range
        while (lo < hi) {
          const int m = (lo + hi) / 2;
          if (col[at(j, m)] < i) lo = m + 1; else hi = m;
        }
        if (lo >= je || col[at(j, lo)] != i) throw HipError{"PD neighbor lists are not symmetric"};
        ps[at(i, k)] = (uint16_t)(base_[j - t0] + lo - jb);
      }
    }
    tiles.push_back(RtTile{(int)t0, (int)t1, base, 1, pbase});
    lds = std::max(lds, (size_t)(base + (2 * c->dim + 1) * (t1 - t0)) * sizeof(double) + (size_t)base * sizeof(float));
  }
  // pass-2 lane order: inside every tile, nodes by degree (most first), then id
  std::vector<uint16_t> perm((size_t)N, 0);
  for (const auto& t : tiles) {
    const int nt = t.t1 - t.t0;
    std::vector<int> loc(nt);
    for (int q = 0; q < nt; q++) loc[q] = q;
    std::stable_sort(loc.begin(), loc.end(),
                     [&](int x, int y) { return c->h_deg[t.t0 + x] > c->h_deg[t.t0 + y]; });
    for (int q = 0; q < nt; q++) perm[(size_t)t.t0 + q] = (uint16_t)loc[q];
  }
  RnpTileArgs& a = c->rt;
  a.perm = c->upload(perm.data(), perm.size());
  a.tiles = c->upload(tiles.data(), tiles.size());
  a.pl = c->upload(pl.data(), pl.size());
  a.ps = c->upload(ps.data(), ps.size());
  a.ntiles = (int)tiles.size();
  a.lds_bytes = lds;
  dispatch(
      c,
      [&] {
        for (auto f : {(const void*)k_pd_rnp_tile<2, 0>, (const void*)k_pd_rnp_tile<2, 1>,
                       (const void*)k_pd_rnp_tile<2, 2>})
          HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      },
      [&] {
        for (auto f : {(const void*)k_pd_rnp_tile<3, 0>, (const void*)k_pd_rnp_tile<3, 1>,
                       (const void*)k_pd_rnp_tile<3, 2>})
          HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      });
  c->rnp_tile = true;
}

// Tiles of the mixed kernels: for every state-based particle, the largest
// tile size (power of two, 64..tile_max) whose neighbor windows all fit in
// LDS, else tiles of tile_max nodes that gather from HBM; LDS tiles first.
void build_tiles(pdem_ctx* c, int64_t tile_max) {
  const auto& lo = c->nbr_lo;
  const auto& hi = c->nbr_hi;
  std::vector<DevTile> tiles;
  for (int p = 0; p < c->np; p++) {
    const auto& q = c->hp[p];
    if (!q.compute_force || q.material_type != PDEM_MAT_PDSTATE) continue;
    auto window = [&](int64_t ts, int64_t te, int& ws, int& we) {
      ws = (int)ts;
      we = (int)te;
      for (int64_t i = ts; i < te; i++) {
        ws = std::min(ws, lo[i]);
        we = std::max(we, hi[i] + 1);
      }
    };
    int64_t tile = tile_max;
    bool fits = false;
    for (; tile >= 64; tile /= 2) {
      fits = true;
      for (int64_t ts = q.glob_start; ts < q.glob_end && fits; ts += tile) {
        int ws, we;
        window(ts, std::min<int64_t>(q.glob_end, ts + tile), ws, we);
        fits = (we - ws) <= LDS_NODES;
      }
      if (fits) break;
    }
    // no tile size down to 64 nodes fits (a long, thin numbering such as a
    // meshed wall): full-size tiles that gather from HBM
    if (!fits) tile = tile_max;
    for (int64_t ts = q.glob_start; ts < q.glob_end; ts += tile) {
      const int64_t te = std::min<int64_t>(q.glob_end, ts + tile);
      int ws, we;
      window(ts, te, ws, we);
      const bool lds = (we - ws) <= LDS_NODES;
      const bool closed = lds && ts == q.glob_start && te == q.glob_end;
      tiles.push_back(DevTile{(int)ts, (int)te, ws, we, p, closed ? 2 : lds ? 1 : 0});
    }
  }
  // thread slots: inside every tile, nodes by 8-bond block count (most
  // first), then by id; the 16-bit id lists follow the slot order
  const int64_t N = c->ds.N;
  std::vector<int> slot(N);
  for (int64_t q = 0; q < N; q++) slot[q] = (int)q;
  for (const auto& t : tiles)
    std::stable_sort(slot.begin() + t.ts, slot.begin() + t.te, [&](int a, int b) {
      return (c->h_deg[a] + KB - 1) / KB > (c->h_deg[b] + KB - 1) / KB;
    });
  std::vector<int64_t> sp16(c->nslices + 1, 0);
  for (int64_t sl = 0; sl < c->nslices; sl++) {
    int w = 0;
    for (int64_t q = sl * WAVE; q < std::min<int64_t>(N, (sl + 1) * WAVE); q++) w = std::max(w, c->h_deg[slot[q]]);
    sp16[sl + 1] = sp16[sl] + (int64_t)((w + KB - 1) / KB) * KB * WAVE;
  }
  const size_t need = (size_t)sp16[c->nslices] + 8;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (need > c->col16_cap) {
    if (c->d_col16) c->dfree(c->d_col16);
    c->d_col16 = c->dalloc<uint16_t>(need);
    c->col16_cap = need;
  }
  HIPCHK(hipMemcpy(c->d_slot, slot.data(), N * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_sp16, sp16.data(), sp16.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_col16, dim3(blocks(N)), dim3(TPB), 0, c->stream, c->ds, c->d_col16, c->d_sp16,
                     (const int*)c->d_slot);
  check_launch(c);
  c->mixed.slot = c->d_slot;
  c->mixed.col16 = c->d_col16;
  c->mixed.sptr16 = c->d_sp16;
  // closed tiles (fused kernel), then open LDS tiles, then HBM tiles
  std::stable_sort(tiles.begin(), tiles.end(), [](const DevTile& a, const DevTile& b) { return a.lds > b.lds; });
  if (tiles.size() > c->d_tiles_cap) throw HipError{"tile buffer too small"};
  c->ntiles = (int)tiles.size();
  c->ntiles_global = 0;
  int nclosed = 0;
  for (auto& t : tiles) {
    c->ntiles_global += t.lds ? 0 : 1;
    nclosed += t.lds == 2 ? 1 : 0;
  }
  if (!c->state_fusion) {  // every LDS tile through the two-pass kernels
    nclosed = 0;
    for (auto& t : tiles) t.lds = t.lds ? 1 : 0;
  }
  // particles with open LDS tiles: their windows' fixed-point step comes from
  // a per-particle maximum computed before the passes
  {
    std::vector<int> open;
    for (auto& t : tiles)
      if (t.lds == 1 && (open.empty() || open.back() != t.pid)) open.push_back(t.pid);
    std::sort(open.begin(), open.end());
    open.erase(std::unique(open.begin(), open.end()), open.end());
    if (!c->mixed.open_pids) c->mixed.open_pids = c->dalloc<int>(std::max(1, c->np));
    if (!c->mixed.pumax) c->mixed.pumax = c->dalloc<float>(std::max(1, c->np));
    if (!open.empty())
      HIPCHK(hipMemcpy((void*)c->mixed.open_pids, open.data(), open.size() * sizeof(int), hipMemcpyHostToDevice));
    c->mixed.n_open_pids = (int)open.size();
  }
  c->mixed.ntiles_closed = nclosed;
  c->mixed.ntiles_lds = c->ntiles - c->ntiles_global - nclosed;
  int maxw = 0;
  for (auto& t : tiles)
    if (t.lds) maxw = std::max(maxw, t.we - t.ws);
  c->mixed.small = maxw <= LDS_SMALL ? 1 : 0;
  c->mixed.ntiles_hbm = c->ntiles_global;
  c->tile_nodes = (int)tile_max;
  if (!tiles.empty())
    HIPCHK(hipMemcpy(c->d_tiles, tiles.data(), tiles.size() * sizeof(DevTile), hipMemcpyHostToDevice));
  c->mixed.tiles = c->d_tiles;
}

void build(pdem_ctx* c, const pdem_desc* d) {
  const int64_t N = d->n_nodes;
  const int dim = d->dim;
  c->N = N;
  c->dim = dim;
  c->precision = d->precision;
  c->multi = d->multi_particle;
  HIPCHK(hipGetDevice(&c->device));
  c->scratch_dev = scratch_address(c->device);
  c->np = d->n_particles;
  c->nz = d->n_zones;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  DS& s = c->ds;
  s.N = N;
  s.np = c->np;
  s.nz = c->nz;
  // ---- node state, AoS host -> SoA device
  std::vector<double> tmp(N);
  auto split = [&](const double* aos, double** out, bool add_ref) {
    for (int k = 0; k < dim; k++) {
      for (int64_t i = 0; i < N; i++) tmp[i] = aos ? aos[3 * i + k] : 0.;
      if (add_ref)
        for (int64_t i = 0; i < N; i++) tmp[i] = d->x_ref[3 * i + k] + tmp[i];
      out[k] = c->upload(tmp.data(), N);
    }
    for (int k = dim; k < 3; k++) out[k] = nullptr;
  };
  split(d->x_ref, s.xr, false);
  split(d->u, s.u, false);
  split(d->v, s.v, false);
  split(d->u, s.x, true);
  for (int k = 0; k < dim; k++) s.f[k] = c->dalloc<double>(N);
  for (int k = dim; k < 3; k++) s.f[k] = nullptr;
  s.vol = c->upload(d->vol, N);
  s.mx = c->dalloc<double>(N);
  s.theta = c->dalloc<double>(N);
  s.vmag = c->dalloc<double>(N);
  s.Z = c->dalloc<float>(N);
  s.fix = c->upload(d->fix, N);
  std::vector<int> pt32(N);
  for (int64_t i = 0; i < N; i++) pt32[i] = (int)d->pt_id[i];
  s.ptid = c->upload(pt32.data(), N);
  // ---- particles
  c->hp.assign(d->particles, d->particles + c->np);
  std::vector<DevParticle> dp(c->np);
  std::vector<int> nonwall;
  for (int p = 0; p < c->np; p++) {
    const auto& q = c->hp[p];
    DevParticle& o = dp[p];
    o.glob_start = q.glob_start;
    o.glob_end = q.glob_end;
    o.center = q.center_node;
    o.zone = q.zone;
    o.type_index = q.type_index;
    o.compute_force = q.compute_force;
    o.material_type = q.material_type;
    o.irreversible = q.irreversible;
    o.searches = (q.type_index != 1 && !q.all_dofs_constrained && q.compute_force) ? 1 : 0;
    o.nonwall_index = -1;
    if (q.type_index == 0) {
      o.nonwall_index = (int)nonwall.size();
      nonwall.push_back(p);
    }
    o.rho = q.density;
    o.horizon = q.horizon;
    o.h = q.mesh_size;
    o.Rb = q.bounding_radius;
    o.Kn = q.Kn;
    o.Rc = q.Rc;
    for (int k = 0; k < 6; k++) o.mat[k] = q.mat[k];
    if (q.material_type == PDEM_MAT_PDSTATE) c->state_any = true;
    if (q.material_type == PDEM_MAT_RNP) c->rnp_any = true;
    if (q.material_type == PDEM_MAT_PMB) c->pmb_any = true;
    if (q.material_type == PDEM_MAT_PDELASTIC) c->elastic_any = true;
  }
  s.P = c->upload(dp.data(), dp.size());
  std::vector<DevContact> dc((size_t)c->nz * c->nz);
  for (size_t k = 0; k < dc.size(); k++) {
    const auto& q = d->contact[k];
    dc[k] = DevContact{q.contact_r, q.Kn, q.mu, q.betan, q.kappa};
  }
  s.CT = c->upload(dc.data(), dc.size());
  s.inf_type = d->influence_fn_type;
  s.inf_p0 = d->influence_fn_params[0];
  s.inf_p1 = d->influence_fn_params[1];
  s.state_global = c->hp[0].material_type == PDEM_MAT_PDSTATE ? 1 : 0;
  double glen = std::sqrt(d->gravity[0] * d->gravity[0] + d->gravity[1] * d->gravity[1] +
                          (dim == 3 ? d->gravity[2] * d->gravity[2] : 0.));
  s.has_gravity = glen > 1.0E-8 ? 1 : 0;
  for (int k = 0; k < 3; k++) s.gravity[k] = k < dim ? d->gravity[k] : 0.;
  // ---- BC sets
  for (int k = 0; k < d->n_ubc; k++) {
    const auto& b = d->ubc[k];
    DevBc o{};
    o.time_fn = b.time_fn;
    o.dir_mask = b.dir_mask;
    for (int q = 0; q < 4; q++) o.tp[q] = b.time_params[q];
    o.n = b.n_entries;
    o.nodes = c->upload(b.nodes, b.n_entries);
    o.amp = c->upload(b.amp, b.n_entries);
    c->ubc.push_back(o);
    c->ubc_zero.push_back(b.is_zero ? 1 : 0);
    if (b.time_fn == 3 || b.time_fn == 4) c->bc_trig = true;
  }
  c->zero_done.assign(c->ubc.size(), 0);
  {
    // per-node entries of the sets applied every step (not the zero sets),
    // in set order, for the fused node update
    std::vector<int> cnt(N + 1, 0);
    std::vector<BcSetParams> sets(d->n_ubc + 1);
    for (int k = 0; k < d->n_ubc; k++) {
      const auto& b = d->ubc[k];
      sets[k].time_fn = b.time_fn;
      sets[k].dir_mask = b.dir_mask;
      for (int q = 0; q < 4; q++) sets[k].tp[q] = b.time_params[q];
      if (b.is_zero) continue;
      for (int64_t e = 0; e < b.n_entries; e++) cnt[b.nodes[e] + 1]++;
    }
    for (int64_t i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    if (cnt[N] > INT_MAX / 2) throw HipError{"too many displacement-BC entries"};
    std::vector<BcEntry> ent(cnt[N] + 1);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < d->n_ubc; k++) {
      const auto& b = d->ubc[k];
      if (b.is_zero) continue;
      for (int64_t e = 0; e < b.n_entries; e++) ent[fill[b.nodes[e]]++] = BcEntry{b.amp[e], k, 0};
    }
    s.bc_ptr = c->upload(cnt.data(), cnt.size());
    s.bc_ent = c->upload(ent.data(), ent.size());
    s.bc_sets = c->upload(sets.data(), sets.size());
  }
  for (int k = 0; k < d->n_fbc; k++) {
    const auto& b = d->fbc[k];
    DevBc o{};
    o.time_fn = b.time_fn;
    o.dir_mask = b.dir_mask;
    for (int q = 0; q < 4; q++) o.tp[q] = b.time_params[q];
    o.n = b.n_entries;
    o.nodes = c->upload(b.nodes, b.n_entries);
    o.amp = c->upload(b.amp, b.n_entries);
    c->fbc.push_back(o);
  }
  // ---- control block
  Ctrl h{};
  h.dt = d->dt;
  h.s_factor = d->search_factor;
  h.deck_interval = d->neigh_update_interval;
  double maxR = 0.;
  if (c->multi)
    for (auto& q : dc)
      if (maxR < q.contact_r) maxR = q.contact_r;
  h.max_contact_r = maxR;
  h.dsum = INFINITY;  // no rebuild yet: the contact skin bound is unknown
  s.ctrl = c->upload(&h, 1);

  // ---- PD neighbor list, sliced ELL built on device
  s.deg = c->dalloc<int>(N);
  dispatch(
      c, [&] { hipLaunchKernelGGL(k_pd_count<2>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); },
      [&] { hipLaunchKernelGGL(k_pd_count<3>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); });
  check_launch(c);
  c->h_deg.resize(N);
  HIPCHK(hipMemcpyAsync(c->h_deg.data(), s.deg, N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nslices = (N + WAVE - 1) / WAVE;
  c->h_sptr.assign(c->nslices + 1, 0);
  c->h_bptr.assign(c->nslices + 1, 0);
  c->nnz = 0;
  for (int64_t sl = 0; sl < c->nslices; sl++) {
    int w = 0;
    for (int64_t i = sl * WAVE; i < std::min(N, (sl + 1) * WAVE); i++) {
      w = std::max(w, c->h_deg[i]);
      c->nnz += c->h_deg[i];
    }
    c->h_sptr[sl + 1] = c->h_sptr[sl] + (int64_t)w * WAVE;
    c->h_bptr[sl + 1] = c->h_bptr[sl] + (int64_t)((w + 31) / 32) * WAVE;
  }
  c->nnz_padded = c->h_sptr[c->nslices];
  c->nbitwords = c->h_bptr[c->nslices];
  s.sptr = c->upload(c->h_sptr.data(), c->h_sptr.size());
  s.bptr = c->upload(c->h_bptr.data(), c->h_bptr.size());
  s.col = c->dalloc<int>(c->nnz_padded);
  s.bits = c->dalloc<unsigned>(c->nbitwords);
  dispatch(
      c, [&] { hipLaunchKernelGGL(k_pd_fill<2>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); },
      [&] { hipLaunchKernelGGL(k_pd_fill<3>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); });
  check_launch(c);
  if (c->rnp_any && !c->state_any && !c->pmb_any && !c->elastic_any) build_rnp_tiles(c);

  // ---- mixed precision: 16-bit neighbor ids and particle tiles
  if (c->precision == PDEM_PREC_MIXED && c->state_any) {
    if (!s.state_global) throw HipError{"mixed precision needs a state-based first particle"};
    for (auto& q : c->hp)
      if (q.glob_end - q.glob_start > 65536)
        throw HipError{"mixed precision needs particles of at most 65536 nodes"};
    HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    c->d_slot = c->dalloc<int>(N, false);
    c->d_sp16 = c->dalloc<int64_t>(c->nslices + 1, false);
    int* dlo = c->dalloc<int>(N);
    int* dhi = c->dalloc<int>(N);
    hipLaunchKernelGGL(k_nbr_range, dim3(blocks(N)), dim3(TPB), 0, c->stream, s, dlo, dhi);
    check_launch(c);
    c->nbr_lo.resize(N);
    c->nbr_hi.resize(N);
    HIPCHK(hipMemcpyAsync(c->nbr_lo.data(), dlo, N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->nbr_hi.data(), dhi, N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // tile buffer sized for the smallest tile size the autotuner tries
    int64_t maxt = 1;
    for (auto& q : c->hp)
      if (q.compute_force && q.material_type == PDEM_MAT_PDSTATE) maxt += (q.glob_end - q.glob_start + 63) / 64;
    c->d_tiles = c->dalloc<DevTile>(maxt);
    c->d_tiles_cap = maxt;
    build_tiles(c, TILE_NODES);
    c->mixed.fa = c->dalloc<float>(N);
    c->mixed.fb = c->dalloc<float>(N);
  }

  // ---- contact structures
  if (c->multi) {
    // uniform grid: csub cells span the largest search radius the controller
    // can pick; the query scans +-csub cells around a node
    s.csub = dim == 3 ? 2 : 1;
    double cell = std::max(1.0, d->search_factor) * maxR / s.csub;
    if (!(cell > 0.)) cell = 1.;
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = 1e300;
      hi[k] = -1e300;
    }
    for (int64_t i = 0; i < N; i++)
      for (int k = 0; k < dim; k++) {
        lo[k] = std::min(lo[k], d->x_ref[3 * i + k]);
        hi[k] = std::max(hi[k], d->x_ref[3 * i + k]);
      }
    int64_t ncells = 1;
    for (;;) {
      ncells = 1;
      for (int k = 0; k < dim; k++) {
        double ext = hi[k] - lo[k];
        double margin = 0.05 * ext + 2 * cell;
        s.glo[k] = lo[k] - margin;
        s.gn[k] = (int)std::ceil((ext + 2 * margin) / cell) + 1;
        ncells *= s.gn[k];
      }
      if (ncells <= (int64_t)1 << 25) break;
      cell *= 1.25;
    }
    for (int k = dim; k < 3; k++) {
      s.glo[k] = 0.;
      s.gn[k] = 1;
    }
    s.cell = cell;
    s.ncells = ncells;
    s.ncell_blocks = (ncells + CELL_BLOCK - 1) / CELL_BLOCK;
    s.cell_count = c->dalloc<int>(ncells + 1);
    s.cell_owner = c->dalloc<int>(ncells + 1);
    hipLaunchKernelGGL(k_cells_init, dim3((unsigned)std::min<int64_t>(blocks(ncells), 4096)), dim3(TPB), 0,
                       c->stream, s);
    s.cell_start = c->dalloc<int>(ncells + 1, false);
    s.cell_btot = c->dalloc<int>(s.ncell_blocks, false);
    s.cell_boff = c->dalloc<int>(s.ncell_blocks, false);
    s.node_cell = c->dalloc<int>(N, false);
    s.node_slot = c->dalloc<int>(N, false);
    s.cell_nodes = c->dalloc<int>(N, false);
    s.cell_xf = c->dalloc<float4>(N, false);
    s.xa = c->dalloc<double4>(N, false);
    s.xa_epoch = c->dalloc<uint8_t>(N);
    HIPCHK(hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming));
    // candidate capacity: nodes of a ball of the largest search radius at the
    // finest mesh spacing
    double hmin = 1e300;
    for (auto& q : c->hp) hmin = std::min(hmin, q.mesh_size);
    double rr = std::max(1.0, d->search_factor) * maxR / hmin;
    double est = dim == 2 ? 1.25 * M_PI * rr * rr : 0.75 * 4.18879 * rr * rr * rr;
    int64_t cap = d->contact_capacity > 0 ? d->contact_capacity : (int64_t)std::ceil(est);
    cap = std::max<int64_t>(32, (cap + 31) / 32 * 32);
    c->cap = cap;
    s.cap = (int)cap;
    s.ccount = c->dalloc<int>(N);
    s.active = c->dalloc<int>(N);
    s.light = c->dalloc<int>(N);
    s.cand = c->dalloc<int>((size_t)cap * N, false);
    // damping bookkeeping: wall node compact ids, per-particle wall bitmaps
    std::vector<int> widx(N, -1), wgid;
    for (int p = 0; p < c->np; p++)
      if (c->hp[p].type_index == 1)
        for (int64_t i = c->hp[p].glob_start; i < c->hp[p].glob_end; i++) {
          widx[i] = (int)wgid.size();
          wgid.push_back((int)i);
        }
    s.nwall = (int)wgid.size();
    s.nwords = (s.nwall + 31) / 32;
    s.widx = c->upload(widx.data(), N);
    s.n_nonwall = (int)nonwall.size();
    if (wgid.empty()) wgid.push_back(0);
    if (nonwall.empty()) nonwall.push_back(0);
    s.wall_gid = c->upload(wgid.data(), wgid.size());
    s.nonwall_pid = c->upload(nonwall.data(), nonwall.size());
    {
      std::vector<int64_t> cn(nonwall.size());
      for (size_t q = 0; q < nonwall.size(); q++) cn[q] = c->hp[nonwall[q]].center_node;
      s.nonwall_center = c->upload(cn.data(), cn.size());
    }
    s.wbits = c->dalloc<unsigned>((size_t)std::max(1, s.n_nonwall) * std::max(1, s.nwords));
    std::vector<int> lo_init(std::max(1, s.n_nonwall), INT_MAX), hi_init(std::max(1, s.n_nonwall), -1);
    s.wlo = c->upload(lo_init.data(), lo_init.size());
    s.whi = c->upload(hi_init.data(), hi_init.size());
    s.fdamp = c->dalloc<double>((size_t)c->np * 3);
    {
      const int nn = std::max(1, s.n_nonwall);
      std::vector<double> rad(nn, 0.), rho(nn, 0.);
      std::vector<int> zn(nn, 0);
      for (int q = 0; q < s.n_nonwall; q++) {
        rad[q] = c->hp[nonwall[q]].bounding_radius;
        rho[q] = c->hp[nonwall[q]].density;
        zn[q] = c->hp[nonwall[q]].zone;
      }
      s.prad = c->upload(rad.data(), nn);
      s.prho = c->upload(rho.data(), nn);
      s.pzone = c->upload(zn.data(), nn);
      s.pcx = c->dalloc<double>(3 * (size_t)nn);
      s.pcv = c->dalloc<double>(3 * (size_t)nn);
      s.dcap = std::min(nn, 256);
      s.dcount = c->dalloc<int>(nn);
      s.dlist = c->dalloc<int>((size_t)nn * s.dcap);
    }
  }
  {
    std::vector<int> meta(c->np);
    for (int p = 0; p < c->np; p++) meta[p] = (c->hp[p].zone & 0xffff) | ((c->hp[p].type_index == 1) << 16);
    s.pmeta = c->upload(meta.data(), meta.size());
    double rc = 0.;
    for (int k = 0; k < c->nz * c->nz && c->multi; k++) rc = std::max(rc, d->contact[k].contact_r);
    s.rc2_max = rc * rc * (1.0 + 1e-12);
    s.rc_skip = rc * (1.0 + 1e-9);
    s.cdmin = c->dalloc<float>(N, false);
    s.n_ubc_sets = (int)c->ubc.size();
    // a lane per node for nodes with few candidates pays off where many nodes
    // have candidates (3D, and large 2D models such as batched ensembles:
    // 0.66 -> 0.60 ms per step for 512 instances); a small 2D model with few
    // contacts does better with the wavefront-per-node kernel alone (one
    // launch less per step: 87 -> 79 us for the 500-particle compressive test)
    s.light_max = (dim == 3 || N >= LIGHT_2D_MIN_NODES) ? LIGHT_MAX_C : 0;
    c->light_used = s.light_max > 0;
    s.nblockmax = (int)blocks(N);
    s.blockmax = c->dalloc<unsigned long long>(s.nblockmax);
    s.blockmv = c->dalloc<unsigned long long>(s.nblockmax);
    std::vector<int64_t> wn;
    for (int p = 0; p < c->np; p++)
      if (c->hp[p].type_index != 0)
        for (int64_t i = c->hp[p].glob_start; i < c->hp[p].glob_end; i++) wn.push_back(i);
    s.n_wall_nodes = (int64_t)wn.size();
    s.wall_nodes = c->upload(wn.data(), wn.size());
  }
  {
    std::vector<int64_t> nf;
    for (int p = 0; p < c->np; p++)
      if (!c->hp[p].compute_force)
        for (int64_t i = c->hp[p].glob_start; i < c->hp[p].glob_end; i++) nf.push_back(i);
    s.n_noforce = (int64_t)nf.size();
    s.noforce_nodes = c->upload(nf.data(), nf.size());
    // the mixed bond kernels leave f of these nodes alone: the node update
    // of a step zeroes it (other force passes launch k_zero_noforce)
    s.nf_zero = (c->precision == PDEM_PREC_MIXED && c->ntiles > 0 && s.n_noforce > 0) ? 1 : 0;
  }
  // state carried between steps: node fields, bond bits, damage, controller,
  // candidate and partner lists of the last rebuild (everything else is
  // constant or recomputed inside every step)
  for (int k = 0; k < dim; k++)
    for (double* a : {s.x[k], s.u[k], s.v[k], s.f[k]}) c->persist.push_back({a, (size_t)N * 8});
  c->persist.push_back({s.vmag, (size_t)N * 8});
  c->persist.push_back({s.theta, (size_t)N * 8});
  c->persist.push_back({s.Z, (size_t)N * 4});
  c->persist.push_back({s.bits, (size_t)c->nbitwords * 4});
  c->persist.push_back({s.ctrl, sizeof(Ctrl)});
  c->persist.push_back({s.blockmax, (size_t)s.nblockmax * 8});
  if (c->multi) {
    c->persist.push_back({s.ccount, (size_t)N * 4});
    c->persist.push_back({s.cand, (size_t)c->cap * N * 4});
    c->persist.push_back({s.active, (size_t)N * 4});
    c->persist.push_back({s.light, (size_t)N * 4});
    const size_t nn = (size_t)std::max(1, s.n_nonwall);
    c->persist.push_back({s.dcount, nn * 4});
    c->persist.push_back({s.dlist, nn * s.dcap * 4});
    c->persist.push_back({s.xa_epoch, (size_t)N});
  }
  // weighted volume for every node (computeStateMx), then the init-time
  // contact search (DEMModel::init calls updateContactNeighborlist once)
  dispatch(
      c, [&] { hipLaunchKernelGGL(k_mx<2>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); },
      [&] { hipLaunchKernelGGL(k_mx<3>, dim3(blocks(N)), dim3(TPB), 0, c->stream, s); });
  // nodes that never search keep an empty candidate list
  check_launch(c);
  if (c->multi) {
    Ctrl h2 = h;
    h2.interval = d->neigh_update_interval;
    h2.radius = d->search_factor * maxR;
    HIPCHK(hipMemcpyAsync(s.ctrl, &h2, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    dispatch(c, [&] { Launch<2>::contact_search(c); }, [&] { Launch<3>::contact_search(c); });
    refresh_xa(c);
    check_launch(c);
    check_overflow(c);
  }
  std::vector<double> mx(N);
  HIPCHK(hipMemcpyAsync(mx.data(), s.mx, N * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < N; i++) {
    double m = mx[i];
    double fa = std::fabs(m), fb = 1.0E-18;
    if ((fb - m) > ((fa < fb ? fb : fa) * 1e-5))
      throw HipError{"weighted nodal volume m_x too close to zero at node " + std::to_string(i)};
  }
}

template <class F>
int guarded(pdem_ctx* c, F f) {
  if (!c) return set_err(nullptr, "null context");
  try {
    f();
    return 0;
  } catch (const HipError& e) {
    return set_err(c, e.msg);
  } catch (const std::exception& e) {
    return set_err(c, e.what());
  }
}

void drop_graph(pdem_ctx* c);

void ensure_graph(pdem_ctx* c, bool verlet) {
  if (c->graph_ready && c->graph_verlet == verlet) return;
  drop_graph(c);
  // zero-displacement sets are applied once, before any capture
  bool pending = false;
  for (size_t s = 0; s < c->ubc.size(); s++)
    if (c->ubc_zero[s] && !c->zero_done[s]) pending = true;
  if (pending) throw HipError{"call pdem_integrate_begin before replaying the captured step"};
  bool prof = c->profiling;
  c->profiling = false;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  dispatch(c, [&] { Launch<2>::integrate_step(c, verlet); }, [&] { Launch<3>::integrate_step(c, verlet); });
  HIPCHK(hipStreamEndCapture(c->stream, &c->graph));
  c->profiling = prof;
  HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_ready = true;
  c->graph_verlet = verlet;
}

}  // namespace

// ======================================================================= ABI
extern "C" {

int pdem_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int64_t pdem_abi_sizeof(int32_t which) {
  switch (which) {
    case 0: return (int64_t)sizeof(pdem_desc);
    case 1: return (int64_t)sizeof(pdem_particle);
    case 2: return (int64_t)sizeof(pdem_contact_pair);
    case 3: return (int64_t)sizeof(pdem_bc_set);
    case 4: return (int64_t)sizeof(pdem_contact_state);
    case 5: return (int64_t)sizeof(pdem_material_deck);
    default: return -1;
  }
}

const char* pdem_last_error(const pdem_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int pdem_create(const pdem_desc* d, pdem_ctx** out) {
  if (!out) return set_err(nullptr, "null output pointer");
  *out = nullptr;
  if (!d) return set_err(nullptr, "null descriptor");
  if (d->dim != 2 && d->dim != 3) return set_err(nullptr, "dim must be 2 or 3");
  if (d->n_nodes <= 0 || d->n_nodes >= INT_MAX) return set_err(nullptr, "n_nodes out of range");
  if (d->n_particles <= 0 || !d->particles) return set_err(nullptr, "no particles");
  if (!d->x_ref || !d->vol || !d->fix || !d->pt_id) return set_err(nullptr, "missing node arrays");
  if (d->precision != PDEM_PREC_FP64 && d->precision != PDEM_PREC_MIXED)
    return set_err(nullptr, "unknown precision mode");
  if (d->multi_particle && (!d->contact || d->n_zones <= 0))
    return set_err(nullptr, "multi-particle model without a contact table");
  int64_t expect = 0;
  for (int p = 0; p < d->n_particles; p++) {
    const auto& q = d->particles[p];
    if (q.glob_start != expect || q.glob_end <= q.glob_start || q.glob_end > d->n_nodes)
      return set_err(nullptr, "particle node ranges must be contiguous and ordered");
    if (q.center_node < q.glob_start || q.center_node >= q.glob_end)
      return set_err(nullptr, "center node outside its particle");
    if (q.zone < 0 || (d->multi_particle && q.zone >= d->n_zones)) return set_err(nullptr, "bad zone id");
    if (q.material_type < PDEM_MAT_RNP || q.material_type > PDEM_MAT_PDELASTIC)
      return set_err(nullptr, "material type not supported on the GPU path");
    if (q.glob_end - q.glob_start > INT_MAX / 2) return set_err(nullptr, "particle too large");
    expect = q.glob_end;
  }
  if (expect != d->n_nodes) return set_err(nullptr, "particles do not cover all nodes");
  for (int64_t i = 0; i < d->n_nodes; i++) {
    int64_t p = d->pt_id[i];
    if (p < 0 || p >= d->n_particles || i < d->particles[p].glob_start || i >= d->particles[p].glob_end)
      return set_err(nullptr, "pt_id inconsistent with particle node ranges");
  }
  for (int k = 0; k < d->n_ubc + d->n_fbc; k++) {
    const pdem_bc_set& b = k < d->n_ubc ? d->ubc[k] : d->fbc[k - d->n_ubc];
    for (int64_t e = 0; e < b.n_entries; e++)
      if (b.nodes[e] < 0 || b.nodes[e] >= d->n_nodes) return set_err(nullptr, "BC node out of range");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(nullptr, "no HIP device");
  auto* c = new pdem_ctx();
  try {
    build(c, d);
  } catch (const HipError& e) {
    g_create_err = e.msg;
    delete c;
    return 1;
  }
  *out = c;
  return 0;
}

void pdem_destroy(pdem_ctx* c) { delete c; }

int pdem_integrate_begin(pdem_ctx* c) {
  return guarded(c, [&] {
    dispatch(c, [&] { Launch<2>::displacement_bc(c); }, [&] { Launch<3>::displacement_bc(c); });
    dispatch(c, [&] { Launch<2>::compute_forces(c); }, [&] { Launch<3>::compute_forces(c); });
    check_launch(c);
    if (c->multi) check_overflow(c);
  });
}

int pdem_integrate_cd(pdem_ctx* c, int64_t nsteps) {
  return guarded(c, [&] {
    for (int64_t k = 0; k < nsteps; k++)
      dispatch(c, [&] { Launch<2>::integrate_step(c); }, [&] { Launch<3>::integrate_step(c); });
    check_launch(c);
    if (c->multi) check_overflow(c);
  });
}

int pdem_integrate_cd_graph(pdem_ctx* c, int64_t nsteps) {
  return guarded(c, [&] {
    ensure_graph(c, false);
    for (int64_t k = 0; k < nsteps; k++) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    if (c->multi) check_overflow(c);
  });
}

int pdem_integrate_verlet(pdem_ctx* c, int64_t nsteps) {
  return guarded(c, [&] {
    for (int64_t k = 0; k < nsteps; k++)
      dispatch(c, [&] { Launch<2>::integrate_step(c, true); }, [&] { Launch<3>::integrate_step(c, true); });
    check_launch(c);
    if (c->multi) check_overflow(c);
  });
}

int pdem_integrate_verlet_graph(pdem_ctx* c, int64_t nsteps) {
  return guarded(c, [&] {
    ensure_graph(c, true);
    for (int64_t k = 0; k < nsteps; k++) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    if (c->multi) check_overflow(c);
  });
}

int pdem_compute_forces(pdem_ctx* c) {
  return guarded(c, [&] {
    dispatch(c, [&] { Launch<2>::compute_forces(c); }, [&] { Launch<3>::compute_forces(c); });
    check_launch(c);
    if (c->multi) check_overflow(c);
  });
}

int pdem_compute_peridynamic_forces(pdem_ctx* c) {
  return guarded(c, [&] {
    // the PD passes set f of the nodes that take no PD force to zero (the
    // force reset of a step); computePeridynamicForces alone leaves their f
    // as it is: it is saved, final (pending terms applied), and put back
    finalize_pending(c);
    const int64_t nnf = c->ds.n_noforce;
    if (nnf > 0 && !c->d_nfsave) c->d_nfsave = c->dalloc<double>(3 * nnf, false);
    const unsigned nb = (unsigned)std::min<int64_t>(blocks(std::max<int64_t>(nnf, 1)), BZ_BLOCKS);
    auto copy = [&](int save) {
      if (nnf == 0) return;
      dispatch(
          c, [&] { hipLaunchKernelGGL(k_noforce_copy<2>, dim3(nb), dim3(TPB), 0, c->stream, c->ds, c->d_nfsave, save); },
          [&] { hipLaunchKernelGGL(k_noforce_copy<3>, dim3(nb), dim3(TPB), 0, c->stream, c->ds, c->d_nfsave, save); });
    };
    copy(1);
    dispatch(c, [&] { Launch<2>::pd_forces(c); }, [&] { Launch<3>::pd_forces(c); });
    copy(0);
    check_launch(c);
  });
}

int pdem_update_contact_neighborlist(pdem_ctx* c) {
  return guarded(c, [&] {
    if (!c->multi) return;
    finalize_pending(c);  // the controller kernel below clears the pending flag
    dispatch(c, [&] { Launch<2>::contact_search(c); }, [&] { Launch<3>::contact_search(c); });
    check_launch(c);
    check_overflow(c);
  });
}

int pdem_compute_contact_forces(pdem_ctx* c) {
  return guarded(c, [&] {
    if (!c->multi) return;
    finalize_pending(c);
    dispatch(c, [&] { Launch<2>::contact_forces(c, true); }, [&] { Launch<3>::contact_forces(c, true); });
    check_launch(c);
  });
}

int pdem_compute_external_forces(pdem_ctx* c) {
  return guarded(c, [&] {
    finalize_pending(c);
    dispatch(c, [&] { Launch<2>::external_forces(c, false); }, [&] { Launch<3>::external_forces(c, false); });
    check_launch(c);
  });
}

int pdem_compute_external_displacement_bc(pdem_ctx* c) {
  return guarded(c, [&] {
    dispatch(c, [&] { Launch<2>::displacement_bc(c); }, [&] { Launch<3>::displacement_bc(c); });
    // moved outside a step: contact calls may follow before any controller
    // step adds the move to the contact skin bound
    hipLaunchKernelGGL(k_skin_invalidate, dim3(1), dim3(1), 0, c->stream, c->ds);
    check_launch(c);
  });
}

int pdem_synchronize(pdem_ctx* c) {
  return guarded(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

int pdem_get_field(pdem_ctx* c, int32_t field, void* host) {
  return guarded(c, [&] {
    const int64_t N = c->N;
    DS& s = c->ds;
    if (field == PDEM_FIELD_F) finalize_pending(c);
    auto getp = [&](double* const* comp) {
      std::vector<double> t(N);
      double* o = (double*)host;
      for (int k = 0; k < 3; k++) {
        if (k < c->dim) {
          HIPCHK(hipMemcpyAsync(t.data(), comp[k], N * 8, hipMemcpyDeviceToHost, c->stream));
          HIPCHK(hipStreamSynchronize(c->stream));
        } else {
          std::fill(t.begin(), t.end(), 0.);
        }
        for (int64_t i = 0; i < N; i++) o[3 * i + k] = t[i];
      }
    };
    auto get1 = [&](const void* d, size_t bytes) {
      HIPCHK(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    };
    switch (field) {
      case PDEM_FIELD_X: getp(s.x); break;
      case PDEM_FIELD_U: getp(s.u); break;
      case PDEM_FIELD_V: getp(s.v); break;
      case PDEM_FIELD_F: getp(s.f); break;
      case PDEM_FIELD_XREF: getp(s.xr); break;
      case PDEM_FIELD_VOL: get1(s.vol, N * 8); break;
      case PDEM_FIELD_Z: get1(s.Z, N * 4); break;
      case PDEM_FIELD_THETA: get1(s.theta, N * 8); break;
      case PDEM_FIELD_MX: get1(s.mx, N * 8); break;
      case PDEM_FIELD_VMAG: get1(s.vmag, N * 8); break;
      case PDEM_FIELD_BONDS: {
        std::vector<unsigned> w(c->nbitwords);
        HIPCHK(hipMemcpyAsync(w.data(), s.bits, w.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        uint8_t* o = (uint8_t*)host;
        int64_t pos = 0;
        for (int64_t i = 0; i < N; i++) {
          int64_t sl = i / WAVE, lane = i % WAVE;
          for (int k = 0; k < c->h_deg[i]; k++)
            o[pos++] = (w[c->h_bptr[sl] + (k >> 5) * WAVE + lane] >> (k & 31)) & 1u;
        }
        break;
      }
      default: throw HipError{"unknown field id"};
    }
  });
}

int pdem_set_field(pdem_ctx* c, int32_t field, const void* host) {
  return guarded(c, [&] {
    const int64_t N = c->N;
    DS& s = c->ds;
    auto setp = [&](double* const* comp) {
      std::vector<double> t(N);
      const double* o = (const double*)host;
      for (int k = 0; k < c->dim; k++) {
        for (int64_t i = 0; i < N; i++) t[i] = o[3 * i + k];
        HIPCHK(hipMemcpyAsync(comp[k], t.data(), N * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
    };
    switch (field) {
      case PDEM_FIELD_X:
        setp(s.x);
        refresh_xa(c);
        HIPCHK(hipStreamSynchronize(c->stream));
        break;
      case PDEM_FIELD_U: setp(s.u); break;
      case PDEM_FIELD_V: setp(s.v); break;
      case PDEM_FIELD_F:
        setp(s.f);
        clear_pending(c);
        HIPCHK(hipStreamSynchronize(c->stream));
        break;
      case PDEM_FIELD_THETA:
        HIPCHK(hipMemcpyAsync(s.theta, host, N * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        break;
      case PDEM_FIELD_BONDS: {
        std::vector<unsigned> w(c->nbitwords, 0u);
        const uint8_t* o = (const uint8_t*)host;
        int64_t pos = 0;
        for (int64_t i = 0; i < N; i++) {
          int64_t sl = i / WAVE, lane = i % WAVE;
          for (int k = 0; k < c->h_deg[i]; k++)
            if (o[pos++]) w[c->h_bptr[sl] + (k >> 5) * WAVE + lane] |= 1u << (k & 31);
        }
        HIPCHK(hipMemcpyAsync(s.bits, w.data(), w.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        break;
      }
      default: throw HipError{"field is read-only"};
    }
  });
}

int pdem_get_time(pdem_ctx* c, int64_t* n, double* t) {
  return guarded(c, [&] {
    Ctrl h = read_ctrl(c);
    *n = h.n;
    *t = h.time;
  });
}

int pdem_get_neighbor_count(pdem_ctx* c, int32_t kind, int64_t* nnz) {
  return guarded(c, [&] {
    if (kind == PDEM_NEIGH_PD) {
      *nnz = c->nnz;
      return;
    }
    if (!c->multi) {
      *nnz = 0;
      return;
    }
    std::vector<int> cnt(c->N);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->ds.ccount, c->N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int64_t sum = 0;
    for (int v : cnt) sum += v;
    *nnz = sum;
  });
}

int pdem_get_neighbors(pdem_ctx* c, int32_t kind, int64_t* rowptr, int64_t* cols) {
  return guarded(c, [&] {
    const int64_t N = c->N;
    if (kind == PDEM_NEIGH_PD) {
      std::vector<int> col(c->nnz_padded);
      HIPCHK(hipMemcpyAsync(col.data(), c->ds.col, col.size() * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      rowptr[0] = 0;
      for (int64_t i = 0; i < N; i++) {
        int64_t sl = i / WAVE, lane = i % WAVE;
        for (int k = 0; k < c->h_deg[i]; k++) cols[rowptr[i] + k] = col[c->h_sptr[sl] + (int64_t)k * WAVE + lane];
        rowptr[i + 1] = rowptr[i] + c->h_deg[i];
      }
      return;
    }
    rowptr[0] = 0;
    if (!c->multi) {
      for (int64_t i = 0; i < N; i++) rowptr[i + 1] = 0;
      return;
    }
    std::vector<int> cnt(N);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->ds.ccount, N * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<int> cand((size_t)c->cap * N);
    HIPCHK(hipMemcpyAsync(cand.data(), c->ds.cand, cand.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < N; i++) {
      for (int k = 0; k < cnt[i]; k++) cols[rowptr[i] + k] = cand[(size_t)cand_idx(N, i, k)];
      rowptr[i + 1] = rowptr[i] + cnt[i];
    }
  });
}

int pdem_get_contact_state(pdem_ctx* c, pdem_contact_state* st) {
  return guarded(c, [&] {
    Ctrl h = read_ctrl(c);
    st->interval = h.interval;
    st->counter = h.counter;
    st->search_radius = h.radius;
    st->max_velocity = h.max_velocity;
    st->max_contact_r = h.max_contact_r;
    st->n_rebuilds = h.n_rebuilds;
    st->overflow = h.overflow;
    st->capacity = c->cap;
  });
}

int pdem_pp_compressive_test(pdem_ctx* c, int32_t wall_id, int32_t f_dir, double* pen, double* reac) {
  return guarded(c, [&] {
    if (wall_id < 0 || wall_id >= c->np || f_dir < 0 || f_dir >= c->dim)
      throw HipError{"bad wall id or direction"};
    finalize_pending(c);
    dispatch(
        c, [&] { hipLaunchKernelGGL(k_pp_compressive<2>, dim3(1), dim3(64), 0, c->stream, c->ds, wall_id, f_dir); },
        [&] { hipLaunchKernelGGL(k_pp_compressive<3>, dim3(1), dim3(64), 0, c->stream, c->ds, wall_id, f_dir); });
    check_launch(c);
    Ctrl h = read_ctrl(c);
    *pen = h.pp_pen;
    *reac = h.pp_reac;
  });
}

int pdem_pp_two_particle_test(pdem_ctx* c, double* out5) {
  return guarded(c, [&] {
    if (c->np < 2 || !c->multi) throw HipError{"the two-particle test needs a multi-particle model of two particles or more"};
    if (!out5) throw HipError{"null output"};
    if (!c->pp_out) c->pp_out = c->dalloc<double>(8);
    hipLaunchKernelGGL(k_pp_two_particle, dim3(1), dim3(64), 0, c->stream, c->ds, c->pp_out);
    check_launch(c);
    HIPCHK(hipMemcpyAsync(out5, c->pp_out, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

// util::isGreater (function.cpp:15-23) on the host
static bool host_is_greater(double a, double b) {
  const double fa = std::fabs(a), fb = std::fabs(b);
  return (a - b) > ((fa < fb ? fb : fa) * COMPARE_EPS);
}

int pdem_check_stop(pdem_ctx* c, int32_t criterion, double param, int32_t* stop, double* value, int64_t* node) {
  return guarded(c, [&] {
    if (!stop || !value || !node) throw HipError{"null output"};
    if (!c->pp_out) c->pp_out = c->dalloc<double>(8);
    *node = -1;
    if (criterion == PDEM_STOP_MAX_PARTICLE_DIST) {
      if (c->np < 2) throw HipError{"max_particle_dist needs two particles"};
      hipLaunchKernelGGL(k_stop_center_dist, dim3(1), dim3(64), 0, c->stream, c->ds, c->pp_out);
      check_launch(c);
      HIPCHK(hipMemcpyAsync(value, c->pp_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    } else if (criterion == PDEM_STOP_MAX_NODE_DIST) {
      unsigned long long init[2] = {0ull, ~0ull};
      auto* u = (unsigned long long*)c->pp_out;
      HIPCHK(hipMemcpyAsync(u, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(k_stop_maxlen, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds, u);
      hipLaunchKernelGGL(k_stop_argmax, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds, u, u + 1);
      check_launch(c);
      unsigned long long h[2];
      HIPCHK(hipMemcpyAsync(h, u, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      std::memcpy(value, &h[0], sizeof(double));
      *node = (int64_t)h[1];
    } else {
      throw HipError{"unknown stop criterion"};
    }
    *stop = host_is_greater(*value, param) ? 1 : 0;
  });
}

int pdem_restart(pdem_ctx* c, int64_t step) {
  return guarded(c, [&] {
    if (step < 0) throw HipError{"negative restart step"};
    Ctrl h = read_ctrl(c);
    h.n = step;
    h.time = double(step) * h.dt;
    h.restart_active = 1;
    h.restart_step = step;
    HIPCHK(hipMemcpyAsync(c->ds.ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int pdem_broken_bonds(pdem_ctx* c, int64_t* broken, int64_t* total) {
  return guarded(c, [&] {
    unsigned long long zero = 0;
    HIPCHK(hipMemcpyAsync(&c->ds.ctrl->broken, &zero, 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_count_broken, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, c->ds);
    check_launch(c);
    Ctrl h = read_ctrl(c);
    *broken = (int64_t)h.broken;
    *total = c->nnz;
  });
}

// Reference-element quadrature points of TriElem::init (triElem.cpp:192-417)
// and TetElem::init (tetElem.cpp:300-409); the shape values at a point are
// N = {1 - xi - eta [- zeta], xi, eta[, zeta]} (triElem.cpp:119-123,
// tetElem.cpp:180-184).
static std::vector<std::array<double, 3>> fe_points(int dim, int order) {
  using P = std::array<double, 3>;
  if (dim == 2) {
    switch (order) {
      case 1: return {P{1. / 3., 1. / 3., 0.}};
      case 2: return {P{1. / 6., 1. / 6., 0.}, P{2. / 3., 1. / 6., 0.}, P{1. / 6., 2. / 3., 0.}};
      case 3:
        return {P{1. / 3., 1. / 3., 0.}, P{1. / 5., 3. / 5., 0.}, P{1. / 5., 1. / 5., 0.}, P{3. / 5., 1. / 5., 0.}};
      case 4: {
        const double a = 0.44594849091597, b = 0.10810301816807, c = 0.09157621350977, d = 0.81684757298046;
        return {P{a, a, 0.}, P{a, b, 0.}, P{b, a, 0.}, P{c, c, 0.}, P{c, d, 0.}, P{d, c, 0.}};
      }
      case 5: {
        const double a = 0.33333333333333, b = 0.47014206410511, c = 0.05971587178977, d = 0.10128650732346,
                     f = 0.79742698535309;
        return {P{a, a, 0.}, P{b, b, 0.}, P{b, c, 0.}, P{c, b, 0.}, P{d, d, 0.}, P{d, f, 0.}, P{f, d, 0.}};
      }
    }
  } else {
    switch (order) {
      case 1: return {P{1. / 4., 1. / 4., 1. / 4.}};
      case 2: {
        const double a = 0.585410196624969, b = 0.138196601125011;
        return {P{a, b, b}, P{b, a, b}, P{b, b, a}, P{b, b, b}};
      }
      case 3: {
        const double a = 0.25, b = 0.5, c = 1. / 6.;
        return {P{a, a, a}, P{b, c, c}, P{c, b, c}, P{c, c, b}, P{c, c, c}};
      }
    }
  }
  throw HipError{"quadrature order must be 1-5 (triangles) or 1-3 (tetrahedra)"};
}

int pdem_fe_setup(pdem_ctx* c, int32_t quad_order, int64_t n_elem, const int64_t* elem_nodes,
                  const int32_t* elem_particle, const double* elastic) {
  return guarded(c, [&] {
    if (n_elem <= 0 || !elem_nodes || !elem_particle || !elastic) throw HipError{"empty element list"};
    const int dim = c->dim, nv = dim + 1;
    const auto pts = fe_points(dim, quad_order);
    const int nq = (int)pts.size();
    if (n_elem > (int64_t)1 << 40) throw HipError{"too many elements"};
    for (int64_t e = 0; e < n_elem; e++) {
      const int32_t p = elem_particle[e];
      if (p < 0 || p >= c->np) throw HipError{"element particle id out of range"};
      for (int k = 0; k < nv; k++) {
        const int64_t g = elem_nodes[e * nv + k];
        if (g < c->hp[p].glob_start || g >= c->hp[p].glob_end)
          throw HipError{"element node outside its particle's node range"};
      }
    }
    std::vector<double> qpt(3 * nq), qsh((size_t)nv * nq);
    for (int q = 0; q < nq; q++) {
      const auto& x = pts[q];
      for (int d = 0; d < 3; d++) qpt[3 * q + d] = x[d];
      if (dim == 2) {
        qsh[3 * q + 0] = 1. - x[0] - x[1];
        qsh[3 * q + 1] = x[0];
        qsh[3 * q + 2] = x[1];
      } else {
        qsh[4 * q + 0] = 1. - x[0] - x[1] - x[2];
        qsh[4 * q + 1] = x[0];
        qsh[4 * q + 2] = x[1];
        qsh[4 * q + 3] = x[2];
      }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void* b : c->fe_bufs) c->dfree(b);
    c->fe_bufs.clear();
    auto keep = [&](auto* ptr) {
      c->fe_bufs.push_back((void*)ptr);
      return ptr;
    };
    FeArgs a{};
    a.n_elem = n_elem;
    a.nq = nq;
    a.nodes = keep(c->upload(elem_nodes, (size_t)n_elem * nv));
    a.part = keep(c->upload((const int*)elem_particle, (size_t)n_elem));
    a.elastic = keep(c->upload(elastic, (size_t)4 * c->np));
    a.qpt = keep(c->upload(qpt.data(), qpt.size()));
    a.qshape = keep(c->upload(qsh.data(), qsh.size()));
    a.xq = keep(c->dalloc<double>((size_t)n_elem * nq * 3));
    a.strain = keep(c->dalloc<float>((size_t)n_elem * nq * 6));
    a.stress = keep(c->dalloc<float>((size_t)n_elem * nq * 6));
    c->fe = a;
  });
}

int pdem_fe_quad_count(pdem_ctx* c, int64_t* nq) {
  return guarded(c, [&] { *nq = c->fe.n_elem * c->fe.nq; });
}

int pdem_strain_stress(pdem_ctx* c, double* xq, float* strain, float* stress) {
  return guarded(c, [&] {
    const FeArgs& a = c->fe;
    if (a.n_elem == 0) throw HipError{"call pdem_fe_setup first"};
    const unsigned nb = blocks(a.n_elem);
    dispatch(
        c, [&] { hipLaunchKernelGGL(k_strain_stress<2>, dim3(nb), dim3(TPB), 0, c->stream, c->ds, a); },
        [&] { hipLaunchKernelGGL(k_strain_stress<3>, dim3(nb), dim3(TPB), 0, c->stream, c->ds, a); });
    check_launch(c);
    const size_t nq = (size_t)a.n_elem * a.nq;
    if (xq) HIPCHK(hipMemcpyAsync(xq, a.xq, nq * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (strain) HIPCHK(hipMemcpyAsync(strain, a.strain, nq * 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (stress) HIPCHK(hipMemcpyAsync(stress, a.stress, nq * 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int pdem_profile_enable(pdem_ctx* c, int32_t on) {
  return guarded(c, [&] {
    c->profiling = on != 0;
    c->prof_ms = 0.;
    c->prof_n = 0;
  });
}

int pdem_profile_get(pdem_ctx* c, double* ms, int64_t* n) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->side) HIPCHK(hipStreamSynchronize(c->side));
    // a region's time is the longer of its pairs (they run side by side)
    for (size_t k = 0; k < c->ev_used.size();) {
      float t = 0.f;
      size_t m = k;
      for (; m < c->ev_used.size() && c->ev_group[m] == c->ev_group[k]; m++) {
        float tm = 0.f;
        HIPCHK(hipEventElapsedTime(&tm, c->ev_used[m].first, c->ev_used[m].second));
        t = std::max(t, tm);
        c->ev_free.push_back(c->ev_used[m]);
      }
      c->prof_ms += t;
      c->prof_n++;
      k = m;
    }
    c->ev_used.clear();
    c->ev_group.clear();
    *ms = c->prof_n ? c->prof_ms / (double)c->prof_n : 0.;
    *n = c->prof_n;
  });
}

}  // extern "C"

namespace {

void drop_graph(pdem_ctx* c) {
  if (!c->graph_ready) return;
  HIPCHK(hipGraphExecDestroy(c->gexec));
  HIPCHK(hipGraphDestroy(c->graph));
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_ready = false;
}

// (re)build the closure pair list and wall flags from the current candidate
// and partner lists
void closure_lists(pdem_ctx* c) {
  DS& s = c->ds;
  const int nn = s.n_nonwall;
  HIPCHK(hipMemsetAsync(s.wall_near, 0, (size_t)nn * sizeof(int), c->stream));
  const uint32_t zero = 0;
  HIPCHK(hipMemcpyAsync(&s.ctrl->npairs, &zero, 4, hipMemcpyHostToDevice, c->stream));
  DS t = s;
  t.cl_on = 1;
  HIPCHK(hipMemsetAsync(s.cmode, 0, (size_t)nn * s.dcap, c->stream));
  HIPCHK(hipMemsetAsync(s.cconn, 0xff, (size_t)nn * s.dcap * sizeof(int), c->stream));
  HIPCHK(hipMemcpyAsync(&s.ctrl->nconn, &zero, 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_cl_wall_near, dim3(blocks(c->N)), dim3(TPB), 0, c->stream, t);
  hipLaunchKernelGGL(k_cl_pairs, dim3(blocks(nn)), dim3(TPB), 0, c->stream, t);
  hipLaunchKernelGGL(k_cl_connect, dim3((unsigned)std::min<int64_t>(blocks(c->N), CL_CONNECT_BLOCKS)), dim3(TPB), 0,
                     c->stream, t, 1);
  check_launch(c);
  check_overflow(c);
}

// surface nodes, orientation reference nodes, mean nodal volumes, codes and
// the closure's device buffers
void closure_setup(pdem_ctx* c, const float* particle_code) {
  if (!c->multi) throw HipError{"the closure needs a multi-particle model"};
  if (c->dim != 2) throw HipError{"the closure is implemented for 2D models"};
  DS& s = c->ds;
  const int64_t N = c->N;
  std::vector<int> nonwall;
  for (int p = 0; p < c->np; p++)
    if (c->hp[p].type_index == 0) nonwall.push_back(p);
  const int nn = (int)nonwall.size();
  if (nn == 0) throw HipError{"the closure needs non-wall particles"};
  std::vector<float> code(nn, 0.f);
  for (int q = 0; q < nn; q++) code[q] = particle_code ? particle_code[nonwall[q]] : 0.f;
  HIPCHK(hipStreamSynchronize(c->stream));  // steps in flight may read the codes
  if (c->cl_ready) {
    HIPCHK(hipMemcpy((void*)c->ca.code, code.data(), nn * sizeof(float), hipMemcpyHostToDevice));
    return;
  }
  std::vector<double> xr0(N), xr1(N), vol(N);
  HIPCHK(hipMemcpy(xr0.data(), s.xr[0], N * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(xr1.data(), s.xr[1], N * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vol.data(), s.vol, N * 8, hipMemcpyDeviceToHost));
  std::vector<uint8_t> surf(N, 0);
  std::vector<int> sptr(nn + 1, 0), sidx, ref(nn);
  std::vector<double> vbar(nn);
  for (int q = 0; q < nn; q++) {
    const auto& P = c->hp[nonwall[q]];
    int dmax = 0;
    for (int64_t i = P.glob_start; i < P.glob_end; i++) dmax = std::max(dmax, c->h_deg[i]);
    // surface: nodes whose horizon is truncated by the boundary
    double far = -1., vs = 0.;
    int64_t fi = P.center_node;
    for (int64_t i = P.glob_start; i < P.glob_end; i++) {
      if (4 * c->h_deg[i] < 3 * dmax) {
        surf[i] = 1;
        sidx.push_back((int)i);
      }
      const double dx = xr0[i] - xr0[P.center_node], dy = xr1[i] - xr1[P.center_node];
      if (dx * dx + dy * dy > far) {
        far = dx * dx + dy * dy;
        fi = i;
      }
      vs += vol[i];
    }
    sptr[q + 1] = (int)sidx.size();
    ref[q] = (int)fi;
    vbar[q] = vs / (double)(P.glob_end - P.glob_start);
  }
  if (sidx.empty()) sidx.push_back(0);
  s.surf = c->upload(surf.data(), N);
  s.surf_ptr = c->upload(sptr.data(), sptr.size());
  s.surf_idx = c->upload(sidx.data(), sidx.size());
  c->ca.ref_node = c->upload(ref.data(), nn);
  c->ca.vbar = c->upload(vbar.data(), nn);
  c->ca.code = c->upload(code.data(), nn);
  s.cmode = c->dalloc<int8_t>((size_t)nn * s.dcap);
  s.cout = c->dalloc<double>((size_t)nn * s.dcap * CL_REC);
  s.pflags = c->dalloc<int>(nn);
  s.wall_near = c->dalloc<int>(nn);
  s.cconn = c->dalloc<int>((size_t)nn * s.dcap, false);
  HIPCHK(hipMemsetAsync(s.cconn, 0xff, (size_t)nn * s.dcap * sizeof(int), c->stream));  // -1: never connected
  s.cpair_cap = (int)std::min<int64_t>((int64_t)nn * 16, (int64_t)nn * s.dcap);
  s.cpair_cap = std::max(s.cpair_cap, 1);
  s.cpairs = c->dalloc<int>(2 * (size_t)s.cpair_cap);
  s.conn_list = c->dalloc<int>((size_t)s.cpair_cap);
  s.cl_pend = c->dalloc<double>((size_t)s.cpair_cap * CL_PEND, false);
  c->ca.stats = c->dalloc<unsigned long long>(8);
  c->ca.err_sum = c->dalloc<double>(2);
  c->persist.push_back({s.cpairs, 2 * (size_t)s.cpair_cap * 4});
  c->persist.push_back({s.wall_near, (size_t)nn * 4});
  c->persist.push_back({s.cmode, (size_t)nn * s.dcap});
  c->persist.push_back({s.cconn, (size_t)nn * s.dcap * 4});
  c->persist.push_back({s.conn_list, (size_t)s.cpair_cap * 4});
  c->cl_ready = true;
}

}  // namespace

extern "C" {

int pdem_closure_setup(pdem_ctx* c, const float* particle_code) {
  return guarded(c, [&] { closure_setup(c, particle_code); });
}

int pdem_closure_enable(pdem_ctx* c, const pdem_mlp* net, double pen_max_h, double dz_max, int32_t check_every) {
  return guarded(c, [&] {
    if (!c->cl_ready) closure_setup(c, nullptr);
    if (net) {
      int dev = 0;
      HIPCHK(hipGetDevice(&dev));
      if (dev != net->device) throw HipError{"closure network lives on another device"};
    }
    if (check_every < 0) throw HipError{"check_every must be >= 0"};
    finalize_pending(c);  // the closure's transfer runs in the finalisation of later steps
    HIPCHK(hipStreamSynchronize(c->stream));  // steps in flight still read the previous network
    c->ca.use_net = (net && pen_max_h > 0.) ? 1 : 0;
    c->ca.net = net ? mlp::make_net(net->wb, net->wtb, net->p, net->norm) : mlp::Net{};
    c->ca.pen_max_h = pen_max_h;
    c->ca.dz_max = (float)dz_max;
    c->ca.check_every = check_every;
    if (net != c->cl_net) {
      if (net) net->users.fetch_add(1);
      if (c->cl_net) c->cl_net->users.fetch_sub(1);
      c->cl_net = net;
    }
    if (net) HIPCHK(hipStreamSynchronize(net->stream));  // training finished before use
    HIPCHK(hipMemsetAsync(c->ca.stats, 0, 8 * sizeof(unsigned long long), c->stream));
    HIPCHK(hipMemsetAsync(c->ca.err_sum, 0, 2 * sizeof(double), c->stream));
    closure_lists(c);
    c->ds.cl_on = 1;
    drop_graph(c);
  });
}

int pdem_closure_disable(pdem_ctx* c) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));  // steps in flight still read the network
    c->ds.cl_on = 0;
    drop_graph(c);
    if (c->cl_net) c->cl_net->users.fetch_sub(1);
    c->cl_net = nullptr;
    c->ca.net = mlp::Net{};
    c->ca.use_net = 0;
  });
}

int pdem_closure_stats(pdem_ctx* c, int64_t* pairs_by_mode, int64_t* checked, double* err_mean,
                       double* err_max, double* plain_mean) {
  return guarded(c, [&] {
    if (!c->cl_ready) throw HipError{"closure not set up"};
    unsigned long long st[8];
    double es[2];
    HIPCHK(hipMemcpyAsync(st, c->ca.stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(es, c->ca.err_sum, sizeof(es), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; k++) pairs_by_mode[k] = (int64_t)st[k];
    *checked = (int64_t)st[3];
    *err_mean = st[3] ? es[0] / (double)st[3] : 0.;
    *plain_mean = st[3] ? es[1] / (double)st[3] : 0.;
    double mx;
    std::memcpy(&mx, &st[4], 8);
    *err_max = mx;
  });
}

}  // extern "C"

namespace {
// The pair-log launches and copies of pdem_pair_log and
// pdem_pair_log_noalloc: rows staged in df / dt / db (room for at least cap
// rows), the row counter in dn (cleared here by zero_count, which enqueues on the
// engine stream); rows beyond cap are dropped, *n gets the number found.
template <class Zero>
void pair_log_staged(pdem_ctx* c, double band_h, int64_t cap, float* df, float* dt, float* db, unsigned* dn,
                     Zero zero_count, float* feat, float* targ, float* base, int64_t* n) {
  closure_lists(c);
  zero_count();
  DS t = c->ds;
  t.cl_on = 1;
  hipLaunchKernelGGL(k_pair_log, dim3(closure_blocks(c)), dim3(CL_WAVES * WAVE), 0, c->stream, t, c->ca, band_h, df,
                     dt, db, dn,
                     (int)std::min<int64_t>(cap, INT_MAX));
  check_launch(c);
  unsigned cnt = 0;
  HIPCHK(hipMemcpyAsync(&cnt, dn, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t m = std::min<int64_t>(cnt, cap);
  if (m > 0) {
    HIPCHK(hipMemcpy(feat, df, m * CL_NFEAT * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(targ, dt, m * 4 * sizeof(float), hipMemcpyDeviceToHost));
    if (base) HIPCHK(hipMemcpy(base, db, m * 3 * sizeof(float), hipMemcpyDeviceToHost));
  }
  *n = (int64_t)cnt;
}
}  // namespace

extern "C" {

int pdem_pair_log(pdem_ctx* c, double band_h, int64_t cap, float* feat, float* targ, float* base, int64_t* n) {
  return guarded(c, [&] {
    if (!c->cl_ready) throw HipError{"call pdem_closure_setup before pdem_pair_log"};
    if (cap < 0) throw HipError{"negative capacity"};
    float* df = c->dalloc<float>((size_t)std::max<int64_t>(cap, 1) * CL_NFEAT);
    float* dt = c->dalloc<float>((size_t)std::max<int64_t>(cap, 1) * 4);
    float* db = c->dalloc<float>((size_t)std::max<int64_t>(cap, 1) * 3);
    unsigned* dn = c->dalloc<unsigned>(1, false);
    pair_log_staged(
        c, band_h, cap, df, dt, db, dn,
        [&] { HIPCHK(hipMemsetAsync(dn, 0, sizeof(unsigned), c->stream)); }, feat, targ, base, n);
    c->dfree(df);
    c->dfree(dt);
    c->dfree(db);
    c->dfree(dn);
  });
}

int pdem_pair_log_noalloc(pdem_ctx* c, double band_h, int64_t cap, float* feat, float* targ, float* base,
                          int64_t* n) {
  return guarded(c, [&] {
    if (!c->cl_ready) throw HipError{"call pdem_closure_setup before pdem_pair_log_noalloc"};
    if (cap < 0) throw HipError{"negative capacity"};
    if (cap > NOALLOC_ROWS)
      throw HipError{"pdem_pair_log_noalloc: cap " + std::to_string(cap) + " exceeds the staging maximum of " +
                     std::to_string(NOALLOC_ROWS) + " rows"};
    if (!c->scratch_dev) throw HipError{"staging array not set up"};
    unsigned char* sp = (unsigned char*)c->scratch_dev;
    unsigned* dn = (unsigned*)sp;
    float* df = (float*)(sp + 256);
    float* dt = df + NOALLOC_ROWS * CL_NFEAT;
    float* db = dt + NOALLOC_ROWS * 4;
    // the counter keeps the previous call's value: cleared by a fill kernel
    pair_log_staged(
        c, band_h, cap, df, dt, db, dn,
        [&] { hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, dn, 1); }, feat, targ, base, n);
  });
}

// Checkpoint layout: "PDEMCKP1", int64 N, int32 dim, int32 number of
// blocks, int64 number of zero-set flags, int64 size of every block, the
// zero-set flags (one byte each), then the blocks.
static int64_t ckp_bytes(const pdem_ctx* c) {
  int64_t b = 8 + 8 + 4 + 4 + 8 + 8 * (int64_t)c->persist.size() + (int64_t)c->zero_done.size();
  for (auto& q : c->persist) b += (int64_t)q.second;
  return b;
}

int pdem_checkpoint_size(pdem_ctx* c, int64_t* bytes) {
  return guarded(c, [&] { *bytes = ckp_bytes(c); });
}

int pdem_checkpoint_save(pdem_ctx* c, void* buf, int64_t bytes) {
  return guarded(c, [&] {
    if (bytes < ckp_bytes(c)) throw HipError{"checkpoint buffer too small"};
    finalize_pending(c);  // the checkpoint holds final forces
    char* o = (char*)buf;
    std::memcpy(o, "PDEMCKP1", 8);
    o += 8;
    const int64_t N = c->N, nz = (int64_t)c->zero_done.size();
    const int32_t dim = c->dim, nb = (int32_t)c->persist.size();
    std::memcpy(o, &N, 8);
    std::memcpy(o + 8, &dim, 4);
    std::memcpy(o + 12, &nb, 4);
    std::memcpy(o + 16, &nz, 8);
    o += 24;
    for (auto& q : c->persist) {
      const int64_t sz = (int64_t)q.second;
      std::memcpy(o, &sz, 8);
      o += 8;
    }
    for (char z : c->zero_done) *o++ = z;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& q : c->persist) {
      HIPCHK(hipMemcpy(o, q.first, q.second, hipMemcpyDeviceToHost));
      o += q.second;
    }
  });
}

int pdem_checkpoint_load(pdem_ctx* c, const void* buf, int64_t bytes) {
  return guarded(c, [&] {
    if (bytes < 32 || std::memcmp(buf, "PDEMCKP1", 8) != 0) throw HipError{"not a checkpoint"};
    const char* o = (const char*)buf + 8;
    int64_t N, nz;
    int32_t dim, nb;
    std::memcpy(&N, o, 8);
    std::memcpy(&dim, o + 8, 4);
    std::memcpy(&nb, o + 12, 4);
    std::memcpy(&nz, o + 16, 8);
    o += 24;
    if (N != c->N || dim != c->dim || nb != (int32_t)c->persist.size() || nz != (int64_t)c->zero_done.size() ||
        bytes != ckp_bytes(c))
      throw HipError{"checkpoint does not match this model"};
    for (auto& q : c->persist) {
      int64_t sz;
      std::memcpy(&sz, o, 8);
      if (sz != (int64_t)q.second) throw HipError{"checkpoint does not match this model"};
      o += 8;
    }
    for (auto& z : c->zero_done) z = *o++;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& q : c->persist) {
      HIPCHK(hipMemcpy(q.first, o, q.second, hipMemcpyHostToDevice));
      o += q.second;
    }
    refresh_xa(c);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int pdem_autotune(pdem_ctx* c, int32_t reps, double* best_ms) {
  return guarded(c, [&] {
    if (reps < 1) throw HipError{"reps must be positive"};
    // Time one peridynamic force pass (idempotent: same positions, same
    // bond bits) per candidate launch shape and keep the fastest.
    auto time_pass = [&]() {
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      auto run = [&] { dispatch(c, [&] { Launch<2>::pd_forces(c); }, [&] { Launch<3>::pd_forces(c); }); };
      run();  // warm-up
      HIPCHK(hipEventRecord(a, c->stream));
      for (int r = 0; r < reps; r++) run();
      HIPCHK(hipEventRecord(b, c->stream));
      HIPCHK(hipEventSynchronize(b));
      check_launch(c);
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      HIPCHK(hipEventDestroy(a));
      HIPCHK(hipEventDestroy(b));
      return (double)ms / reps;
    };
    // The trial passes overwrite the force density (bond part only), the
    // dilation and the damage: keep the step's values and put them back.
    const size_t nN = (size_t)c->N;
    double* keep = nullptr;
    HIPCHK(hipMalloc(&keep, (4 * nN + 1) * sizeof(double)));
    for (int d = 0; d < c->dim; d++)
      HIPCHK(hipMemcpyAsync(keep + d * nN, c->ds.f[d], nN * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(keep + 3 * nN, c->ds.theta, nN * 8, hipMemcpyDeviceToDevice, c->stream));
    float* keepZ = nullptr;
    HIPCHK(hipMalloc(&keepZ, (nN + 1) * sizeof(float)));
    HIPCHK(hipMemcpyAsync(keepZ, c->ds.Z, nN * 4, hipMemcpyDeviceToDevice, c->stream));
    const bool prof = c->profiling;
    c->profiling = false;
    double best = 1e300;
    if (c->precision == PDEM_PREC_MIXED && c->ntiles > 0) {
      int bt = c->mixed_tpb, bn = c->tile_nodes;
      for (int tpb : {256, 1024})
        for (int tn : {256, 512, 1024}) {
          c->mixed_tpb = tpb;
          build_tiles(c, tn);
          const double ms = time_pass();
          if (ms < best) {
            best = ms;
            bt = tpb;
            bn = tn;
          }
        }
      c->mixed_tpb = bt;
      build_tiles(c, bn);
    } else {
      int bp = c->pd_tpb;
      const bool have_tile = c->rt.ntiles > 0;
      bool bt = false;
      c->rnp_tile = false;
      for (int tpb : {64, 128, 256}) {
        c->pd_tpb = tpb;
        const double ms = time_pass();
        if (ms < best) {
          best = ms;
          bp = tpb;
        }
      }
      if (have_tile) {
        c->rnp_tile = true;
        const double ms = time_pass();
        if (ms < best) {
          best = ms;
          bt = true;
        }
      }
      c->pd_tpb = bp;
      c->rnp_tile = bt;
    }
    c->profiling = prof;
    for (int d = 0; d < c->dim; d++)
      HIPCHK(hipMemcpyAsync(c->ds.f[d], keep + d * nN, nN * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ds.theta, keep + 3 * nN, nN * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ds.Z, keepZ, nN * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(keep));
    HIPCHK(hipFree(keepZ));
    // launch shapes are baked into a captured step: capture again
    drop_graph(c);
    if (best_ms) *best_ms = best;
  });
}

int pdem_set_launch_config(pdem_ctx* c, int32_t pd_tpb, int32_t mixed_tpb, int32_t tile_nodes) {
  return guarded(c, [&] {
    if (pd_tpb != 64 && pd_tpb != 128 && pd_tpb != 256 && pd_tpb != RT_THREADS)
      throw HipError{"pd_tpb must be 64, 128, 256 or 512 (the RNP pair-tile kernel)"};
    if (pd_tpb == RT_THREADS && c->rt.ntiles == 0) throw HipError{"the RNP pair-tile kernel is not set up for this model"};
    if (mixed_tpb != 256 && mixed_tpb != 1024) throw HipError{"mixed_tpb must be 256 or 1024"};
    if (tile_nodes != 256 && tile_nodes != 512 && tile_nodes != 1024) throw HipError{"tile_nodes must be 256, 512 or 1024"};
    HIPCHK(hipStreamSynchronize(c->stream));
    c->rnp_tile = pd_tpb == RT_THREADS;
    if (!c->rnp_tile) c->pd_tpb = pd_tpb;
    c->mixed_tpb = mixed_tpb;
    if (c->precision == PDEM_PREC_MIXED && c->ntiles > 0) build_tiles(c, tile_nodes);
    drop_graph(c);
  });
}

int pdem_set_contact_light_max(pdem_ctx* c, int32_t light_max) {
  return guarded(c, [&] {
    if (light_max < 0 || light_max > LIGHT_MAX_C)
      throw HipError{"light_max must be in 0.." + std::to_string(LIGHT_MAX_C)};
    if (!c->multi) throw HipError{"contact needs a multi-particle model"};
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ds.light_max = light_max;  // the lists follow at the next rebuild
    if (light_max > 0) c->light_used = true;
    drop_graph(c);
  });
}

int pdem_get_tile_counts(pdem_ctx* c, int32_t* lds_tiles, int32_t* hbm_tiles) {
  return guarded(c, [&] {
    if (!lds_tiles || !hbm_tiles) throw HipError{"null output pointer"};
    *lds_tiles = c->mixed.ntiles_closed + c->mixed.ntiles_lds;
    *hbm_tiles = c->mixed.ntiles_hbm;
  });
}

int pdem_set_step_fusion(pdem_ctx* c, int32_t on) {
  return guarded(c, [&] {
    if (on < 0 || on > 2) throw HipError{"step fusion mode must be 0, 1 or 2"};
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step_fusion = on;
    drop_graph(c);
  });
}

int pdem_set_state_fusion(pdem_ctx* c, int32_t on) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->side) HIPCHK(hipStreamSynchronize(c->side));
    c->state_fusion = on != 0;
    if (c->precision == PDEM_PREC_MIXED && c->ntiles > 0) build_tiles(c, c->tile_nodes);
    drop_graph(c);
  });
}

int pdem_get_launch_config(pdem_ctx* c, int32_t* pd_tpb, int32_t* mixed_tpb, int32_t* tile_nodes) {
  return guarded(c, [&] {
    *pd_tpb = c->rnp_tile ? RT_THREADS : c->pd_tpb;
    *mixed_tpb = c->mixed_tpb;
    *tile_nodes = c->tile_nodes;
  });
}

}  // extern "C"
