/*
 * pdem_oracle.cpp -- CPU FP64 restatement of PeriDEM's explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU engine: only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product (peridem_amd) never links, loads or calls it.
 *
 * It follows the reference's per-node loops line by line in FP64 (built with
 * -ffp-contract=off, the x86-64 -O3 arithmetic of the reference):
 *   node update            src/model/dem/demModel.cpp:349-390 (central difference),
 *                          :392-460 (velocity Verlet)
 *   computeForces          demModel.cpp:462-600
 *   PD forces              demModel.cpp:602-789; RNP material.h:342-368,405-417,
 *                          PMB :732-793, PdElastic :1090-1136, PdState :1415-1451
 *   external forces        demModel.cpp:791-812, particleFLoading.cpp:85-173
 *   displacement BC        particleULoading.cpp:115-230
 *   contact + damping      demModel.cpp:820-1066
 *   PD neighbor list       demModel.cpp:1659-1694 (nsearch.h:340-355 metric)
 *   contact neighbor list  demModel.cpp:1696-1806 (nsearch.h:357-372 metric)
 *   search cadence         demModel.cpp:1808-1955
 *   weighted volume        materialUtil.cpp:21-73
 *   compressive test pp    demModel.cpp:2202-2250
 *   relative compares      src/util/function.cpp:15-23
 *
 * Deliberate difference from the reference, results-neutral: neighbor lists
 * are kept in ascending node order (the reference keeps nanoflann's traversal
 * order).  Sets are identical; only summation order and the 1e-5-slack damage
 * maximum (demModel.cpp:770-773) can differ in the last bits.  The condensed
 * particle-wall list of the damping loop is likewise kept ascending.
 *
 * Threading: a persistent pool runs each per-node loop as a parallel-for, the
 * way the reference runs one Taskflow for_each_index per loop.
 */
#include "pdem.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

const double COMPARE_EPS = 1e-5;  // src/util/function.h:18

// util::isGreater / util::isLess (src/util/function.cpp:15-23)
inline bool isGreater(double a, double b) {
  return (a - b) > ((std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a)) * COMPARE_EPS);
}
inline bool isLess(double a, double b) {
  return (b - a) > ((std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a)) * COMPARE_EPS);
}
// util::equivalentMass (function.cpp:127)
inline double equivalentMass(double m1, double m2) { return 2. * m1 * m2 / (m1 + m2); }
// util::linearStepFunc (function.cpp:62-85)
inline double linearStepFunc(double x, double x1, double x2) {
  double period = std::floor(x / (x1 + x2));
  if (isLess(x, period * (x1 + x2) + x1)) return x - period * x2;
  return (period + 1.) * x1;
}

struct P3 {
  double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 scale(double s, P3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline P3 divs(P3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(P3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// ---------------------------------------------------------------- thread pool
class Pool {
 public:
  explicit Pool(int n) : nthreads_(n < 1 ? 1 : n) {
    for (int t = 1; t < nthreads_; t++) workers_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto &w : workers_) w.join();
  }
  int size() const { return nthreads_; }
  void run(int64_t n, const std::function<void(int64_t, int64_t)> &fn) {
    if (n <= 0) return;
    if (nthreads_ == 1 || n < 256) {
      fn(0, n);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      active_ = nthreads_ - 1;
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [this] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    const int64_t chunk = 64;
    for (;;) {
      int64_t b = next_.fetch_add(chunk);
      if (b >= n_) break;
      (*fn_)(b, std::min(n_, b + chunk));
    }
  }
  void loop(int) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      {
        std::lock_guard<std::mutex> lk(m_);
        if (--active_ == 0) done_cv_.notify_all();
      }
    }
  }
  int nthreads_;
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int64_t, int64_t)> *fn_ = nullptr;
  int64_t n_ = 0;
  std::atomic<int64_t> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct BcSet {
  pdem_bc_set s;
  std::vector<int64_t> nodes;
  std::vector<double> amp;
};

}  // namespace

struct oracle_ctx {
  int dim = 2;
  int multi = 0;
  int inf_type = 0;
  double inf_p[2] = {0, 0};
  double dt = 0;
  int64_t n = 0;
  std::vector<P3> xr, x, u, v, f;
  std::vector<double> vol, mX, theta, vmag;
  std::vector<float> Z;
  std::vector<uint8_t> fix;
  std::vector<int64_t> ptid;
  std::vector<pdem_particle> P;
  int nz = 0;
  std::vector<pdem_contact_pair> C;
  P3 gravity{0, 0, 0};
  double sFactor = 1;
  int64_t deckInterval = 1;
  std::vector<BcSet> ubc, fbc;
  std::vector<char> zeroApplied;

  // PD neighbor list (d_neighPd) and bond states (geometry::Fracture), one byte per bond
  std::vector<int64_t> prow, pcol;
  std::vector<uint8_t> bits;
  // contact candidates (d_neighC) and wall candidates (d_neighWallNodes)
  std::vector<std::vector<int64_t>> neighC;
  std::vector<std::vector<std::vector<int64_t>>> neighWall;
  std::vector<char> searched;

  // cadence controller state (ModelData fields, modelData.h:537-583)
  int64_t step = 0;
  double time = 0;
  double currentDt = 0;
  double maxContactR = 0;
  int64_t contInterval = 0;
  int64_t contCounter = 0;
  double contRadius = 0;
  double maxVelocity = 0;
  std::vector<double> maxVelParticle;
  int initCallCount = 0;
  bool restartActive = false;
  int64_t restartStep = 0;
  int64_t nRebuilds = 0;

  Pool *pool = nullptr;
  std::string err;

  const pdem_particle &part(int64_t i) const { return P[ptid[i]]; }
  double infFn(double r) const {  // getGlobalInfFn(r/delta), influenceFn.cpp
    if (inf_type == PDEM_INF_CONST) return inf_p[0];
    if (inf_type == PDEM_INF_LINEAR) return inf_p[0] + inf_p[1] * r;
    return inf_p[0] * std::exp(-r * r / inf_p[1]);
  }
  bool computeForceNode(int64_t i) const { return part(i).compute_force != 0; }
  bool isState() const { return P[0].material_type == PDEM_MAT_PDSTATE; }
};

namespace {
thread_local std::string g_err;

int fail(oracle_ctx *c, const std::string &m) {
  if (c) c->err = m;
  g_err = m;
  return 1;
}

inline P3 ld(const double *a, int64_t i, int dim) {
  return {a[3 * i], a[3 * i + 1], dim == 3 ? a[3 * i + 2] : 0.};
}

// ------------------------------------------------------------ neighbor search
// PD list (demModel.cpp:1659-1694): j in the same particle, j != i and
// 0 + dx^2 + dy^2 + dz^2 < horizon^2 (nanoflann L2_Simple_Adaptor, strict <).
void buildPdNeighbors(oracle_ctx *c) {
  const int64_t N = c->n;
  std::vector<std::vector<int64_t>> rows(N);
  c->pool->run(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const auto &p = c->part(i);
      const double r2 = p.horizon * p.horizon;
      for (int64_t j = p.glob_start; j < p.glob_end; j++) {
        if (j == i) continue;
        double d2 = 0.;
        double d = c->xr[i].x - c->xr[j].x;
        d2 += d * d;
        d = c->xr[i].y - c->xr[j].y;
        d2 += d * d;
        d = c->xr[i].z - c->xr[j].z;
        d2 += d * d;
        if (d2 < r2) rows[i].push_back(j);
      }
    }
  });
  c->prow.assign(N + 1, 0);
  for (int64_t i = 0; i < N; i++) c->prow[i + 1] = c->prow[i] + (int64_t)rows[i].size();
  c->pcol.resize(c->prow[N]);
  for (int64_t i = 0; i < N; i++)
    std::copy(rows[i].begin(), rows[i].end(), c->pcol.begin() + c->prow[i]);
  c->bits.assign(c->prow[N], 0);
}

// m_x (materialUtil.cpp:21-73, computeStateMx :211-245), all nodes
int computeMx(oracle_ctx *c) {
  std::atomic<int64_t> bad{-1};
  c->pool->run(c->n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const auto &p = c->part(i);
      const double horizon = p.horizon, h = p.mesh_size;
      const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
      double m = 0.;
      for (int64_t k = c->prow[i]; k < c->prow[i + 1]; k++) {
        int64_t j = c->pcol[k];
        double rji = len(sub(c->xr[j], c->xr[i]));
        if (isGreater(rji, horizon)) continue;
        double volj = c->vol[j];
        if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
        m += rji * rji * c->infFn(rji / horizon) * volj;
      }
      if (isLess(m, 1.0E-18)) bad.store(i);
      c->mX[i] = m;
    }
  });
  if (bad.load() >= 0)
    return fail(c, "weighted nodal volume m_x too close to zero at node " +
                       std::to_string(bad.load()));
  return 0;
}

// contact candidates over all nodes of other particles, 0+dx^2+dy^2+dz^2 < r^2
void contactSearch(oracle_ctx *c) {
  const int64_t N = c->n;
  const double r = c->contRadius, r2 = r * r;
  // uniform grid with cell = r (any cell >= r gives the same sets)
  P3 lo = c->x[0], hi = c->x[0];
  for (int64_t i = 0; i < N; i++) {
    lo.x = std::min(lo.x, c->x[i].x);
    lo.y = std::min(lo.y, c->x[i].y);
    lo.z = std::min(lo.z, c->x[i].z);
    hi.x = std::max(hi.x, c->x[i].x);
    hi.y = std::max(hi.y, c->x[i].y);
    hi.z = std::max(hi.z, c->x[i].z);
  }
  double cell = r > 0 ? r : 1.;
  auto ncell = [&](double a, double b) {
    double k = std::floor((b - a) / cell) + 1;
    return (int64_t)std::min(k, 2048.);
  };
  int64_t nx = ncell(lo.x, hi.x), ny = ncell(lo.y, hi.y), nzc = c->dim == 3 ? ncell(lo.z, hi.z) : 1;
  auto cidx = [&](double a, double l, int64_t nn) {
    int64_t k = (int64_t)std::floor((a - l) / cell);
    return std::max<int64_t>(0, std::min(nn - 1, k));
  };
  std::vector<int64_t> key(N), start(nx * ny * nzc + 1, 0), order(N);
  for (int64_t i = 0; i < N; i++) {
    key[i] = (cidx(c->x[i].z, lo.z, nzc) * ny + cidx(c->x[i].y, lo.y, ny)) * nx +
             cidx(c->x[i].x, lo.x, nx);
    start[key[i] + 1]++;
  }
  for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
  {
    std::vector<int64_t> fillp(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < N; i++) order[fillp[key[i]]++] = i;  // ascending within cell
  }
  c->neighC.resize(N);
  c->searched.assign(N, 0);
  c->pool->run(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const auto &p = c->part(i);
      // demModel.cpp:1722-1729: walls, all-constrained and force-off particles do not search
      bool search = !(p.type_index == 1) && !(p.all_dofs_constrained || !p.compute_force);
      if (!search) continue;
      c->searched[i] = 1;
      auto &lst = c->neighC[i];
      lst.clear();
      int64_t cx = cidx(c->x[i].x, lo.x, nx), cy = cidx(c->x[i].y, lo.y, ny),
              cz = cidx(c->x[i].z, lo.z, nzc);
      for (int64_t z = std::max<int64_t>(0, cz - 1); z <= std::min(nzc - 1, cz + 1); z++)
        for (int64_t y = std::max<int64_t>(0, cy - 1); y <= std::min(ny - 1, cy + 1); y++)
          for (int64_t xx = std::max<int64_t>(0, cx - 1); xx <= std::min(nx - 1, cx + 1); xx++) {
            int64_t cc = (z * ny + y) * nx + xx;
            for (int64_t s = start[cc]; s < start[cc + 1]; s++) {
              int64_t j = order[s];
              if (j == i || c->ptid[j] == c->ptid[i]) continue;
              double d2 = 0., d = c->x[i].x - c->x[j].x;
              d2 += d * d;
              d = c->x[i].y - c->x[j].y;
              d2 += d * d;
              d = c->x[i].z - c->x[j].z;
              d2 += d * d;
              if (d2 < r2) lst.push_back(j);
            }
          }
      std::sort(lst.begin(), lst.end());
    }
  });
  // wall sublists per non-wall particle (demModel.cpp:1760-1805)
  c->neighWall.assign(c->P.size(), {});
  for (size_t pi = 0; pi < c->P.size(); pi++) {
    const auto &p = c->P[pi];
    if (p.type_index == 1) continue;
    auto &w = c->neighWall[pi];
    w.resize(p.glob_end - p.glob_start);
    for (int64_t il = 0; il < p.glob_end - p.glob_start; il++) {
      int64_t ig = p.glob_start + il;
      for (int64_t j : c->neighC[ig])
        if (c->P[c->ptid[j]].type_index == 1) w[il].push_back(j);
    }
  }
  c->nRebuilds++;
}

// updateContactNeighborSearchParameters (demModel.cpp:1808-1955)
bool searchParameters(oracle_ctx *c) {
  const int64_t deckI = c->deckInterval;
  if (c->contInterval == 0 && isLess(c->contRadius, 1.e-16)) {
    c->contInterval = deckI;
    c->contCounter = c->step % c->contInterval;
    c->contRadius = c->maxContactR * c->sFactor;
  }
  if (c->step == 0) {
    c->initCallCount++;
    if (c->initCallCount == 1) return true;
    if (c->initCallCount == 2) {
      c->contCounter++;
      return (c->contCounter - 1) % c->contInterval == 0;
    }
  }
  // restart step (demModel.cpp:1833-1837)
  if (c->restartActive && c->step == c->restartStep) c->contCounter = c->step % c->contInterval;
  if (c->contInterval == 1) {
    c->contRadius = c->maxContactR;
    c->contCounter++;
    return (c->contCounter - 1) % c->contInterval == 0;
  }
  int64_t upi = c->contInterval > 5 ? (int64_t)(0.2 * (double)c->contInterval) : 1;
  if (c->contCounter > 0 && c->contCounter % upi != 0) {
    c->contCounter++;
    return (c->contCounter - 1) % c->contInterval == 0;
  }
  for (size_t pi = 0; pi < c->P.size(); pi++) {
    double mv = c->vmag[c->P[pi].glob_start];  // first maximum (stable sort, methods.h:117-149)
    for (int64_t i = c->P[pi].glob_start; i < c->P[pi].glob_end; i++)
      if (c->vmag[i] > mv) mv = c->vmag[i];
    c->maxVelParticle[pi] = mv;
  }
  double mv = c->maxVelParticle[0];
  for (double a : c->maxVelParticle)
    if (a > mv) mv = a;
  c->maxVelocity = mv;
  int64_t up_old = c->contInterval;
  double safety = c->sFactor > 5 ? c->sFactor : 10;
  double r_contact = c->sFactor * c->maxContactR;
  double r_vel = c->maxVelocity * c->currentDt * (double)deckI * safety;
  if (isGreater(r_vel, r_contact)) {
    c->contInterval = (int64_t)(c->maxContactR / (c->maxVelocity * c->currentDt));
    (void)up_old;
    c->contRadius = r_contact;
    c->contCounter = 0;
    if (c->contInterval < 1) {
      c->contInterval = 1;
      c->contRadius = c->maxContactR;
    }
  } else {
    c->contRadius = c->contInterval < 2 ? c->maxContactR : r_contact;
  }
  c->contCounter++;
  return (c->contCounter - 1) % c->contInterval == 0;
}

void updateContactNeighborlist(oracle_ctx *c) {
  if (searchParameters(c)) contactSearch(c);
}

// ------------------------------------------------------------- bond laws
// RnpMaterial::getBondEF(r, s, fs, break_bonds = true), force part
// (material.h:342-368); fs is updated in place.
inline double rnpBondF(double J, double r, double S, const double *mat, int irrev, bool &fs) {
  const double Cc = mat[0], beta = mat[1], rbar = mat[2], invF = mat[3], fsc = mat[4];
  if (irrev && !fs && isGreater(std::abs(S), fsc * (rbar / std::sqrt(r)))) fs = true;
  if (fs) return 0.;
  return J * 4. * S * Cc * beta * std::exp(-beta * r * S * S) / invF;
}
// PmbMaterial::getBondEF(r, s, fs, break_bonds = true), force part
// (material.h:732-757); PdElastic::getBondEF (material.h:1090-1097) never
// breaks.  J = getInfFn(r); fs updated in place.
inline double pmbBondF(double J, double S, double cc, double s0, bool &fs) {
  if (!fs && isGreater(std::abs(S), s0 + 1.0e-10)) fs = true;
  if (fs) return 0.;
  return J * cc * S;
}
inline double pdelasticBondF(double J, double S, double cc) { return J * cc * S; }
// PdState::getBondEF(r, s, fs, mx, thetax), force part (material.h:1415-1429)
inline double pdstateBondF(double J, double r, double S, double K, double G, double mx,
                           double thetax) {
  double cl = S * r;
  double alpha = 15. * G / mx;
  double factor = (3. * K / mx) - alpha / 3.;
  return J * (r * thetax * factor + cl * alpha);
}

// ---------------------------------------------------------------- PD forces
void computePeridynamicForces(oracle_ctx *c) {
  const int64_t N = c->n;
  if (c->isState()) {
    // dilation pass, demModel.cpp:610-679
    c->pool->run(N, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) {
        if (!c->computeForceNode(i)) continue;
        const auto &p = c->part(i);
        if (p.material_type != PDEM_MAT_PDSTATE) continue;
        const double horizon = p.horizon, h = p.mesh_size;
        const P3 xi = c->xr[i], ui = c->u[i];
        const double m = c->mX[i];
        double theta = 0.;
        const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
        const double sc = p.mat[2];
        for (int64_t k = c->prow[i]; k < c->prow[i + 1]; k++) {
          int64_t j = c->pcol[k];
          const P3 xj = c->xr[j], uj = c->u[j];
          double rji = len(sub(xj, xi));
          double change_length = len(sub(add(sub(xj, xi), uj), ui)) - rji;
          double s = change_length / rji;
          bool fs = c->bits[k];
          if (!fs && isGreater(std::abs(s), sc + 1.0e-10)) fs = true;
          c->bits[k] = fs;
          if (!fs) {
            double volj = c->vol[j];
            if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
            theta += rji * change_length * c->infFn(rji / horizon) * volj;
          }
        }
        c->theta[i] = 3. * theta / m;
      }
    });
  }
  // force pass, demModel.cpp:682-788
  c->pool->run(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (!c->computeForceNode(i)) continue;
      const auto &p = c->part(i);
      P3 force{0, 0, 0};
      float Zi = 0.;
      const double horizon = p.horizon, h = p.mesh_size;
      const P3 xi = c->xr[i], ui = c->u[i];
      const double mi = c->mX[i], thetai = c->theta[i];
      const double check_up = horizon + 0.5 * h, check_low = horizon - 0.5 * h;
      const bool state = p.material_type == PDEM_MAT_PDSTATE;
      // PMB and PdElastic: actual stretch and deformed direction like PdState,
      // pairwise force like RNP (demModel.cpp:741-756)
      const bool deformed = p.material_type == PDEM_MAT_PMB || p.material_type == PDEM_MAT_PDELASTIC;
      for (int64_t k = c->prow[i]; k < c->prow[i + 1]; k++) {
        int64_t j = c->pcol[k];
        bool fs = c->bits[k];
        const P3 xj = c->xr[j], uj = c->u[j];
        double volj = c->vol[j];
        const P3 dx = sub(xj, xi), du = sub(uj, ui);
        double rji = len(dx);
        double Sji;
        if (state || deformed) {
          const P3 dy = add(dx, du);
          Sji = (len(dy) - len(dx)) / len(dx);  // getS, material.h:791-793, 1134-1136, 1449-1451
        } else {
          Sji = dot(dx, du) / dot(dx, dx);  // RnpMaterial::getS, material.h:405-407
        }
        if (!fs) {
          if (isGreater(rji, check_low)) volj *= (check_up - rji) / h;
          double J = c->infFn(rji / horizon);
          if (state) {
            // PdState::getBondEF, material.h:1415-1429 (evaluated for i and j)
            const double K = p.mat[0], G = p.mat[1];
            const double mj = c->mX[j], thetaj = c->theta[j];
            double ef_i = pdstateBondF(J, rji, Sji, K, G, mi, thetai);
            double ef_j = pdstateBondF(J, rji, Sji, K, G, mj, thetaj);
            double scalar_f = (ef_i + ef_j) * volj;
            const P3 dy = add(dx, du);
            force = add(force, scale(scalar_f, divs(dy, len(dy))));
          } else if (deformed) {
            double ef = p.material_type == PDEM_MAT_PMB ? pmbBondF(J, Sji, p.mat[0], p.mat[1], fs)
                                                         : pdelasticBondF(J, Sji, p.mat[0]);
            c->bits[k] = fs;
            double scalar_f = ef * volj;
            const P3 dy = add(dx, du);
            force = add(force, scale(scalar_f, divs(dy, len(dy))));
          } else {
            // RnpMaterial::getBondEF with break_bonds = true, material.h:342-368
            double ef = rnpBondF(J, rji, Sji, p.mat, p.irreversible, fs);
            c->bits[k] = fs;
            double scalar_f = ef * volj;
            force = add(force, scale(scalar_f, divs(dx, len(dx))));
          }
        } else {
          // broken bond: short-range repulsion, demModel.cpp:760-768
          P3 yji = sub(add(xj, uj), add(xi, ui));
          double Rji = len(yji);
          double scalar_f = p.Kn * volj * (Rji - p.Rc) / Rji;
          if (scalar_f > 0.) scalar_f = 0.;
          force = add(force, scale(scalar_f, yji));
        }
        double Sc = state ? p.mat[2] : p.mat[2] / std::sqrt(rji);
        if (p.material_type == PDEM_MAT_PMB) Sc = p.mat[1];
        if (p.material_type == PDEM_MAT_PDELASTIC) Sc = std::numeric_limits<double>::max();
        if (isGreater(std::abs(Sji / Sc), Zi)) Zi = (float)std::abs(Sji / Sc);
      }
      c->f[i] = force;
      c->Z[i] = Zi;
    }
  });
}

// ----------------------------------------------------------- contact forces
void computeContactForces(oracle_ctx *c) {
  const int64_t N = c->n;
  c->pool->run(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (!c->computeForceNode(i)) continue;  // d_fContCompNodes
      const auto &pi = c->part(i);
      P3 force{0, 0, 0};
      const P3 yi = c->x[i], vi = c->v[i];
      for (int64_t j : c->neighC[i]) {
        const P3 yj = c->x[j];
        double Rji = len(sub(yj, yi));
        const auto &pj = c->part(j);
        bool both_walls = pi.type_index == 1 && pj.type_index == 1;
        if (j == i || c->ptid[j] == c->ptid[i] || both_walls) continue;
        const auto &ct = c->C[pi.zone * c->nz + pj.zone];
        if (!isLess(Rji, ct.contact_r)) continue;
        P3 yji = sub(c->x[j], yi);
        double volj = c->vol[j];
        P3 vji = sub(c->v[j], vi);
        P3 en = divs(yji, Rji);
        double vn_mag = dot(vji, en);
        P3 et = sub(vji, scale(vn_mag, en));
        if (isGreater(len(et), 0.))
          et = divs(et, len(et));
        else
          et = {0, 0, 0};
        double scalar_f = ct.Kn * (Rji - ct.contact_r) * volj;
        if (scalar_f > 0.) scalar_f = 0.;
        force = add(force, scale(scalar_f, en));
        force = add(force, scale(ct.mu * scalar_f, et));
      }
      c->f[i] = add(c->f[i], force);
    }
  });

  // damping, demModel.cpp:937-1065 (serial over particles)
  const double PI = M_PI;
  for (size_t pid = 0; pid < c->P.size(); pid++) {
    const auto &pi = c->P[pid];
    if (pi.type_index == 1) continue;
    double Ri = pi.bounding_radius;
    double vol_pi = PI * Ri * Ri;
    P3 pi_xc = c->x[pi.center_node], pi_vc = c->v[pi.center_node];
    double rhoi = pi.density;
    P3 force_i{0, 0, 0};
    for (size_t qid = 0; qid < c->P.size(); qid++) {
      const auto &pj = c->P[qid];
      if (pj.type_index == 1 || qid == pid) continue;
      double Rj = pj.bounding_radius;
      P3 xc_ji = sub(c->x[pj.center_node], pi_xc);
      double dist = len(xc_ji);
      const auto &ct = c->C[pi.zone * c->nz + pj.zone];
      if (isLess(dist, Rj + Ri + 1.01 * ct.contact_r)) {
        double vol_pj = PI * Rj * Rj;
        double meq = equivalentMass(rhoi * vol_pi, pj.density * vol_pj);
        double beta_n = ct.betan * std::sqrt(ct.kappa * ct.contact_r * meq);
        P3 hat{0, 0, 0};
        if (isGreater(dist, 0.)) hat = divs(xc_ji, dist);
        P3 vc_ji = sub(c->v[pj.center_node], pi_vc);
        double vc_mag = dot(vc_ji, hat);
        if (vc_mag > 0.) vc_mag = 0.;
        force_i = add(force_i, divs(scale(beta_n * vc_mag, hat), vol_pi));
      }
    }
    // condensed wall nodes (ascending)
    std::vector<int64_t> cond;
    if (!c->neighWall.empty() && !c->neighWall[pid].empty()) {
      for (int64_t jl = 0; jl < (int64_t)c->neighWall[pid].size(); jl++) {
        const P3 yj = c->x[pi.glob_start + jl];
        for (int64_t k_id : c->neighWall[pid][jl]) {
          const auto &pk = c->part(k_id);
          double Rjk = len(sub(c->x[k_id], yj));
          const auto &ct = c->C[pi.zone * c->nz + pk.zone];
          if (isLess(Rjk, ct.contact_r)) cond.push_back(k_id);
        }
      }
      std::sort(cond.begin(), cond.end());
      cond.erase(std::unique(cond.begin(), cond.end()), cond.end());
    }
    for (int64_t j : cond) {
      const auto &pj = c->part(j);
      double meq = rhoi * vol_pi;
      const auto &ct = c->C[pi.zone * c->nz + pj.zone];
      double beta_n = ct.betan * std::sqrt(ct.kappa * ct.contact_r * meq);
      P3 xc_ji = sub(c->x[j], pi_xc);
      P3 hat{0, 0, 0};
      if (isGreater(len(xc_ji), 0.)) hat = divs(xc_ji, len(xc_ji));
      P3 vc_ji = sub(c->v[j], pi_vc);
      double vc_mag = dot(vc_ji, hat);
      if (vc_mag > 0.) vc_mag = 0.;
      force_i = add(force_i, divs(scale(beta_n * vc_mag, hat), vol_pi));
    }
    for (int64_t i = pi.glob_start; i < pi.glob_end; i++) c->f[i] = add(c->f[i], force_i);
  }
}

inline void setc(P3 &p, int d, double val) { (d == 0 ? p.x : d == 1 ? p.y : p.z) = val; }
inline double getc(const P3 &p, int d) { return d == 0 ? p.x : d == 1 ? p.y : p.z; }

// ParticleFLoading::apply (particleFLoading.cpp:85-173)
void computeExternalForces(oracle_ctx *c) {
  const P3 g = c->gravity;
  if (len(g) > 1.0E-8) {
    c->pool->run(c->n, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) c->f[i] = add(c->f[i], scale(c->part(i).density, g));
    });
  }
  const double time = c->time;
  for (auto &bs : c->fbc) {
    const auto &s = bs.s;
    for (size_t k = 0; k < bs.nodes.size(); k++) {
      double fmax = bs.amp[k];
      if (s.time_fn == PDEM_FFN_LINEAR)
        fmax *= time;
      else if (s.time_fn == PDEM_FFN_LINEAR_STEP)
        fmax *= linearStepFunc(time, s.time_params[1], s.time_params[2]);
      else if (s.time_fn == PDEM_FFN_LINEAR_SLOW_FAST) {
        if (isGreater(time, s.time_params[1]))
          fmax *= s.time_params[3] * time;
        else
          fmax *= s.time_params[2] * time;
      } else if (s.time_fn == PDEM_FFN_SIN) {
        double a = M_PI * s.time_params[1];
        fmax *= std::sin(a * time);
      }
      fmax *= s.time_params[0];
      P3 fi{0, 0, 0};
      for (int d = 0; d < 3; d++)
        if (s.dir_mask >> d & 1) setc(fi, d, fmax);
      int64_t i = bs.nodes[k];
      c->f[i] = add(c->f[i], fi);
    }
  }
}

// ParticleULoading::apply (particleULoading.cpp:115-230)
void computeExternalDisplacementBC(oracle_ctx *c) {
  const double time = c->time;
  for (size_t si = 0; si < c->ubc.size(); si++) {
    auto &bs = c->ubc[si];
    const auto &s = bs.s;
    if (s.is_zero && c->zeroApplied[si]) continue;
    if (s.is_zero) c->zeroApplied[si] = 1;
    for (size_t k = 0; k < bs.nodes.size(); k++) {
      int64_t i = bs.nodes[k];
      double umax = bs.amp[k], du = 0., dv = 0.;
      if (s.time_fn == PDEM_UFN_CONSTANT)
        du = umax;
      else if (s.time_fn == PDEM_UFN_LINEAR) {
        du = umax * time;
        dv = umax;
      } else if (s.time_fn == PDEM_UFN_QUADRATIC) {
        du = umax * time + s.time_params[1] * time * time;
        dv = umax + s.time_params[1] * time;
      } else if (s.time_fn == PDEM_UFN_SIN) {
        double a = M_PI * s.time_params[1];
        du = umax * std::sin(a * time);
        dv = umax * a * std::cos(a * time);
      }
      P3 ui{0, 0, 0}, vi{0, 0, 0};
      for (int d = 0; d < 3; d++)
        if (s.dir_mask >> d & 1) {
          setc(ui, d, du);
          setc(vi, d, dv);
        }
      if (s.time_fn == PDEM_UFN_ROTATION) {
        const P3 x0{s.time_params[1], s.time_params[2], s.time_params[3]};
        const P3 xx = c->xr[i];
        const P3 dx = sub(xx, x0);
        const double th = s.time_params[0] * time;
        // util::rotate2D / derRotate2D (transformation.cpp)
        const P3 rx{dx.x * std::cos(th) - dx.y * std::sin(th),
                    dx.x * std::sin(th) + dx.y * std::cos(th), 0.};
        const P3 drx{-dx.x * std::sin(th) - dx.y * std::cos(th),
                     dx.x * std::cos(th) - dx.y * std::sin(th), 0.};
        ui = add(ui, sub(rx, dx));
        vi = add(vi, scale(s.time_params[0], drx));
      }
      for (int d = 0; d < 3; d++)
        if (s.dir_mask >> d & 1) {
          setc(c->u[i], d, getc(ui, d));
          setc(c->v[i], d, getc(vi, d));
          setc(c->x[i], d, getc(ui, d) + getc(c->xr[i], d));
        }
    }
  }
}

void computeForces(oracle_ctx *c) {
  for (int64_t i = 0; i < c->n; i++) c->f[i] = {0, 0, 0};
  computePeridynamicForces(c);
  if (c->multi) {
    updateContactNeighborlist(c);
    computeContactForces(c);
  }
  computeExternalForces(c);
}

// integrateCD, demModel.cpp:349-390
void integrateCD(oracle_ctx *c) {
  c->currentDt = c->dt;
  const int dim = c->dim;
  c->pool->run(c->n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (!c->computeForceNode(i)) continue;
      const double rho = c->part(i).density;
      const uint8_t fx = c->fix[i];
      for (int dof = 0; dof < dim; dof++) {
        if (!(fx >> dof & 1)) {
          double vv = getc(c->v[i], dof) + (c->currentDt / rho) * getc(c->f[i], dof);
          setc(c->v[i], dof, vv);
          setc(c->u[i], dof, getc(c->u[i], dof) + c->currentDt * vv);
          setc(c->x[i], dof, getc(c->x[i], dof) + c->currentDt * vv);
        }
      }
      c->vmag[i] = len(c->v[i]);
    }
  });
  c->step++;
  c->time += c->currentDt;
  computeExternalDisplacementBC(c);
  computeForces(c);
}

// integrateVerlet, demModel.cpp:392-460: half kick and drift, n++, t += dt,
// displacement BC, forces, second half kick.  vMag is set after each kick.
void integrateVerlet(oracle_ctx *c) {
  c->currentDt = c->dt;
  const int dim = c->dim;
  c->pool->run(c->n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (!c->computeForceNode(i)) continue;
      const double rho = c->part(i).density;
      const uint8_t fx = c->fix[i];
      for (int dof = 0; dof < dim; dof++) {
        if (!(fx >> dof & 1)) {
          double vv = getc(c->v[i], dof) + 0.5 * (c->currentDt / rho) * getc(c->f[i], dof);
          setc(c->v[i], dof, vv);
          setc(c->u[i], dof, getc(c->u[i], dof) + c->currentDt * vv);
          setc(c->x[i], dof, getc(c->x[i], dof) + c->currentDt * vv);
        }
      }
      c->vmag[i] = len(c->v[i]);
    }
  });
  c->step++;
  c->time += c->currentDt;
  computeExternalDisplacementBC(c);
  computeForces(c);
  c->pool->run(c->n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (!c->computeForceNode(i)) continue;
      const double rho = c->part(i).density;
      const uint8_t fx = c->fix[i];
      for (int dof = 0; dof < dim; dof++)
        if (!(fx >> dof & 1)) setc(c->v[i], dof, getc(c->v[i], dof) + 0.5 * (c->currentDt / rho) * getc(c->f[i], dof));
      c->vmag[i] = len(c->v[i]);
    }
  });
}

}  // namespace

// ======================================================================= ABI
extern "C" {

const char *oracle_last_error(const oracle_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int oracle_create(const pdem_desc *d, int32_t nthreads, oracle_ctx **out) {
  *out = nullptr;
  if (!d || d->n_nodes <= 0 || (d->dim != 2 && d->dim != 3) || d->n_particles <= 0)
    return fail(nullptr, "invalid descriptor");
  auto *c = new oracle_ctx();
  c->pool = new Pool(nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency());
  c->dim = d->dim;
  c->multi = d->multi_particle;
  c->inf_type = d->influence_fn_type;
  c->inf_p[0] = d->influence_fn_params[0];
  c->inf_p[1] = d->influence_fn_params[1];
  c->dt = d->dt;
  const int64_t N = d->n_nodes;
  c->n = N;
  c->xr.resize(N);
  c->x.resize(N);
  c->u.resize(N);
  c->v.resize(N);
  c->f.assign(N, {0, 0, 0});
  for (int64_t i = 0; i < N; i++) {
    c->xr[i] = ld(d->x_ref, i, c->dim);
    c->u[i] = d->u ? ld(d->u, i, c->dim) : P3{0, 0, 0};
    c->v[i] = d->v ? ld(d->v, i, c->dim) : P3{0, 0, 0};
    c->x[i] = add(c->xr[i], c->u[i]);
  }
  c->vol.assign(d->vol, d->vol + N);
  c->fix.assign(d->fix, d->fix + N);
  c->ptid.assign(d->pt_id, d->pt_id + N);
  c->mX.assign(N, 0.);
  c->theta.assign(N, 0.);
  c->vmag.assign(N, 0.);
  c->Z.assign(N, 0.f);
  c->P.assign(d->particles, d->particles + d->n_particles);
  c->nz = d->n_zones;
  c->C.assign(d->contact, d->contact + (size_t)d->n_zones * d->n_zones);
  c->gravity = {d->gravity[0], d->gravity[1], d->dim == 3 ? d->gravity[2] : 0.};
  c->sFactor = d->search_factor;
  c->deckInterval = d->neigh_update_interval;
  for (int s = 0; s < d->n_ubc; s++) {
    BcSet b;
    b.s = d->ubc[s];
    b.nodes.assign(d->ubc[s].nodes, d->ubc[s].nodes + d->ubc[s].n_entries);
    b.amp.assign(d->ubc[s].amp, d->ubc[s].amp + d->ubc[s].n_entries);
    c->ubc.push_back(b);
  }
  for (int s = 0; s < d->n_fbc; s++) {
    BcSet b;
    b.s = d->fbc[s];
    b.nodes.assign(d->fbc[s].nodes, d->fbc[s].nodes + d->fbc[s].n_entries);
    b.amp.assign(d->fbc[s].amp, d->fbc[s].amp + d->fbc[s].n_entries);
    c->fbc.push_back(b);
  }
  c->zeroApplied.assign(c->ubc.size(), 0);
  c->maxVelParticle.assign(c->P.size(), 0.);
  for (int64_t i = 0; i < N; i++)
    if (c->ptid[i] < 0 || c->ptid[i] >= d->n_particles) {
      int r = fail(c, "pt_id out of range");
      delete c->pool;
      delete c;
      return r;
    }

  // DEMModel::init order (demModel.cpp:109-266)
  if (c->multi) {
    c->maxContactR = 0.;
    for (auto &ct : c->C)
      if (c->maxContactR < ct.contact_r) c->maxContactR = ct.contact_r;
  }
  buildPdNeighbors(c);
  if (c->multi) {
    c->contInterval = c->deckInterval;
    c->contRadius = c->sFactor * c->maxContactR;
    updateContactNeighborlist(c);
  }
  if (computeMx(c)) {
    g_err = c->err;
    delete c->pool;
    delete c;
    return 1;
  }
  *out = c;
  return 0;
}

void oracle_destroy(oracle_ctx *c) {
  if (!c) return;
  delete c->pool;
  delete c;
}

int oracle_integrate_begin(oracle_ctx *c) {
  computeExternalDisplacementBC(c);
  computeForces(c);
  return 0;
}

int oracle_integrate_cd(oracle_ctx *c, int64_t nsteps) {
  for (int64_t s = 0; s < nsteps; s++) integrateCD(c);
  return 0;
}

int oracle_integrate_verlet(oracle_ctx *c, int64_t nsteps) {
  for (int64_t s = 0; s < nsteps; s++) integrateVerlet(c);
  return 0;
}

int oracle_compute_forces(oracle_ctx *c) {
  computeForces(c);
  return 0;
}

int oracle_compute_peridynamic_forces(oracle_ctx *c) {
  computePeridynamicForces(c);
  return 0;
}

int oracle_get_field(oracle_ctx *c, int32_t field, void *host) {
  const int64_t N = c->n;
  auto putp = [&](const std::vector<P3> &a) {
    double *o = (double *)host;
    for (int64_t i = 0; i < N; i++) {
      o[3 * i] = a[i].x;
      o[3 * i + 1] = a[i].y;
      o[3 * i + 2] = c->dim == 3 ? a[i].z : 0.;
    }
  };
  switch (field) {
    case PDEM_FIELD_X: putp(c->x); break;
    case PDEM_FIELD_U: putp(c->u); break;
    case PDEM_FIELD_V: putp(c->v); break;
    case PDEM_FIELD_F: putp(c->f); break;
    case PDEM_FIELD_XREF: putp(c->xr); break;
    case PDEM_FIELD_VOL: std::memcpy(host, c->vol.data(), N * 8); break;
    case PDEM_FIELD_Z: std::memcpy(host, c->Z.data(), N * 4); break;
    case PDEM_FIELD_THETA: std::memcpy(host, c->theta.data(), N * 8); break;
    case PDEM_FIELD_MX: std::memcpy(host, c->mX.data(), N * 8); break;
    case PDEM_FIELD_VMAG: std::memcpy(host, c->vmag.data(), N * 8); break;
    case PDEM_FIELD_BONDS: std::memcpy(host, c->bits.data(), c->bits.size()); break;
    default: return fail(c, "unknown field");
  }
  return 0;
}

int oracle_set_field(oracle_ctx *c, int32_t field, const void *host) {
  const int64_t N = c->n;
  auto getp = [&](std::vector<P3> &a) {
    const double *o = (const double *)host;
    for (int64_t i = 0; i < N; i++) a[i] = {o[3 * i], o[3 * i + 1], c->dim == 3 ? o[3 * i + 2] : 0.};
  };
  switch (field) {
    case PDEM_FIELD_X: getp(c->x); break;
    case PDEM_FIELD_U: getp(c->u); break;
    case PDEM_FIELD_V: getp(c->v); break;
    case PDEM_FIELD_F: getp(c->f); break;
    case PDEM_FIELD_THETA: std::memcpy(c->theta.data(), host, N * 8); break;
    case PDEM_FIELD_BONDS: std::memcpy(c->bits.data(), host, c->bits.size()); break;
    default: return fail(c, "field is read-only");
  }
  return 0;
}

int oracle_get_time(oracle_ctx *c, int64_t *n, double *t) {
  *n = c->step;
  *t = c->time;
  return 0;
}

int oracle_get_neighbor_count(oracle_ctx *c, int32_t kind, int64_t *nnz) {
  if (kind == PDEM_NEIGH_PD) {
    *nnz = (int64_t)c->pcol.size();
    return 0;
  }
  int64_t s = 0;
  for (auto &l : c->neighC) s += (int64_t)l.size();
  *nnz = s;
  return 0;
}

int oracle_get_neighbors(oracle_ctx *c, int32_t kind, int64_t *rowptr, int64_t *cols) {
  if (kind == PDEM_NEIGH_PD) {
    std::memcpy(rowptr, c->prow.data(), (c->n + 1) * 8);
    std::memcpy(cols, c->pcol.data(), c->pcol.size() * 8);
    return 0;
  }
  rowptr[0] = 0;
  for (int64_t i = 0; i < c->n; i++) {
    const auto *l = (size_t)i < c->neighC.size() ? &c->neighC[i] : nullptr;
    int64_t m = l ? (int64_t)l->size() : 0;
    if (m) std::memcpy(cols + rowptr[i], l->data(), m * 8);
    rowptr[i + 1] = rowptr[i] + m;
  }
  return 0;
}

int oracle_get_contact_state(oracle_ctx *c, pdem_contact_state *st) {
  st->interval = c->contInterval;
  st->counter = c->contCounter;
  st->search_radius = c->contRadius;
  st->max_velocity = c->maxVelocity;
  st->max_contact_r = c->maxContactR;
  st->n_rebuilds = c->nRebuilds;
  st->overflow = 0;
  st->capacity = 0;
  return 0;
}

int oracle_pp_compressive_test(oracle_ctx *c, int32_t wall_id, int32_t f_dir, double *pen,
                               double *reac) {
  if (wall_id < 0 || wall_id >= (int)c->P.size() || f_dir < 0 || f_dir > 2)
    return fail(c, "bad wall id or direction");
  const auto &w = c->P[wall_id];
  int64_t i0 = w.glob_start;
  *pen = getc(c->x[i0], f_dir) - getc(c->xr[i0], f_dir);
  double r = 0.;
  for (int64_t i = w.glob_start; i < w.glob_end; i++) r += getc(c->f[i], f_dir) * c->vol[i];
  *reac = r;
  return 0;
}

// DEMModel::ppTwoParticleTest, demModel.cpp:2096-2149 (without the
// d_dtTestOut cadence): out5 = pen_dist, contact_area_radius, max_y,
// max_dist, max_y_loc
int oracle_pp_two_particle_test(oracle_ctx *c, double *out5) {
  if (c->P.size() < 2 || !c->multi) return fail(c, "the two-particle test needs two particles");
  const auto &p0 = c->P[0];
  const auto &p1 = c->P[1];
  const P3 xc0 = c->x[p0.center_node], xc1 = c->x[p1.center_node];
  const double r = p0.bounding_radius;
  const auto &ct = c->C[p0.zone * c->nz + p1.zone];
  const double r_e = r + ct.contact_r;
  const double dist = std::sqrt((xc0.x - xc1.x) * (xc0.x - xc1.x) + (xc0.y - xc1.y) * (xc0.y - xc1.y) +
                                (xc0.z - xc1.z) * (xc0.z - xc1.z));
  double pen = dist - r_e - r;
  double area_r = 0.;
  if (isLess(pen, 0.))
    area_r = std::sqrt(std::pow(r_e, 2.) - std::pow(r_e + pen, 2.));
  else if (isGreater(pen, 0.)) {
    pen = 0.;
    area_r = 0.;
  }
  const double max_dist = xc1.y + p1.bounding_radius;
  double max_y_loc = c->x[p1.glob_start].y;
  for (int64_t i = p1.glob_start; i < p1.glob_end; i++)
    if (isLess(max_y_loc, c->x[i].y)) max_y_loc = c->x[i].y;
  double max_y = 0.;
  if (isLess(max_y, max_y_loc)) max_y = max_y_loc;
  out5[0] = pen;
  out5[1] = area_r;
  out5[2] = max_y;
  out5[3] = max_dist;
  out5[4] = max_y_loc;
  return 0;
}

// DEMModel::checkStop, demModel.cpp:2151-2200 (the criterion's quantity;
// util::methods::maxLengthAndMaxLengthIndex, methods.h:38-48, 208-215)
int oracle_check_stop(oracle_ctx *c, int32_t criterion, double param, int32_t *stop, double *value,
                      int64_t *node) {
  *node = -1;
  if (criterion == PDEM_STOP_MAX_PARTICLE_DIST) {
    if (c->P.size() < 2) return fail(c, "max_particle_dist needs two particles");
    const P3 a = c->x[c->P[0].center_node], b = c->x[c->P[1].center_node];
    *value = std::sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
  } else if (criterion == PDEM_STOP_MAX_NODE_DIST) {
    double m = -1.;
    int64_t k = 0;
    for (int64_t i = 0; i < c->n; i++) {
      const P3 p = c->x[i];
      const double l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
      if (l > m) {
        m = l;
        k = i;
      }
    }
    *value = m;
    *node = k;
  } else {
    return fail(c, "unknown stop criterion");
  }
  *stop = isGreater(*value, param) ? 1 : 0;
  return 0;
}

// DEMModel::restart (demModel.cpp:84-103): the step counter and time of the
// restart step; positions, displacements and velocities are set by the
// caller (set_field) from the restart file
int oracle_restart(oracle_ctx *c, int64_t step) {
  if (step < 0) return fail(c, "negative restart step");
  c->step = step;
  c->time = double(step) * c->dt;
  c->restartActive = true;
  c->restartStep = step;
  return 0;
}

int oracle_broken_bonds(oracle_ctx *c, int64_t *broken, int64_t *total) {
  int64_t b = 0;
  for (auto x : c->bits) b += x ? 1 : 0;
  *broken = b;
  *total = (int64_t)c->bits.size();
  return 0;
}

// pieces used by the known-answer tests
int oracle_is_greater(double a, double b) { return isGreater(a, b); }
int oracle_is_less(double a, double b) { return isLess(a, b); }
/* returns fs after the call; force in *f */
int oracle_rnp_bond_f(double J, double r, double S, const double *mat, int irrev, int fs_in,
                      double *f) {
  bool fs = fs_in != 0;
  *f = rnpBondF(J, r, S, mat, irrev, fs);
  return fs ? 1 : 0;
}
/* PMB (law 2) / PdElastic (law 3) force part; returns fs after the call */
int oracle_bond_f(int law, double J, double S, double cc, double s0, int fs_in, double *f) {
  bool fs = fs_in != 0;
  *f = law == PDEM_MAT_PMB ? pmbBondF(J, S, cc, s0, fs) : pdelasticBondF(J, S, cc);
  return fs ? 1 : 0;
}
double oracle_pdstate_bond_f(double J, double r, double S, double K, double G, double mx,
                             double thetax) {
  return pdstateBondF(J, r, S, K, G, mx, thetax);
}
double oracle_influence(int type, double p0, double p1, double r) {
  oracle_ctx c;
  c.inf_type = type;
  c.inf_p[0] = p0;
  c.inf_p[1] = p1;
  return c.infFn(r);
}

}  // extern "C"
