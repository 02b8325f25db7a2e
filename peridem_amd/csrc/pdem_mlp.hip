// pdem_mlp.hip -- C-ABI of the closure network (pdem_mlp_* in include/pdem.h):
// creation and initialisation, parameter and normalisation exchange,
// inference and gradients on host buffers, and the training loop (minibatch
// Adam on the device, one gradient launch and one update launch per step).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "pdem.h"
#include "pdem_host.hpp"
#include "pdem_mlp.hpp"

using namespace pdem;

namespace pdem {
namespace mlp {

// Gradient of loss = (1/B) sum_rows sum_o tw[o] (y_o - t_o)^2 over a batch of
// B rows.  Rows: sample index = rows in order (shuffle == 0, rows >= n
// inactive) or drawn with replacement by a hash of (shuffle, row).
// partial: per workgroup NPARAM gradient entries + 1 loss entry.
__global__ void __launch_bounds__(256) k_mlp_grad(Net n, const float* X, const float* T, const float* tw,
                                                  int64_t nsamp, int64_t batch, uint64_t shuffle,
                                                  float inv_batch, float* partial) {
  __shared__ short XT[NIN * ROWS_WG];
  __shared__ short HT[3][NH * ROWS_WG];
  __shared__ short DYT[NOP * ROWS_WG];
  __shared__ float lsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
  const int64_t brow = (int64_t)blockIdx.x * ROWS_WG + wave * 16 + j;
  const bool active = brow < batch;
  int64_t srow = 0;
  if (active) srow = shuffle ? (int64_t)(splitmix(shuffle * 0x100000001b3ull + (uint64_t)brow) % (uint64_t)nsamp) : brow;
  float x16[NIN];
  for (int f = 0; f < NIN; f++) x16[f] = active ? X[srow * NIN + f] : n.in_mean[f];
  const s4 xb = input_frag(n, x16, lane);
  s4 h1[8], h2[8], h3[8];
  const f4 y = forward(n, xb, h1, h2, h3, lane);
  // output gradient (lanes q == 0 hold the 4 outputs)
  s4 dy;
  float l = 0.f;
  for (int r = 0; r < 4; r++) {
    float d = 0.f;
    if (q == 0 && active) {
      const float e = y[r] - T[srow * NOUT + r] * n.out_scale[r];
      l += tw[r] * e * e;
      d = 2.f * tw[r] * e * inv_batch;
    }
    dy[r] = bf16_rne(d);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if (lane == 0) lsum[wave] = l;
  store_fm(XT, &xb, 1, wave, lane);
  store_fm(HT[0], h1, 8, wave, lane);
  store_fm(HT[1], h2, 8, wave, lane);
  store_fm(HT[2], h3, 8, wave, lane);
  store_fm(DYT, &dy, 1, wave, lane);
  __syncthreads();
  float* g = partial + (size_t)blockIdx.x * (NPARAM + 1);
  if (threadIdx.x == 0) g[NPARAM] = (lsum[0] + lsum[1] + lsum[2] + lsum[3]) * inv_batch;
  weight_grad(HT[2], DYT, NH, NOP, NOUT, g + OW4, wave, lane);
  bias_grad(DYT, NOUT, g + OB4);
  // dz3 = mask(h3) * W4 dy
  s4 dz[8], dzn[8];
  for (int ib = 0; ib < 8; ib++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(frag(n.w + BW4, NOP, ib * 16, 0, lane), dy, acc);
    s4 o;
    for (int r = 0; r < 4; r++) o[r] = bf16_f(h3[ib][r]) > 0.f ? bf16_rne(acc[r]) : (short)0;
    dz[ib] = o;
  }
  __syncthreads();  // H3 consumed: its LDS now holds dZ3
  store_fm(HT[2], dz, 8, wave, lane);
  __syncthreads();
  weight_grad(HT[1], HT[2], NH, NH, NH, g + OW3, wave, lane);
  bias_grad(HT[2], NH, g + OB3);
  back_hidden(n, BW3, dz, h2, dzn, lane);
  __syncthreads();
  store_fm(HT[1], dzn, 8, wave, lane);
  __syncthreads();
  weight_grad(HT[0], HT[1], NH, NH, NH, g + OW2, wave, lane);
  bias_grad(HT[1], NH, g + OB2);
  back_hidden(n, BW2, dzn, h1, dz, lane);
  __syncthreads();
  store_fm(HT[0], dz, 8, wave, lane);
  __syncthreads();
  weight_grad(XT, HT[0], NIN, NH, NH, g + OW1, wave, lane);
  bias_grad(HT[0], NH, g + OB1);
}

// bf16 copies of one master parameter
__device__ __forceinline__ void pack_one(int i, float w, short* wb, short* wtb) {
  int bw, nin, nout_p, nout, k;
  if (i < OB1) { bw = BW1; nin = NIN; nout_p = NH; nout = NH; k = i - OW1; }
  else if (i >= OW2 && i < OB2) { bw = BW2; nin = NH; nout_p = NH; nout = NH; k = i - OW2; }
  else if (i >= OW3 && i < OB3) { bw = BW3; nin = NH; nout_p = NH; nout = NH; k = i - OW3; }
  else if (i >= OW4 && i < OB4) { bw = BW4; nin = NH; nout_p = NOP; nout = NOUT; k = i - OW4; }
  else return;  // biases stay FP32
  const int a = k / nout, b = k % nout;
  const short h = bf16_rne(w);
  wb[bw + a * nout_p + b] = h;
  wtb[bw + b * nin + a] = h;
}

__global__ void k_mlp_pack(const float* p, short* wb, short* wtb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NPARAM) pack_one(i, p[i], wb, wtb);
}

// Sum of the workgroup partials (fixed order) and one Adam step.
__global__ void k_mlp_adam(float* p, float* m, float* v, short* wb, short* wtb, const float* partial,
                           int nwg, float lr, float b1, float b2, float eps, float c1, float c2, float* grad_out,
                           float* loss_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > NPARAM) return;
  float g = 0.f;
  for (int w = 0; w < nwg; w++) g += partial[(size_t)w * (NPARAM + 1) + i];
  if (i == NPARAM) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (grad_out) grad_out[i] = g;
  if (lr == 0.f) return;
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float w = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = w;
  pack_one(i, w, wb, wtb);
}

// Inference on host-provided rows: Y[n][4] de-normalised (y / out_scale).
__global__ void __launch_bounds__(256) k_mlp_infer(Net n, const float* X, int64_t nrows, float* Y) {
  const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
  const int64_t row = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16 + j;
  const bool active = row < nrows;
  float x16[NIN];
  for (int f = 0; f < NIN; f++) x16[f] = active ? X[row * NIN + f] : n.in_mean[f];
  s4 h1[8], h2[8], h3[8];
  const f4 y = forward(n, input_frag(n, x16, lane), h1, h2, h3, lane);
  if (q == 0 && active)
    for (int r = 0; r < 4; r++) Y[row * NOUT + r] = y[r] / n.out_scale[r];
}

}  // namespace mlp
}  // namespace pdem

namespace {

thread_local std::string g_mlp_create_err;

template <class F>
int mguard(pdem_mlp* m, F f) {
  if (!m) return -1;
  try {
    HIPCHK(hipSetDevice(m->device));
    f();
    return 0;
  } catch (const HipError& e) {
    m->err = e.msg;
  } catch (const std::exception& e) {
    m->err = e.what();
  }
  return -1;
}

mlp::Net net_of(const pdem_mlp* m) { return mlp::make_net(m->wb, m->wtb, m->p, m->norm); }

void ensure_partial(pdem_mlp* m, int64_t nwg) {
  if (nwg <= m->partial_wg) return;
  if (m->partial) HIPCHK(hipFree(m->partial));
  m->partial = nullptr;
  HIPCHK(hipMalloc(&m->partial, (size_t)nwg * (mlp::NPARAM + 1) * sizeof(float)));
  m->partial_wg = nwg;
}

void pack(pdem_mlp* m) {
  hipLaunchKernelGGL(mlp::k_mlp_pack, dim3(blocks(mlp::NPARAM)), dim3(TPB), 0, m->stream, m->p, m->wb, m->wtb);
  HIPCHK(hipGetLastError());
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { HIPCHK(hipMalloc(&p, bytes ? bytes : 4)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const { return (T*)p; }
};

}  // namespace

extern "C" {

int64_t pdem_mlp_param_count(void) { return mlp::NPARAM; }

const char* pdem_mlp_last_error(const pdem_mlp* m) { return m ? m->err.c_str() : g_mlp_create_err.c_str(); }

int pdem_mlp_create(uint64_t seed, pdem_mlp** out) {
  if (!out) return -1;
  *out = nullptr;
  pdem_mlp* m = new pdem_mlp();
  try {
    HIPCHK(hipGetDevice(&m->device));
    HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&m->p, mlp::NPARAM * sizeof(float)));
    HIPCHK(hipMalloc(&m->m, mlp::NPARAM * sizeof(float)));
    HIPCHK(hipMalloc(&m->v, mlp::NPARAM * sizeof(float)));
    HIPCHK(hipMalloc(&m->wb, mlp::NBF * sizeof(short)));
    HIPCHK(hipMalloc(&m->wtb, mlp::NBF * sizeof(short)));
    HIPCHK(hipMalloc(&m->norm, mlp::NNORM * sizeof(float)));
    // zeroed on the network's (non-blocking) stream, so that its kernels see it
    HIPCHK(hipMemsetAsync(m->m, 0, mlp::NPARAM * sizeof(float), m->stream));
    HIPCHK(hipMemsetAsync(m->v, 0, mlp::NPARAM * sizeof(float), m->stream));
    // padded rows/columns of the last layer's BF16 copies stay zero
    HIPCHK(hipMemsetAsync(m->wb, 0, mlp::NBF * sizeof(short), m->stream));
    HIPCHK(hipMemsetAsync(m->wtb, 0, mlp::NBF * sizeof(short), m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    // He-uniform hidden layers, a small output layer, zero biases
    std::vector<float> p(mlp::NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int nin, int nout, float gain) {
      const float a = gain * std::sqrt(6.f / (float)nin);
      std::uniform_real_distribution<float> u(-a, a);
      for (int k = 0; k < nin * nout; k++) p[off + k] = u(rng);
    };
    fill(mlp::OW1, mlp::NIN, mlp::NH, 1.f);
    fill(mlp::OW2, mlp::NH, mlp::NH, 1.f);
    fill(mlp::OW3, mlp::NH, mlp::NH, 1.f);
    fill(mlp::OW4, mlp::NH, mlp::NOUT, 0.1f);
    HIPCHK(hipMemcpy(m->p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> nrm(mlp::NNORM, 1.f);
    for (int f = 0; f < mlp::NIN; f++) nrm[f] = 0.f;
    HIPCHK(hipMemcpy(m->norm, nrm.data(), nrm.size() * sizeof(float), hipMemcpyHostToDevice));
    pack(m);
    HIPCHK(hipStreamSynchronize(m->stream));
  } catch (const HipError& e) {
    g_mlp_create_err = e.msg;
    delete m;
    return -1;
  }
  *out = m;
  return 0;
}

int pdem_mlp_destroy(pdem_mlp* m) {
  if (!m) return 0;
  if (const int u = m->users.load()) {
    m->err = "network still enabled in " + std::to_string(u) +
             " engine context(s): call pdem_closure_disable or pdem_destroy first";
    return -1;
  }
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* q : {(void*)m->p, (void*)m->m, (void*)m->v, (void*)m->wb, (void*)m->wtb, (void*)m->norm,
                  (void*)m->partial})
    if (q) (void)hipFree(q);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int pdem_mlp_get_params(pdem_mlp* m, float* params) {
  return mguard(m, [&] {
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(params, m->p, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int pdem_mlp_set_params(pdem_mlp* m, const float* params) {
  return mguard(m, [&] {
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(m->p, params, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(m->m, 0, mlp::NPARAM * sizeof(float), m->stream));
    HIPCHK(hipMemsetAsync(m->v, 0, mlp::NPARAM * sizeof(float), m->stream));
    m->adam_step = 0;
    pack(m);
    HIPCHK(hipStreamSynchronize(m->stream));
  });
}

int pdem_mlp_set_normalization(pdem_mlp* m, const float* in_mean, const float* in_scale,
                               const float* out_scale, const float* out_weight) {
  return mguard(m, [&] {
    std::vector<float> nrm(mlp::NNORM);
    for (int f = 0; f < mlp::NIN; f++) {
      nrm[f] = in_mean[f];
      nrm[mlp::NIN + f] = in_scale[f];
    }
    for (int o = 0; o < mlp::NOUT; o++) {
      if (!(out_scale[o] > 0.f) || !(out_weight[o] >= 0.f)) throw HipError{"bad output scale or weight"};
      nrm[2 * mlp::NIN + o] = out_scale[o];
      nrm[2 * mlp::NIN + mlp::NOUT + o] = out_weight[o];
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(m->norm, nrm.data(), nrm.size() * sizeof(float), hipMemcpyHostToDevice));
  });
}

int pdem_mlp_get_normalization(pdem_mlp* m, float* norm40) {
  return mguard(m, [&] {
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(norm40, m->norm, mlp::NNORM * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int pdem_mlp_forward(pdem_mlp* m, const float* x, int64_t n, float* y) {
  return mguard(m, [&] {
    if (n < 0) throw HipError{"negative row count"};
    if (n == 0) return;
    DevBuf dx(n * mlp::NIN * sizeof(float)), dy(n * mlp::NOUT * sizeof(float));
    HIPCHK(hipMemcpyAsync(dx.p, x, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(mlp::k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, m->stream, net_of(m),
                       dx.as<float>(), n, dy.as<float>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, dy.p, n * mlp::NOUT * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
  });
}

int pdem_mlp_gradient(pdem_mlp* m, const float* x, const float* t, int64_t n, float* grad, float* loss) {
  return mguard(m, [&] {
    if (n <= 0) throw HipError{"gradient needs at least one row"};
    const int64_t nwg = (n + mlp::ROWS_WG - 1) / mlp::ROWS_WG;
    ensure_partial(m, nwg);
    DevBuf dx(n * mlp::NIN * sizeof(float)), dt(n * mlp::NOUT * sizeof(float));
    DevBuf dg((mlp::NPARAM + 1) * sizeof(float));
    HIPCHK(hipMemcpyAsync(dx.p, x, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(dt.p, t, n * mlp::NOUT * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const mlp::Net net = net_of(m);
    hipLaunchKernelGGL(mlp::k_mlp_grad, dim3((unsigned)nwg), dim3(256), 0, m->stream, net, dx.as<float>(),
                       dt.as<float>(), m->norm + 2 * mlp::NIN + mlp::NOUT, n, n, (uint64_t)0,
                       1.f / (float)n, m->partial);
    HIPCHK(hipGetLastError());
    float* g = dg.as<float>();
    hipLaunchKernelGGL(mlp::k_mlp_adam, dim3(blocks(mlp::NPARAM + 1)), dim3(TPB), 0, m->stream, m->p, m->m, m->v,
                       m->wb, m->wtb, m->partial, (int)nwg, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, g, g + mlp::NPARAM);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, g, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(loss, g + mlp::NPARAM, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
  });
}

int pdem_mlp_train(pdem_mlp* m, const float* x, const float* t, int64_t n, int32_t iters, int32_t batch,
                   float lr, uint64_t seed, float* loss_curve) {
  return mguard(m, [&] {
    if (n <= 0 || iters <= 0 || batch <= 0) throw HipError{"train needs rows, iterations and a batch"};
    if (!(lr > 0.f)) throw HipError{"learning rate must be positive"};
    const int64_t nwg = (batch + mlp::ROWS_WG - 1) / mlp::ROWS_WG;
    ensure_partial(m, nwg);
    DevBuf dx(n * mlp::NIN * sizeof(float)), dt(n * mlp::NOUT * sizeof(float)), dl((size_t)iters * sizeof(float));
    HIPCHK(hipMemcpyAsync(dx.p, x, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(dt.p, t, n * mlp::NOUT * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const mlp::Net net = net_of(m);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const double pi = 3.14159265358979323846;
    for (int32_t it = 0; it < iters; it++) {
      m->adam_step++;
      // cosine decay to 2 % of the initial rate
      const float lr_t = lr * (float)(0.02 + 0.98 * 0.5 * (1.0 + std::cos(pi * it / iters)));
      const float c1 = 1.f / (1.f - std::pow(b1, (float)m->adam_step));
      const float c2 = 1.f / (1.f - std::pow(b2, (float)m->adam_step));
      const uint64_t sh = (seed * 0x2545f4914f6cdd1dull) ^ (uint64_t)(it + 1);
      hipLaunchKernelGGL(mlp::k_mlp_grad, dim3((unsigned)nwg), dim3(256), 0, m->stream, net, dx.as<float>(),
                         dt.as<float>(), m->norm + 2 * mlp::NIN + mlp::NOUT, n, (int64_t)batch, sh ? sh : 1,
                         1.f / (float)batch, m->partial);
      hipLaunchKernelGGL(mlp::k_mlp_adam, dim3(blocks(mlp::NPARAM + 1)), dim3(TPB), 0, m->stream, m->p, m->m,
                         m->v, m->wb, m->wtb, m->partial, (int)nwg, lr_t, b1, b2, eps, c1, c2, (float*)nullptr,
                         dl.as<float>() + it);
    }
    HIPCHK(hipGetLastError());
    if (loss_curve)
      HIPCHK(hipMemcpyAsync(loss_curve, dl.p, (size_t)iters * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
  });
}

}  // extern "C"
