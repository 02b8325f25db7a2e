// pdem_mlp.hpp -- the pair-contact closure network on MFMA.
//
// 16 -> 128 -> 128 -> 128 -> 4 perceptron with ReLU hidden layers.  Operands
// are BF16 (round to nearest even), products accumulate in FP32 on
// v_mfma_f32_16x16x16_bf16; biases, the master weights and the optimiser
// state are FP32.
//
// One wavefront evaluates 16 rows (samples).  Each layer is computed in
// transposed form, H_out^T = W^T H_in^T: the accumulator of a 16x16 output
// block then holds, in lane l, features 4(l/16)..4(l/16)+3 of row l%16 --
// exactly the lane layout of the B operand of the next layer's K-block, so
// activations stay in registers from layer to layer.
//
// MFMA 16x16x16 lane layout (wave64): A[i][k] in lane i + 16(k/4), element
// k%4; B[k][j] in lane j + 16(k/4), element k%4; D[i][j] in lane j + 16(i/4),
// element i%4.
//
// Training (k_mlp_grad, pdem_mlp.hip): a workgroup of 4 wavefronts takes 64 rows, runs the
// forward and backward passes per wavefront, stages the activations and the
// back-propagated pre-activation gradients feature-major in LDS, and forms
// the weight gradients with the rows as the MFMA K dimension.  Each workgroup
// writes its partial gradient; k_mlp_adam sums the partials in a fixed order
// and applies Adam to the FP32 master weights, then refreshes the BF16 copies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdem {
namespace mlp {

constexpr int NIN = 16, NH = 128, NOUT = 4, NOP = 16;  // NOP: output padded to one MFMA block
constexpr int ROWS_WG = 64;                             // rows per training workgroup
// FP32 master parameters, W row-major [in][out]
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH;
constexpr int OW2 = OB1 + NH, OB2 = OW2 + NH * NH;
constexpr int OW3 = OB2 + NH, OB3 = OW3 + NH * NH;
constexpr int OW4 = OB3 + NH, OB4 = OW4 + NH * NOUT;
constexpr int NPARAM = OB4 + NOUT;
// BF16 copies: per layer W [in][out_p] and Wt [out_p][in] (out_p = 16 for the last layer)
constexpr int BW1 = 0, BW2 = BW1 + NIN * NH, BW3 = BW2 + NH * NH, BW4 = BW3 + NH * NH;
constexpr int NBF = BW4 + NH * NOP;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct Net {
  const short* w;         // BF16 W copies (offsets BW*)
  const short* wt;        // BF16 Wt copies (offsets BW*)
  const float* p;         // FP32 master (biases read from here)
  const float* in_mean;   // feature normalisation: x_n = (x - mean) * scale
  const float* in_scale;
  const float* out_scale; // target normalisation: y_n = y * out_scale
};

// norm = in_mean[16] | in_scale[16] | out_scale[4] | out_weight[4]
inline Net make_net(const short* wb, const short* wtb, const float* p, const float* norm) {
  return Net{wb, wtb, p, norm, norm + NIN, norm + 2 * NIN};
}
constexpr int NNORM = 2 * NIN + 2 * NOUT;

__device__ __forceinline__ short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_f(short h) {
  return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}
__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// lane's fragment of a row-major [M][ld] BF16 matrix at block (m0, k0):
// row m0 + l%16, columns k0 + 4(l/16) .. +3 (A operand; B operand of a
// matrix stored transposed)
__device__ __forceinline__ s4 frag(const short* A, int ld, int m0, int k0, int lane) {
  return *(const s4*)(A + (size_t)(m0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}
__device__ __forceinline__ s4 relu_bf16(f4 a) {
  s4 o;
  for (int r = 0; r < 4; r++) o[r] = bf16_rne(a[r] > 0.f ? a[r] : 0.f);
  return o;
}

// input fragment of row `row` (features 4(l/16)..+3), normalised and rounded
__device__ __forceinline__ s4 input_frag(const Net& n, const float* x16, int lane) {
  const int q = lane >> 4;
  s4 o;
  for (int r = 0; r < 4; r++) {
    const int f = 4 * q + r;
    o[r] = bf16_rne((x16[f] - n.in_mean[f]) * n.in_scale[f]);
  }
  return o;
}

// hidden layer with K = 128: out[ob] = relu(b + sum_kb Wt[ob][kb] h[kb])
__device__ __forceinline__ void hidden(const Net& n, int bw, int ob_off, const s4 (&hin)[8], s4 (&hout)[8],
                                       int lane) {
  const int q = lane >> 4;
  for (int ob = 0; ob < 8; ob++) {
    f4 acc;
    for (int r = 0; r < 4; r++) acc[r] = n.p[ob_off + ob * 16 + 4 * q + r];
    for (int kb = 0; kb < 8; kb++) acc = mfma(frag(n.wt + bw, NH, ob * 16, kb * 16, lane), hin[kb], acc);
    hout[ob] = relu_bf16(acc);
  }
}

// Forward pass of one wavefront's 16 rows.  xb: input fragment.  Returns the
// output fragment: lanes 0..15 hold outputs 0..3 (normalised) of row lane.
__device__ __forceinline__ f4 forward(const Net& n, s4 xb, s4 (&h1)[8], s4 (&h2)[8], s4 (&h3)[8], int lane) {
  const int q = lane >> 4;
  for (int ob = 0; ob < 8; ob++) {
    f4 acc;
    for (int r = 0; r < 4; r++) acc[r] = n.p[OB1 + ob * 16 + 4 * q + r];
    acc = mfma(frag(n.wt + BW1, NIN, ob * 16, 0, lane), xb, acc);
    h1[ob] = relu_bf16(acc);
  }
  hidden(n, BW2, OB2, h1, h2, lane);
  hidden(n, BW3, OB3, h2, h3, lane);
  f4 y;
  for (int r = 0; r < 4; r++) y[r] = q == 0 ? n.p[OB4 + r] : 0.f;
  for (int kb = 0; kb < 8; kb++) y = mfma(frag(n.wt + BW4, NH, 0, kb * 16, lane), h3[kb], y);
  return y;
}

// back-propagation through a hidden layer: dz_in[ib] = mask(h_in[ib]) * sum_ob W[ib][ob] dz_out[ob]
__device__ __forceinline__ void back_hidden(const Net& n, int bw, const s4 (&dz_out)[8], const s4 (&h_in)[8],
                                            s4 (&dz_in)[8], int lane) {
  for (int ib = 0; ib < 8; ib++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ob = 0; ob < 8; ob++) acc = mfma(frag(n.w + bw, NH, ib * 16, ob * 16, lane), dz_out[ob], acc);
    s4 o;
    for (int r = 0; r < 4; r++) o[r] = bf16_f(h_in[ib][r]) > 0.f ? bf16_rne(acc[r]) : (short)0;
    dz_in[ib] = o;
  }
}

// store a fragment array feature-major into LDS: T[f][row], 64 rows per workgroup
__device__ __forceinline__ void store_fm(short* T, const s4* fr, int nblk, int wave, int lane) {
  const int q = lane >> 4, j = lane & 15;
  for (int b = 0; b < nblk; b++)
    for (int r = 0; r < 4; r++) T[(b * 16 + 4 * q + r) * ROWS_WG + wave * 16 + j] = fr[b][r];
}

// weight gradient block loop: dW[in][out] = sum_rows HT[in][row] dZT[out][row];
// blocks split over the 4 wavefronts; out_real limits the columns written.
__device__ __forceinline__ void weight_grad(const short* HT, const short* dZT, int nin, int nout_p, int nout,
                                            float* gw, int wave, int lane) {
  const int q = lane >> 4, j = lane & 15;
  const int nbi = nin / 16, nbo = nout_p / 16;
  for (int blk = wave; blk < nbi * nbo; blk += 4) {
    const int ib = blk / nbo, ob = blk % nbo;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < ROWS_WG / 16; ks++)
      acc = mfma(frag(HT, ROWS_WG, ib * 16, ks * 16, lane), frag(dZT, ROWS_WG, ob * 16, ks * 16, lane), acc);
    const int o = ob * 16 + j;
    if (o < nout)
      for (int r = 0; r < 4; r++) gw[(ib * 16 + 4 * q + r) * nout + o] = acc[r];
  }
}
__device__ __forceinline__ void bias_grad(const short* dZT, int nout, float* gb) {
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < ROWS_WG; r++) s += bf16_f(dZT[o * ROWS_WG + r]);
    gb[o] = s;
  }
}

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

}  // namespace mlp
}  // namespace pdem
