// mmf_learn.hip -- learned face reconstruction on gfx950 (2D, order 2).
//
// The reference evaluates face values of its limited linear polynomial
// (reconstruction.cpp:67-97: mean +/- slope/2, mean when the face density or
// pressure is not positive).  The learned form replaces those face values by
// the output of a small network evaluated on the matrix cores and keeps the
// limited linear values as the fallback:
//
//   k_recon_learned  one block = 16x4 cells; the 20x8 region (2-cell halo) of
//                    the stage state is staged in LDS; per cell the 5x5
//                    stencil of rho, rho u, rho v, rho E is normalised
//                    (centre c, scale s per component) and fed, 16 cells per
//                    wave, through the 100-64-64-16 network with bf16
//                    v_mfma_f32_16x16x16 and fp32 accumulation.  Output z of
//                    face f, component v sets the face value P + sigmoid(z)
//                    (Q - P) between the minmod value P and the fifth-order
//                    upwind-biased linear value Q (a learned limiter).  The network
//                    is applied transposed (H^T = W X^T): the accumulator of
//                    one layer is, lane for lane, the B operand of the next,
//                    so activations never leave registers.  A cell falls back
//                    to the plain face values when its stencil crosses a
//                    non-periodic border, when the shock sensor fires, or
//                    when a learned face state has non-positive density or
//                    pressure.  Face states go to a [face][component] SoA.
//   k_flux_fs        four lanes per cell, one face each, from the face states
//                    (border states from the BC), Rusanov flux, residual in
//                    the reference face order, fused RK stage update.  With
//                    every cell on the fallback this is bit-identical to the
//                    order-2 gather kernel.
//
// Training (fp32, deterministic): k_train evaluates forward and backward for
// 32 rows per block with the weights staged in LDS and writes per-block
// gradient partials; k_reduce sums them in block order; k_adam updates.
// k_collect builds training rows from a fine run: coarse stencils by
// conservative averaging, targets from the fine plain face values, eight
// dihedral images of each row.
#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_learn.hpp"
#include "mmf_stage.hpp"

namespace mmf {

using namespace learn;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// conservative component of stencil variable v (rho, rho u, rho v, rho E)
__device__ __forceinline__ int var_of(int v) { return v == 3 ? FID_RHO_E : v; }

// logistic function 1 / (1 + e^-z) from the hardware exp2 and reciprocal
// (v_exp_f32, v_rcp_f32: about 1 ulp each); z -> -inf gives 0, +inf gives 1
__device__ __forceinline__ float sigm_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}

// GELU in its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
// = z sigmoid(v), v = 2 sqrt(2/pi) (z + 0.044715 z^3) (Hendrycks & Gimpel,
// "Gaussian Error Linear Units", 2016); within 4.8e-4 of the erf form.
constexpr float GELU_K = 1.5957691216057308f, GELU_C = 0.044715f;
__device__ __forceinline__ float gelu_f(float z) {
  // z / (1 + 2^(z (a + b z^2))), a = -K log2(e), b = -K C log2(e)
  const float e = __builtin_amdgcn_exp2f(z * fmaf(-GELU_K * GELU_C * 1.4426950408889634f, z * z,
                                                   -GELU_K * 1.4426950408889634f));
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float gelu_d(float z) {
  const float sg = sigm_f(GELU_K * (z + GELU_C * (z * z * z)));
  return sg + z * (sg * (1.0f - sg)) * (GELU_K * (1.0f + 3.0f * GELU_C * (z * z)));
}

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// four fp32 values -> one bf16 fragment (two packed conversions)
__device__ __forceinline__ s4 bf4(float a, float b, float c, float d) {
  const f2v lo = {a, b}, hi = {c, d};
  const uint32_t l = __builtin_bit_cast(uint32_t, __builtin_convertvector(lo, bf2));
  const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(hi, bf2));
  s4 r;
  r[0] = (short)(l & 0xFFFFu);
  r[1] = (short)(l >> 16);
  r[2] = (short)(h & 0xFFFFu);
  r[3] = (short)(h >> 16);
  return r;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 16 cells per wave through the network on the matrix cores.  Lane l holds
// cell l%16 and input group g = l/16: xb[kb] = inputs 16 kb + 4 g + (0..3).
// Returns y[i] = output 4 g + i (face g, variable i) of cell l%16.  The
// biases (B: b1 (64), b2 (64), b3 (16)) initialise the accumulators.
__device__ __forceinline__ f4 mlp16(const s4 *F, const float *B, const s4 *xb, int lane) {
  const int g = lane >> 4;
  s4 h1[4], h2[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const float *b = B + 16 * mt + 4 * g;
    f4 acc = {b[0], b[1], b[2], b[3]};  // the biases start the sums
#pragma unroll
    for (int kb = 0; kb < 7; ++kb)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F[(mt * 7 + kb) * 64 + lane], xb[kb], acc, 0, 0, 0);
    h1[mt] = bf4(gelu_f(acc[0]), gelu_f(acc[1]), gelu_f(acc[2]), gelu_f(acc[3]));
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const float *b = B + NH + 16 * mt + 4 * g;
    f4 acc = {b[0], b[1], b[2], b[3]};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F[(28 + mt * 4 + kb) * 64 + lane], h1[kb], acc, 0, 0, 0);
    h2[mt] = bf4(gelu_f(acc[0]), gelu_f(acc[1]), gelu_f(acc[2]), gelu_f(acc[3]));
  }
  const float *b = B + 2 * NH + 4 * g;
  f4 y = {b[0], b[1], b[2], b[3]};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F[(44 + kb) * 64 + lane], h2[kb], y, 0, 0, 0);
  return y;
}

__device__ __forceinline__ float minmodf(float a, float b) {
  if (a > 0.f && b > 0.f) return (a < b) ? a : b;
  if (a < 0.f && b < 0.f) return (a > b) ? a : b;
  return 0.f;
}

// Limiter candidates of output o from a normalised stencil row: P = minmod
// face value, Q = fifth-order upwind-biased linear face value, along the
// face's axis with offsets counted towards the face.
__device__ __forceinline__ void candidates_norm(const float *x, int o, float &P, float &Q) {
  const int f = o >> 2, v = o & 3, a = f >> 1, side = f & 1;
  float u[5];
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    const int dd = side ? d : -d;
    u[d + 2] = a ? x[v * 25 + (dd + 2) * 5 + 2] : x[v * 25 + 2 * 5 + (dd + 2)];
  }
  P = u[2] + 0.5f * minmodf(u[3] - u[2], u[2] - u[1]);
  Q = ((((2.0f * u[0] - 13.0f * u[1]) + 47.0f * u[2]) + 27.0f * u[3]) - 3.0f * u[4]) / 60.0f;
}

}  // namespace

// ---- weight packing ------------------------------------------------------------
// Fragment fi (layer 1: mt*7+kb; layer 2: 28+mt*4+kb; layer 3: 44+kb), lane l,
// element i holds W[16 mt + l%16][k] with, for layers 2 and 3, k = 16 kb +
// 4 (l/16) + i.  Layer 1 reads its inputs in a permuted order: lane group
// g = l/16 carries stencil variable g, stencil position r = 4 kb + i (zero
// for r >= 25), i.e. input k = 25 g + r of the stored row layout.  The B
// operand built by the kernels follows the same order, so the product is the
// network's; each lane then only touches its own variable's stencil.
__global__ __launch_bounds__(64) void k_learn_pack(const float *P, short *frags) {
  const int fi = blockIdx.x, lane = threadIdx.x;
  int off, K, mt, kb;
  if (fi < 28) {
    off = OFF_W1; K = NIN; mt = fi / 7; kb = fi % 7;
  } else if (fi < 44) {
    off = OFF_W2; K = NH; mt = (fi - 28) / 4; kb = (fi - 28) % 4;
  } else {
    off = OFF_W3; K = NH; mt = 0; kb = fi - 44;
  }
  const int m = 16 * mt + (lane & 15);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float w;
    if (fi < 28) {
      const int r = 4 * kb + i;
      w = (r < 25) ? P[off + m * K + 25 * (lane >> 4) + r] : 0.0f;
    } else {
      const int k = 16 * kb + 4 * (lane >> 4) + i;
      w = (k < K) ? P[off + m * K + k] : 0.0f;
    }
    frags[(fi * 64 + lane) * 4 + i] = f2bf(w);
  }
}

hipError_t launch_learn_pack(const float *params, short *frags, hipStream_t s) {
  hipLaunchKernelGGL(k_learn_pack, dim3(NFRAG), dim3(64), 0, s, params, frags);
  return hipGetLastError();
}

// ---- inference on given rows (tests, offline use) -------------------------------
__global__ __launch_bounds__(64) void k_learn_forward_mfma(const float *P, const s4 *F, const float *X, int64_t n,
                                                           float *Y) {
  __shared__ float sB[2 * NH + NOUT];
  const int lane = threadIdx.x, c = lane & 15, g = lane >> 4;
  for (int e = lane; e < 2 * NH + NOUT; e += 64)
    sB[e] = (e < NH) ? P[OFF_B1 + e] : (e < 2 * NH) ? P[OFF_B2 + e - NH] : P[OFF_B3 + e - 2 * NH];
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * 16 + c;
  const bool live = row < n;
  s4 xb[7];
#pragma unroll
  for (int kb = 0; kb < 7; ++kb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * kb + i;  // stencil position of variable g
      xb[kb][i] = f2bf((live && r < 25) ? X[row * NIN + 25 * g + r] : 0.0f);
    }
  const f4 y = mlp16(F, sB, xb, lane);
  if (live) {
#pragma unroll
    for (int i = 0; i < 4; ++i) Y[row * NOUT + 4 * g + i] = y[i];
  }
}

__global__ __launch_bounds__(256) void k_learn_forward_f32(const float *P, const float *X, int64_t n, float *Y) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const float *x = X + row * NIN;
  float h1[NH], h2[NH];
  for (int j = 0; j < NH; ++j) {
    float z = 0.f;
    for (int i = 0; i < NIN; ++i) z += P[OFF_W1 + j * NIN + i] * x[i];
    h1[j] = gelu_f(z + P[OFF_B1 + j]);
  }
  for (int j = 0; j < NH; ++j) {
    float z = 0.f;
    for (int i = 0; i < NH; ++i) z += P[OFF_W2 + j * NH + i] * h1[i];
    h2[j] = gelu_f(z + P[OFF_B2 + j]);
  }
  for (int o = 0; o < NOUT; ++o) {
    float z = 0.f;
    for (int i = 0; i < NH; ++i) z += P[OFF_W3 + o * NH + i] * h2[i];
    Y[row * NOUT + o] = z + P[OFF_B3 + o];
  }
}

hipError_t launch_learn_forward(const float *params, const short *frags, const float *X, int64_t n, int use_mfma,
                                float *Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (use_mfma)
    hipLaunchKernelGGL(k_learn_forward_mfma, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, s, params,
                       (const s4 *)frags, X, n, Y);
  else
    hipLaunchKernelGGL(k_learn_forward_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params, X, n, Y);
  return hipGetLastError();
}

// ---- training: forward + backward of TRAIN_ROWS rows per block ------------------
// Rows: r = block * TRAIN_ROWS + t for r < rows; row r reads sample r, or with
// nsamples > 0 a sample drawn by hashing (seed, r).  Loss sum w (y - t)^2 /
// (16 rows); each block writes its gradient partial and its loss sum.
constexpr int TR = TRAIN_ROWS;

__global__ __launch_bounds__(256) void k_learn_train(const float *P, const float *X, const float *T, const float *Wt,
                                                     int64_t rows, int64_t nsamples, uint64_t seed, float *partial,
                                                     float *ploss) {
  __shared__ float sW1[NH][NIN + 1];
  __shared__ float sW2[NH][NH + 1];
  __shared__ float sW3[NOUT][NH + 1];
  __shared__ float sB[2 * NH + NOUT];
  __shared__ float sx[TR][NIN];
  __shared__ float sz1[TR][NH], sa1[TR][NH], sz2[TR][NH], sa2[TR][NH];
  __shared__ float sdy[TR][NOUT], st[TR][NOUT], sw[TR];
  __shared__ float sd2[TR][NH], sd1[TR][NH];
  __shared__ float sred[256];
  __shared__ int64_t ssrc[TR];
  const int tid = threadIdx.x;
  for (int e = tid; e < NH * NIN; e += 256) sW1[e / NIN][e % NIN] = P[OFF_W1 + e];
  for (int e = tid; e < NH * NH; e += 256) sW2[e / NH][e % NH] = P[OFF_W2 + e];
  for (int e = tid; e < NOUT * NH; e += 256) sW3[e / NH][e % NH] = P[OFF_W3 + e];
  if (tid < 2 * NH + NOUT)
    sB[tid] = (tid < NH) ? P[OFF_B1 + tid] : (tid < 2 * NH) ? P[OFF_B2 + tid - NH] : P[OFF_B3 + tid - 2 * NH];
  if (tid < TR) {
    const int64_t r = (int64_t)blockIdx.x * TR + tid;
    int64_t src = -1;
    if (r < rows)
      src = (nsamples > 0) ? (int64_t)(mix64(seed + (uint64_t)r * 0x632BE59BD9B4E019ull) % (uint64_t)nsamples) : r;
    ssrc[tid] = src;
  }
  __syncthreads();
  for (int e = tid; e < TR * NIN; e += 256) {
    const int s = e / NIN, i = e % NIN;
    const int64_t src = ssrc[s];
    sx[s][i] = (src >= 0) ? X[src * NIN + i] : 0.0f;
  }
  for (int e = tid; e < TR * NOUT; e += 256) {
    const int s = e / NOUT, o = e % NOUT;
    const int64_t src = ssrc[s];
    st[s][o] = (src >= 0) ? T[src * NOUT + o] : 0.0f;
  }
  if (tid < TR) sw[tid] = (ssrc[tid] >= 0) ? Wt[ssrc[tid]] : 0.0f;
  __syncthreads();

  for (int e = tid; e < TR * NH; e += 256) {
    const int s = e >> 6, j = e & 63;
    float z = 0.f;
    for (int i = 0; i < NIN; ++i) z += sW1[j][i] * sx[s][i];
    z += sB[j];
    sz1[s][j] = z;
    sa1[s][j] = gelu_f(z);
  }
  __syncthreads();
  for (int e = tid; e < TR * NH; e += 256) {
    const int s = e >> 6, j = e & 63;
    float z = 0.f;
    for (int i = 0; i < NH; ++i) z += sW2[j][i] * sa1[s][i];
    z += sB[NH + j];
    sz2[s][j] = z;
    sa2[s][j] = gelu_f(z);
  }
  __syncthreads();
  const float scale = 2.0f / (float)(NOUT * rows);
  float lsum = 0.f;
  for (int e = tid; e < TR * NOUT; e += 256) {
    const int s = e >> 4, o = e & 15;
    float z = 0.f;
    for (int i = 0; i < NH; ++i) z += sW3[o][i] * sa2[s][i];
    z += sB[2 * NH + o];
    // model output: P + sigmoid(z) (Q - P)
    float Pn, Qn;
    candidates_norm(&sx[s][0], o, Pn, Qn);
    const float sg = sigm_f(z);
    const float r = (Pn + sg * (Qn - Pn)) - st[s][o];
    lsum += sw[s] * r * r;
    sdy[s][o] = scale * sw[s] * r * (sg * (1.0f - sg)) * (Qn - Pn);
  }
  sred[tid] = lsum;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) sred[tid] += sred[tid + o];
    __syncthreads();
  }
  if (tid == 0) ploss[blockIdx.x] = sred[0];

  float *G = partial + (int64_t)blockIdx.x * NPARAMS;
  for (int e = tid; e < NOUT * NH; e += 256) {
    const int o = e >> 6, i = e & 63;
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sdy[s][o] * sa2[s][i];
    G[OFF_W3 + e] = a;
  }
  if (tid < NOUT) {
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sdy[s][tid];
    G[OFF_B3 + tid] = a;
  }
  for (int e = tid; e < TR * NH; e += 256) {
    const int s = e >> 6, i = e & 63;
    float a = 0.f;
    for (int o = 0; o < NOUT; ++o) a += sW3[o][i] * sdy[s][o];
    sd2[s][i] = a * gelu_d(sz2[s][i]);
  }
  __syncthreads();
  for (int e = tid; e < NH * NH; e += 256) {
    const int j = e >> 6, i = e & 63;
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sd2[s][j] * sa1[s][i];
    G[OFF_W2 + e] = a;
  }
  if (tid < NH) {
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sd2[s][tid];
    G[OFF_B2 + tid] = a;
  }
  for (int e = tid; e < TR * NH; e += 256) {
    const int s = e >> 6, i = e & 63;
    float a = 0.f;
    for (int j = 0; j < NH; ++j) a += sW2[j][i] * sd2[s][j];
    sd1[s][i] = a * gelu_d(sz1[s][i]);
  }
  __syncthreads();
  for (int e = tid; e < NH * NIN; e += 256) {
    const int j = e / NIN, i = e % NIN;
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sd1[s][j] * sx[s][i];
    G[OFF_W1 + e] = a;
  }
  if (tid < NH) {
    float a = 0.f;
    for (int s = 0; s < TR; ++s) a += sd1[s][tid];
    G[OFF_B1 + tid] = a;
  }
}

int64_t learn_train_blocks(int64_t rows) { return (rows + TR - 1) / TR; }

hipError_t launch_learn_train(const float *params, const float *X, const float *T, const float *W, int64_t rows,
                              int64_t nsamples, uint64_t seed, float *partial, float *ploss, hipStream_t s) {
  hipLaunchKernelGGL(k_learn_train, dim3((unsigned)learn_train_blocks(rows)), dim3(256), 0, s, params, X, T, W, rows,
                     nsamples, seed, partial, ploss);
  return hipGetLastError();
}

// gradient = sum of the block partials: each of RSPLIT groups of blocks is
// summed in block order by one thread, the group sums in group order by a
// second; loss = block loss sums in block order, in fp64, / (16 rows)
constexpr int RSPLIT = 16;

__global__ __launch_bounds__(256) void k_learn_reduce(const float *partial, const float *ploss, int64_t nblocks,
                                                      int64_t rows, float *grad, double *loss) {
  __shared__ float part[RSPLIT][16];
  const int pl = threadIdx.x & 15, grp = threadIdx.x >> 4;  // 16 parameters x 16 groups per block
  const int p = blockIdx.x * 16 + pl;
  const int64_t per = (nblocks + RSPLIT - 1) / RSPLIT;
  const int64_t b0 = grp * per, b1 = (b0 + per < nblocks) ? b0 + per : nblocks;
  float a = 0.f;
  if (p < NPARAMS)
    for (int64_t b = b0; b < b1; ++b) a += partial[b * NPARAMS + p];
  part[grp][pl] = a;
  __syncthreads();
  if (grp == 0 && p < NPARAMS) {
    float t = 0.f;
    for (int g = 0; g < RSPLIT; ++g) t += part[g][pl];
    grad[p] = t;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && loss) {
    double l = 0.0;
    for (int64_t b = 0; b < nblocks; ++b) l += (double)ploss[b];
    *loss = l / (double)(NOUT * rows);
  }
}

hipError_t launch_learn_reduce(const float *partial, const float *ploss, int64_t nblocks, int64_t rows, float *grad,
                               double *loss, hipStream_t s) {
  hipLaunchKernelGGL(k_learn_reduce, dim3((NPARAMS + 15) / 16), dim3(256), 0, s, partial, ploss, nblocks, rows, grad,
                     loss);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_learn_adam(float *P, const float *G, float *M, float *V, float lr, float bc1,
                                                    float bc2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NPARAMS) return;
  const float g = G[p];
  const float m = 0.9f * M[p] + 0.1f * g;
  const float v = 0.999f * V[p] + 0.001f * (g * g);
  M[p] = m;
  V[p] = v;
  P[p] = P[p] - lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

hipError_t launch_learn_adam(float *params, const float *grad, float *m, float *v, float lr, float bc1, float bc2,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_learn_adam, dim3((NPARAMS + 255) / 256), dim3(256), 0, s, params, grad, m, v, lr, bc1, bc2);
  return hipGetLastError();
}

// ---- training rows from a fine run ------------------------------------------------
// One block per selected coarse cell of the grid restricted by 'factor'.
__global__ __launch_bounds__(256) void k_learn_collect(GridDev fg, const double *Uf, int r, const int64_t *cells,
                                                       double theta, float *X, float *T, float *W) {
  __shared__ double S[NF][5][5];
  __shared__ uint8_t sex[5][5];
  __shared__ double FT[4][8][NF];
  __shared__ double cs[4][2];
  __shared__ float xs[NIN], ts[NOUT];
  __shared__ float wsh;
  const int tid = threadIdx.x;
  const int64_t fs = fg.fs, nfx = fg.n[0];
  const int64_t ncx = fg.n[0] / r, ncy = fg.n[1] / r;
  const int64_t cell = cells[blockIdx.x];
  const int64_t I = cell % ncx, J = cell / ncx;
  if (tid < NF * 25) {
    const int k = tid / 25, rr = tid % 25, dj = rr / 5 - 2, di = rr % 5 - 2;
    int64_t ci = I + di, cj = J + dj;
    const bool ex = ((ci >= 0 && ci < ncx) || fg.periodic[0]) && ((cj >= 0 && cj < ncy) || fg.periodic[1]);
    ci = ((ci % ncx) + ncx) % ncx;
    cj = ((cj % ncy) + ncy) % ncy;
    double acc = 0.0;
    for (int b = 0; b < r; ++b)
      for (int a = 0; a < r; ++a) acc += Uf[k * fs + (r * cj + b) * nfx + (r * ci + a)];
    S[k][dj + 2][di + 2] = acc * (1.0 / (double)(r * r));
    if (k == 0) sex[dj + 2][di + 2] = ex;
  }
  if (tid >= 128 && tid < 128 + 4 * r) {
    const int q = tid - 128, f = q / r, b = q % r;
    const int a = f >> 1, side = f & 1;
    int64_t fq[3];
    if (a == 0) {
      fq[0] = r * I + (side ? r - 1 : 0);
      fq[1] = r * J + b;
    } else {
      fq[0] = r * I + b;
      fq[1] = r * J + (side ? r - 1 : 0);
    }
    fq[2] = 0;
    const int64_t idx = fq[0] + nfx * fq[1];
    int64_t nm = 0, np = 0;
    const bool em = neighbour(fg, fq, a, -1, nm);
    const bool ep = neighbour(fg, fq, a, +1, np);
    double X0[NF], v[NF], pr[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      X0[k] = Uf[k * fs + idx];
      const double dm = em ? X0[k] - Uf[k * fs + nm] : 0.0;
      const double dp = ep ? Uf[k * fs + np] - X0[k] : 0.0;
      const double sl = minmod(dp, dm);
      v[k] = side ? X0[k] + 0.5 * sl : X0[k] - 0.5 * sl;
    }
    cons2prim(v, pr);
    if (!(v[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
#pragma unroll
      for (int k = 0; k < NF; ++k) v[k] = X0[k];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) FT[f][b][k] = v[k];
  }
  __syncthreads();
  if (tid < 4) {
    const int k = var_of(tid);
    const double c = S[k][2][2];
    double dev = 0.0;
    for (int e = 0; e < 25; ++e) {
      const double d = fabs(S[k][e / 5][e % 5] - c);
      dev = (d > dev) ? d : dev;
    }
    const double tau = TAU_REL * (fabs(S[FID_RHO][2][2]) + fabs(S[FID_RHO_E][2][2]));
    cs[tid][0] = c;
    cs[tid][1] = (dev >= tau) ? dev : tau;
  }
  if (tid == 4) {
    bool all = true;
    for (int e = 0; e < 25; ++e) all = all && sex[e / 5][e % 5];
    // the shock sensor: largest relative jump |q_n - q0| / (q_n + q0) of rho
    // and p to the four face neighbours
    double phi = 0.0;
    for (int a = 0; a < 2; ++a) {
      const int mj = a ? 1 : 2, mi = a ? 2 : 1, pj = a ? 3 : 2, pi = a ? 2 : 3;
      for (int qn = 0; qn < 2; ++qn) {
        double q[3];
        const int jj[3] = {mj, 2, pj}, ii[3] = {mi, 2, pi};
        for (int t = 0; t < 3; ++t) {
          if (qn == 0) {
            q[t] = S[FID_RHO][jj[t]][ii[t]];
          } else {
            double c5[NF], p5[NF];
            for (int k = 0; k < NF; ++k) c5[k] = S[k][jj[t]][ii[t]];
            cons2prim(c5, p5);
            q[t] = p5[FID_P];
          }
        }
        for (int t = 0; t < 3; t += 2) {
          const double val = fabs(q[t] - q[1]) / (q[t] + q[1]);
          phi = (val > phi) ? val : phi;
        }
      }
    }
    wsh = (all && phi <= theta) ? 1.0f : 0.0f;
  }
  __syncthreads();
  if (tid < NIN) {
    const int v = tid / 25, rr = tid % 25;
    xs[tid] = (float)((S[var_of(v)][rr / 5][rr % 5] - cs[v][0]) / cs[v][1]);
  }
  if (tid >= 128 && tid < 128 + NOUT) {
    const int o = tid - 128, f = o >> 2, v = o & 3;
    double acc = 0.0;
    for (int b = 0; b < r; ++b) acc += FT[f][b][var_of(v)];
    ts[o] = (float)((acc * (1.0 / (double)r) - cs[v][0]) / cs[v][1]);
  }
  __syncthreads();
  // eight dihedral images: transpose (bit 0), then reflect x (bit 1), then
  // reflect y (bit 2); the source of each destination element is found by
  // undoing the steps in reverse order
  const int64_t base = (int64_t)blockIdx.x * 8;
  for (int e = tid; e < 8 * NIN; e += 256) {
    const int sg = e / NIN, d = e % NIN;
    int v = d / 25, dj = (d % 25) / 5, di = d % 5;
    float sgn = 1.0f;
    if (sg & 4) { if (v == 2) sgn = -sgn; dj = 4 - dj; }
    if (sg & 2) { if (v == 1) sgn = -sgn; di = 4 - di; }
    if (sg & 1) { v = (v == 1) ? 2 : (v == 2) ? 1 : v; const int t = dj; dj = di; di = t; }
    X[(base + sg) * NIN + d] = sgn * xs[v * 25 + dj * 5 + di];
  }
  for (int e = tid; e < 8 * NOUT; e += 256) {
    const int sg = e / NOUT, d = e % NOUT;
    int f = d >> 2, v = d & 3;
    float sgn = 1.0f;
    if (sg & 4) { if (v == 2) sgn = -sgn; f = (f == 2) ? 3 : (f == 3) ? 2 : f; }
    if (sg & 2) { if (v == 1) sgn = -sgn; f = (f == 0) ? 1 : (f == 1) ? 0 : f; }
    if (sg & 1) { f = (f + 2) & 3; v = (v == 1) ? 2 : (v == 2) ? 1 : v; }
    T[(base + sg) * NOUT + d] = sgn * ts[f * 4 + v];
  }
  if (tid < 8) W[base + tid] = wsh;
}

hipError_t launch_learn_collect(const GridDev &fine, const double *Uf, int factor, const int64_t *cells, int ncells,
                                double theta, float *X, float *T, float *W, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_learn_collect, dim3((unsigned)ncells), dim3(256), 0, s, fine, Uf, factor, cells, theta, X, T, W);
  return hipGetLastError();
}

// conservative average over factor x factor blocks -> AoS [ncoarse][5]
__global__ __launch_bounds__(256) void k_learn_restrict(GridDev fg, const double *Uf, int r, double *out) {
  const int64_t ncx = fg.n[0] / r, ncy = fg.n[1] / r;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ncx * ncy * NF) return;
  const int64_t cell = e / NF;
  const int k = (int)(e % NF);
  const int64_t I = cell % ncx, J = cell / ncx;
  double acc = 0.0;
  for (int b = 0; b < r; ++b)
    for (int a = 0; a < r; ++a) acc += Uf[k * fg.fs + (r * J + b) * fg.n[0] + (r * I + a)];
  out[e] = acc * (1.0 / (double)(r * r));
}

hipError_t launch_learn_restrict(const GridDev &fine, const double *Uf, int factor, double *coarse, hipStream_t s) {
  const int64_t n = (fine.n[0] / factor) * (fine.n[1] / factor) * NF;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_learn_restrict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fine, Uf, factor, coarse);
  return hipGetLastError();
}

// ---- reconstruction kernels ----------------------------------------------------------
// Stage state of an RX x RY region whose first cell is (xo, yo), wrapped on
// periodic axes; ex = 0 (and zero state) where a cell lies beyond a
// non-periodic border.  Stage 2 rebuilds W1 = U + dt R / V (main.cpp:421).
// With want_p, pre ([RY][RX]) receives the pressure of every cell (the shock
// sensor's).
template <int STAGE, int RX, int RY, int NT>
__device__ __forceinline__ void load_region(const GridDev &g, int xo, int yo, const double *U, const double *R,
                                            const double *Xin, double dt, double (&reg)[NF][RY][RX],
                                            uint8_t (&rex)[RY][RX], double *pre, bool want_p, int tid) {
  const int nx = (int)g.n[0], ny = (int)g.n[1];
  const int64_t fs = g.fs;
  for (int e = tid; e < RX * RY; e += NT) {
    const int ry = e / RX, rx = e % RX;
    int gx = xo + rx, gy = yo + ry;
    const bool ex = ((gx >= 0 && gx < nx) || g.periodic[0]) && ((gy >= 0 && gy < ny) || g.periodic[1]);
    if (gx < 0 || gx >= nx) gx = ((gx % nx) + nx) % nx;
    if (gy < 0 || gy >= ny) gy = ((gy % ny) + ny) % ny;
    rex[ry][rx] = ex;
    const int64_t i = gx + (int64_t)nx * gy;
    double v[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      v[k] = !ex ? 0.0 : (STAGE == 2) ? U[k * fs + i] + div_vol(g, dt * R[k * fs + i]) : Xin[k * fs + i];
      reg[k][ry][rx] = v[k];
    }
    // p = (gamma - 1) (rho E - |rho u|^2 / (2 rho)) for the shock sensor
    if (want_p) {
      const double m2 = (v[FID_RHO_U] * v[FID_RHO_U] + v[FID_RHO_V] * v[FID_RHO_V]) + v[FID_RHO_W] * v[FID_RHO_W];
      pre[ry * RX + rx] = ex ? (GAMMA - 1.0) * (v[FID_RHO_E] - 0.5 * m2 / v[FID_RHO]) : 0.0;
    }
  }
}

// Packed weight fragments and biases of the network into LDS.
template <int NT>
__device__ __forceinline__ void load_net(const LearnRun &run, s4 *sF, float *sB, int tid) {
  if (run.net.frags == nullptr) return;
  const s4 *gF = (const s4 *)run.net.frags;
  for (int e = tid; e < NFRAG * 64; e += NT) sF[e] = gF[e];
  const float *P = run.net.params;
  for (int e = tid; e < 2 * NH + NOUT; e += NT)
    sB[e] = (e < NH) ? P[OFF_B1 + e] : (e < 2 * NH) ? P[OFF_B2 + e - NH] : P[OFF_B3 + e - 2 * NH];
}

// Face grp = lane / 16 (0 x-, 1 x+, 2 y-, 3 y+) of the cell at region position
// (rx, ry); lane % 16 is the cell's column of the network.  Fv receives the
// learned value where the cell is learnable, its sensor is quiet and all
// four learned face states are positive, else the plain limited linear value
// (reconstruction.cpp:67-97).  The network runs when any lane of the wave is
// live (wave_live, uniform over the wave).  Returns whether the cell's faces
// are the learned ones.  Every cell's result depends only on its own
// stencil, so any kernel that places the cell in some lane gets the same bits.
template <bool NET, int RX, int RY>
__device__ __forceinline__ bool cell_face(const GridDev &g, const LearnRun &run, const double (&reg)[NF][RY][RX],
                                          const uint8_t (&rex)[RY][RX], const double *pre, const s4 *sF,
                                          const float *sB, int rx, int ry, bool live, bool learnable, bool wave_live,
                                          int lane, double *Fv) {
  const int c = lane & 15, grp = lane >> 4;
  const bool use_net = NET && run.net.frags != nullptr;
  bool fires = false;
  // this lane's stencil variable (lane group grp), its centre and 1 / scale
  const int kv = var_of(grp);
  double cv = 0.0;
  float inv_s = 0.0f;
  if (use_net) {
    cv = reg[kv][ry][rx];
    double dev = 0.0;
#pragma unroll
    for (int dj = -2; dj <= 2; ++dj)
#pragma unroll
      for (int di = -2; di <= 2; ++di) {
        dev = fmax(dev, fabs(reg[kv][ry + dj][rx + di] - cv));
      }
    const double tau = TAU_REL * (fabs(reg[FID_RHO][ry][rx]) + fabs(reg[FID_RHO_E][ry][rx]));
    inv_s = (float)(1.0 / ((dev >= tau) ? dev : tau));
    // shock sensor: lane group 0 on rho, group 1 on p; fires where the jump
    // to a face neighbour |q_n - q0| exceeds theta (q_n + q0) along either axis
    if (grp < 2) {
      const double *q = (grp == 0) ? &reg[FID_RHO][0][0] : pre;
      const double q0 = q[ry * RX + rx];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int d = -1; d <= 1; d += 2) {
          const double qn = a ? q[(ry + d) * RX + rx] : q[ry * RX + rx + d];
          const double num = fabs(qn - q0);
          const double den = qn + q0;
          fires = fires || !(den > 0.0) || !(num <= run.theta * den);
        }
      }
    }
  }

  // plain limited linear face value of face grp (reconstruction.cpp:67-97)
  auto plain_face = [&](double *P) {
    const int a = grp >> 1, side = grp & 1;
    const int my = ry - a, mx = rx - (1 - a), py = ry + a, px = rx + (1 - a);
    const bool em = rex[my][mx], ep = rex[py][px];
    double X0[NF], pr[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      X0[k] = reg[k][ry][rx];
      const double dm = em ? X0[k] - reg[k][my][mx] : 0.0;
      const double dp = ep ? reg[k][py][px] - X0[k] : 0.0;
      const double sl = minmod(dp, dm);
      P[k] = side ? X0[k] + 0.5 * sl : X0[k] - 0.5 * sl;
    }
    cons2prim(P, pr);
    if (!(P[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
#pragma unroll
      for (int k = 0; k < NF; ++k) P[k] = X0[k];
    }
  };
  double P[NF];
  const bool need_plain_all = run.blend != 1.0;
  if (need_plain_all) plain_face(P);

  bool pos = false;
  if (use_net && wave_live) {
    s4 xb[7];
#pragma unroll
    for (int kb = 0; kb < 7; ++kb) {
      float xv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * kb + i;  // stencil position, compile-time after unrolling
        xv[i] = 0.0f;
        if (r < 25 && live) xv[i] = (float)(reg[kv][ry - 2 + r / 5][rx - 2 + r % 5] - cv) * inv_s;
      }
      xb[kb] = bf4(xv[0], xv[1], xv[2], xv[3]);
    }
    const f4 y = mlp16(sF, sB, xb, lane);
    // face neighbour: the cell across face grp
    const int a = grp >> 1, side = grp & 1;
    const int ny_ = ry + (a ? (side ? 1 : -1) : 0), nx_ = rx + (a ? 0 : (side ? 1 : -1));
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = var_of(v);
      // learned limiter: between the minmod value and the fifth-order
      // upwind-biased value, offsets counted towards the face
      double u[5];
#pragma unroll
      for (int d = -2; d <= 2; ++d) {
        const int dd = side ? d : -d;
        u[d + 2] = a ? reg[k][ry + dd][rx] : reg[k][ry][rx + dd];
      }
      const double Pm = u[2] + 0.5 * minmod(u[3] - u[2], u[2] - u[1]);
      const double Q5 = ((((2.0 * u[0] - 13.0 * u[1]) + 47.0 * u[2]) + 27.0 * u[3]) - 3.0 * u[4]) * (1.0 / 60.0);
      const double sg = (double)sigm_f(y[v]);
      double f = Pm + sg * (Q5 - Pm);
      if (run.clip) {
        // between the means of the two cells sharing the face
        const double u0 = reg[k][ry][rx], u1 = reg[k][ny_][nx_];
        const double lo = (u0 < u1) ? u0 : u1, hi = (u0 < u1) ? u1 : u0;
        f = (f < lo) ? lo : (f > hi) ? hi : f;
      }
      if (run.blend != 1.0) f = P[k] + run.blend * (f - P[k]);
      Fv[k] = f;
    }
    Fv[FID_RHO_W] = reg[FID_RHO_W][ry][rx];
    // rho > 0 and p > 0, with p > 0 <=> 2 rho (rho E) > |rho u|^2 for rho > 0
    const double m2 = (Fv[FID_RHO_U] * Fv[FID_RHO_U] + Fv[FID_RHO_V] * Fv[FID_RHO_V]) + Fv[FID_RHO_W] * Fv[FID_RHO_W];
    pos = (Fv[FID_RHO] > 0.0) && (2.0 * Fv[FID_RHO] * Fv[FID_RHO_E] > m2);
  }
  const unsigned long long fire_m = __ballot(fires);
  const unsigned long long bad_m = __ballot(!pos);
  const bool cell_fire = ((fire_m >> c) | (fire_m >> (c + 16))) & 1ull;
  const bool cell_bad = ((bad_m >> c) | (bad_m >> (c + 16)) | (bad_m >> (c + 32)) | (bad_m >> (c + 48))) & 1ull;
  const bool learned = use_net && learnable && !cell_fire && !cell_bad;

  if (live && !learned) {
    if (!need_plain_all) plain_face(P);
#pragma unroll
    for (int k = 0; k < NF; ++k) Fv[k] = P[k];
  }
  return learned;
}

// Whether the cell at unwrapped position (gx, gy) has its whole 5x5 stencil
// inside the grid or on periodic axes.
__device__ __forceinline__ bool stencil_inside(const GridDev &g, int gx, int gy) {
  return (g.periodic[0] || (gx >= 2 && gx <= (int)g.n[0] - 3)) && (g.periodic[1] || (gy >= 2 && gy <= (int)g.n[1] - 3));
}

// The face-state kernel: one block = 16 x 4 cells, one wave per tile row;
// writes the four face states of every cell (FS[(face * 5 + k) * fs + cell]).
constexpr int RTX = 16, RTY = 4, RGX = RTX + 4, RGY = RTY + 4;

template <int STAGE>
__global__ __launch_bounds__(256) void k_recon_learned(GridDev g, LearnRun run, const double *U, const double *R,
                                                       const double *Xin, Ctrl *ctrl, int mode) {
  __shared__ double reg[NF][RGY][RGX];
  __shared__ double pre[RGY][RGX];
  __shared__ uint8_t rex[RGY][RGX];
  __shared__ s4 sF[NFRAG * 64];
  __shared__ float sB[2 * NH + NOUT];
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // uniform over the grid
  const double dt = (STAGE == 2) ? step_dt(ctrl) : 0.0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 15, grp = lane >> 4;
  const int nx = (int)g.n[0], ny = (int)g.n[1];
  const int64_t fs = g.fs;
  const int x0 = blockIdx.x * RTX, y0 = blockIdx.y * RTY;
  load_region<STAGE, RGX, RGY, 256>(g, x0 - 2, y0 - 2, U, R, Xin, dt, reg, rex, &pre[0][0], run.net.frags != nullptr,
                                    tid);
  load_net<256>(run, sF, sB, tid);
  __syncthreads();

  const int gx = x0 + c, gy = y0 + w;
  const bool wave_live = gy < ny;
  const bool live = wave_live && gx < nx;
  const bool learnable = live && stencil_inside(g, gx, gy);
  double Fv[NF];
  const bool learned =
      cell_face<true, RGX, RGY>(g, run, reg, rex, &pre[0][0], sF, sB, c + 2, w + 2, live, learnable, wave_live, lane, Fv);
  if (live) {
    const int64_t cell = gx + (int64_t)nx * gy;
#pragma unroll
    for (int k = 0; k < NF; ++k) run.fs[(grp * NF + k) * fs + cell] = Fv[k];
  }
  if (STAGE == 1) {
    // per-block counts: one plain update of this block's own slots (no
    // contended atomics; launches on the stream are ordered)
    __shared__ int scount[2][4];
    const unsigned long long lm = __ballot(live && learned && grp == 0);
    const unsigned long long fm = __ballot(live && !learned && grp == 0);
    if (lane == 0) {
      scount[0][w] = __popcll(lm);
      scount[1][w] = __popcll(fm);
    }
    __syncthreads();
    if (tid == 0) {
      const int64_t b = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
      run.count[2 * b] += (unsigned long long)(scount[0][0] + scount[0][1] + scount[0][2] + scount[0][3]);
      run.count[2 * b + 1] += (unsigned long long)(scount[1][0] + scount[1][1] + scount[1][2] + scount[1][3]);
    }
  }
}

// One face of a cell along axis A (lower or upper per lane) from the face
// states: border state from the BC where the neighbour is outside, Rusanov
// flux, the area-weighted flux into LDS.  Returns the face eigenvalue.
template <int A>
__device__ __forceinline__ double fs_face(const GridDev &g, const double *FS, const int64_t *q, int64_t idx,
                                          bool upper, Ctrl *ctrl, double *af, int stride) {
  const int64_t fs = g.fs;
  const int f = 2 * A + (upper ? 1 : 0);
  int64_t nb = 0;
  const bool en = neighbour(g, q, A, upper ? +1 : -1, nb);
  double own[NF], oth[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) own[k] = FS[(f * NF + k) * fs + idx];
  if (en) {
#pragma unroll
    for (int k = 0; k < NF; ++k) oth[k] = FS[((f ^ 1) * NF + k) * fs + nb];
  } else if (upper) {
    border_state(g, 2 * A + 1, A, own, oth);
  } else {
    border_state(g, 2 * A, A, own, oth);
  }
  double cl[NF], cr[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    cl[k] = upper ? own[k] : oth[k];
    cr[k] = upper ? oth[k] : own[k];
  }
  Tup TL, TR;
  make_tup(cl, TL);
  make_tup(cr, TR);
  double F[NF];
  const double lam = face_flux(TL, TR, A, F);
  // each face counted once: a cell's lower face only at a border
  if (upper || !en) count_neg(TL, TR, ctrl);
  const double Ar = g.A;
#pragma unroll
  for (int k = 0; k < NF; ++k) af[k * stride] = Ar * F[k];
  return lam;
}

// 64 cells per block of four waves.  Wave w takes axis w&1 for the 32 cells
// of half w>>1, lanes 0-31 their lower faces and lanes 32-63 their upper
// faces, so the axis is uniform across a wave.  The area-weighted fluxes go
// through LDS; then thread t sums component t>>6 (threads 0-63 also
// component 4) of cell t&63 in the reference face order and writes that
// component's RK stage update.
template <int STAGE>
__global__ __launch_bounds__(256) void k_flux_fs(GridDev g, StageIn<STAGE> in, const double *FS, const double *U,
                                                 double *out, Ctrl *ctrl, int mode) {
  __shared__ double saf[4][NF][64];  // [face][component][cell of the block]
  const int slot = (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;
  if (STAGE >= 2) in.dt = step_dt(ctrl);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int axis = wv & 1;
  const bool upper = lane >= 32;
  const int cb = 32 * (wv >> 1) + (lane & 31);
  const int64_t base = (int64_t)blockIdx.x * 64;
  double lam_max = 0.0;
  {
    const int64_t idx = base + cb;
    if (idx < g.N) {
      const int64_t q[3] = {idx % g.n[0], idx / g.n[0], 0};
      double *af = &saf[2 * axis + (upper ? 1 : 0)][0][cb];
      lam_max = (axis == 0) ? fs_face<0>(g, FS, q, idx, upper, ctrl, af, 64)
                            : fs_face<1>(g, FS, q, idx, upper, ctrl, af, 64);
    }
  }
  __syncthreads();
  const int c = threadIdx.x & 63;
  const int64_t idx = base + c;
  if (idx < g.N) {
    const int64_t fs = g.fs;
    const bool wrap_x = g.periodic[0] && (idx % g.n[0]) == g.n[0] - 1;
    const bool wrap_y = g.periodic[1] && (idx / g.n[0]) == g.n[1] - 1;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r == 1 && wv != 0) break;
      const int k = (r == 0) ? wv : 4;
      double rhs = 0.0;
      if (wrap_x) { rhs -= saf[1][k][c]; rhs += saf[0][k][c]; }
      else { rhs += saf[0][k][c]; rhs -= saf[1][k][c]; }
      if (wrap_y) { rhs -= saf[3][k][c]; rhs += saf[2][k][c]; }
      else { rhs += saf[2][k][c]; rhs -= saf[3][k][c]; }
      // the stage input of component k (main.cpp:421 for stage 2)
      const double x = (STAGE == 2) ? in.U[k * in.fs + idx] + div_vol(g, in.dt * in.R[k * in.fs + idx])
                                    : in.X[k * in.fs + idx];
      const double u = (STAGE == 1) ? 0.0 : U[k * fs + idx];
      out[k * fs + idx] = rk_update<STAGE>(g, u, x, rhs, in.dt);
    }
  }
  block_max_to<256>(ctrl, slot, lam_max, threadIdx.x);
}

// Rusanov flux of an axis-A face from the face states of its lower cell (that
// cell's upper face, sl) and its upper cell (that cell's lower face, sr); a
// side beyond a non-periodic border (el or er false) takes the border state
// of the BC from the other side (euler.cpp:261-376).  Writes A F to af and
// returns the face eigenvalue; counts non-positive states when count is set.
template <int A>
__device__ __forceinline__ double face_of_states(const GridDev &g, const double *sl, const double *sr, bool el,
                                                 bool er, bool count, Ctrl *ctrl, double *af) {
  double cl[NF], cr[NF];
  if (!el) {
#pragma unroll
    for (int k = 0; k < NF; ++k) cr[k] = sr[k];
    border_state(g, 2 * A, A, cr, cl);
  } else if (!er) {
#pragma unroll
    for (int k = 0; k < NF; ++k) cl[k] = sl[k];
    border_state(g, 2 * A + 1, A, cl, cr);
  } else {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      cl[k] = sl[k];
      cr[k] = sr[k];
    }
  }
  Tup TL, TR;
  make_tup(cl, TL);
  make_tup(cr, TR);
  double F[NF];
  const double lam = face_flux(TL, TR, A, F);
  if (count) count_neg(TL, TR, ctrl);
  const double Ar = g.A;
#pragma unroll
  for (int k = 0; k < NF; ++k) af[k] = Ar * F[k];
  return lam;
}

// The tile stage kernel of 2D order-2 grids (plain face values without a
// network, learned ones with): one block = an 8 x 8 tile, six waves.  Waves 0-3
// take the tile's cells (two rows each), wave 4 the ring columns left and
// right of the tile, wave 5 the ring rows below and above it: the 32 cells
// across the tile's border faces, whose face states the neighbouring blocks
// also compute, bit for bit.  Face states stay in LDS; waves 0-1 then
// evaluate the tile's x faces and waves 2-3 its y faces, each face once, and
// every (cell, component) of the tile sums its residual in the reference
// face order and writes its RK stage update: the same values as
// k_recon_learned followed by k_flux_fs, in one launch and without the
// face-state round trip through HBM.  The stage-3 launch of a step also
// advances the time (its last block), in place of k_advance.
constexpr int TW = 8;  // tile width (two tile rows per wave of 16 cells)
constexpr int TH = 8;  // tile height of the launched kernel

template <int H>
struct TileShape {
  static constexpr int RX = TW + 6, RY = H + 6;  // region with a three-cell halo
  static constexpr int NTW = H / 2;              // waves on the tile's cells
  static constexpr int NT = 64 * (NTW + 2);       // + one wave each for ring columns and rows
};

template <int STAGE, bool NET, int H>
__global__ __launch_bounds__(TileShape<H>::NT) void k_stage_learned(GridDev g, LearnRun run, StageIn<STAGE> in,
                                                                     const double *U, double *out, Ctrl *ctrl,
                                                                     int mode) {
  using S = TileShape<H>;
  constexpr int RX = S::RX, RY = S::RY, NTW = S::NTW, NT = S::NT;
  // without a network (NET false) the weights and the sensor's pressure take no LDS
  __shared__ double reg[NF][RY][RX];
  __shared__ double pre[NET ? RY * RX : 1];
  __shared__ uint8_t rex[RY][RX];
  __shared__ s4 sF[NET ? NFRAG * 64 : 1];
  __shared__ float sB[NET ? 2 * NH + NOUT : 1];
  __shared__ double fsl[H + 2][TW + 2][4][NF];  // face states at region positions 2 .. T + 3
  __shared__ double afx[H][TW + 1][NF];         // A F of x face i (left of tile column i) in tile row j
  __shared__ double afy[TW][H + 1][NF];         // A F of y face j (below tile row j) in tile column i
  __shared__ int scount[2][NTW];
  const int slot = (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // uniform over the grid
  if (STAGE >= 2) in.dt = step_dt(ctrl);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 15, grp = lane >> 4;
  const int nx = (int)g.n[0], ny = (int)g.n[1];
  const int64_t fs = g.fs;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * H;
  const int mx = min(TW, nx - x0), my = min(H, ny - y0);  // real columns and rows of the tile

  // this thread's (cell, component) of the update; its inputs fetched early
  constexpr int NC = TW * H;
  const int uc = tid % NC, uk = tid / NC;
  const int ui = uc & 7, uj = uc >> 3;
  const bool ulive = tid < NC * NF && ui < mx && uj < my;
  const int64_t uidx = (x0 + ui) + (int64_t)nx * (y0 + uj);
  double ux = 0.0, uu = 0.0;
  if (ulive && STAGE != 1) {
    ux = (STAGE == 2) ? in.U[uk * fs + uidx] + div_vol(g, in.dt * in.R[uk * fs + uidx]) : in.X[uk * fs + uidx];
    uu = U[uk * fs + uidx];
  }

  load_region<STAGE, RX, RY, NT>(g, x0 - 3, y0 - 3, in.U, in.R, in.X, in.dt, reg, rex, pre,
                                 NET && run.net.frags != nullptr, tid);
  if (NET) load_net<NT>(run, sF, sB, tid);
  __syncthreads();

  // this lane's cell (region coordinates): tile rows 2w, 2w+1 (waves below
  // NTW), ring columns at x0 - 1 and after the last tile column (wave NTW),
  // ring rows at y0 - 1 and after the last tile row (wave NTW + 1); lanes
  // without a cell keep coordinates inside the region
  int rx, ry;
  bool live;
  if (w < NTW) {
    rx = 3 + (c & 7);
    ry = 3 + 2 * w + (c >> 3);
    live = (c & 7) < mx && 2 * w + (c >> 3) < my;
  } else if (w == NTW) {
    const int j = c & 7;
    rx = (c < 8) ? 2 : 3 + mx;
    ry = 3 + ((j < H) ? j : H - 1);
    live = j < my && rex[ry][rx];
  } else {
    rx = 3 + (c & 7);
    ry = (c < 8) ? 2 : 3 + my;
    live = (c & 7) < mx && rex[ry][rx];
  }
  const bool learnable = live && stencil_inside(g, x0 - 3 + rx, y0 - 3 + ry);
  const bool wave_live = __ballot(live) != 0ull;
  double Fv[NF];
  const bool learned = cell_face<NET, RX, RY>(g, run, reg, rex, pre, sF, sB, rx, ry, live, learnable, wave_live, lane, Fv);
  if (live) {
#pragma unroll
    for (int k = 0; k < NF; ++k) fsl[ry - 2][rx - 2][grp][k] = Fv[k];
  }
  if (STAGE == 1 && w < NTW) {
    const unsigned long long lm = __ballot(live && learned && grp == 0);
    const unsigned long long fm = __ballot(live && !learned && grp == 0);
    if (lane == 0) {
      scount[0][w] = __popcll(lm);
      scount[1][w] = __popcll(fm);
    }
  }
  __syncthreads();

  // each face of the tile once, x faces on the first half of the tile waves,
  // y faces on the second (the axis uniform per wave): a cell's upper face,
  // and its lower face only at a border, counts the non-positive states (as
  // the per-cell form does)
  double lam_max = 0.0;
  constexpr int XW = 64 * (NTW / 2);  // threads on x faces
  if (tid < XW) {
    const int j = tid / (TW + 1), i = tid % (TW + 1);
    if (tid < H * (TW + 1) && j < my && i <= mx) {
      const bool el = rex[3 + j][2 + i], er = rex[3 + j][3 + i];
      lam_max = face_of_states<0>(g, fsl[1 + j][i][1], fsl[1 + j][1 + i][0], el, er, i >= 1 || !el, ctrl, afx[j][i]);
    }
  } else if (tid < 2 * XW) {
    const int t = tid - XW;
    const int i = t / (H + 1), j = t % (H + 1);
    if (t < TW * (H + 1) && i < mx && j <= my) {
      const bool el = rex[2 + j][3 + i], er = rex[3 + j][3 + i];
      lam_max = face_of_states<1>(g, fsl[j][1 + i][3], fsl[1 + j][1 + i][2], el, er, j >= 1 || !el, ctrl, afy[i][j]);
    }
  }
  __syncthreads();

  if (ulive) {
    // reference face order; the wrap face first for the last cell of a periodic axis
    const bool wrap_x = g.periodic[0] && x0 + ui == nx - 1;
    const bool wrap_y = g.periodic[1] && y0 + uj == ny - 1;
    const double xm = afx[uj][ui][uk], xp = afx[uj][ui + 1][uk];
    const double ym = afy[ui][uj][uk], yp = afy[ui][uj + 1][uk];
    double rhs = 0.0;
    if (wrap_x) { rhs -= xp; rhs += xm; }
    else { rhs += xm; rhs -= xp; }
    if (wrap_y) { rhs -= yp; rhs += ym; }
    else { rhs += ym; rhs -= yp; }
    out[uk * fs + uidx] = rk_update<STAGE>(g, uu, ux, rhs, in.dt);
  }
  if (STAGE == 1 && tid == 0 && run.count != nullptr) {
    const int64_t b = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    int nl = 0, nf = 0;
#pragma unroll
    for (int v = 0; v < NTW; ++v) {
      nl += scount[0][v];
      nf += scount[1][v];
    }
    run.count[2 * b] += (unsigned long long)nl;
    run.count[2 * b + 1] += (unsigned long long)nf;
  }
  block_max_to<NT>(ctrl, slot, lam_max, tid);
  // stage 3 of a step: the last block advances the time (no separate launch)
  if (STAGE == 3 && mode == MODE_STEP) close_step(ctrl, tid);
}

int64_t learn_count_slots(const GridDev &g) {
  // both tilings' block counts (16x4 face-state tiles, 8x8 stage tiles)
  const int64_t a = ((g.n[0] + RTX - 1) / RTX) * ((g.n[1] + RTY - 1) / RTY);
  const int64_t b = ((g.n[0] + TW - 1) / TW) * ((g.n[1] + TH - 1) / TH);
  return 2 * (a > b ? a : b);
}

template <int STAGE>
static hipError_t stage_learned_t(const GridDev &g, const LearnRun &run, int mode, const double *U, const double *R,
                                  const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  in.X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  // 8x8 tiles: 8x4 tiles (two blocks per CU at 128^2) measured slower, 47.0 vs
  // 42.4 us per learned 128^2 step and 53.7 vs 49.3 per plain 256^2 step
  dim3 grid((unsigned)((g.n[0] + TW - 1) / TW), (unsigned)((g.n[1] + TH - 1) / TH));
  if (run.net.frags != nullptr)
    hipLaunchKernelGGL((k_stage_learned<STAGE, true, TH>), grid, dim3(TileShape<TH>::NT), 0, s, g, run, in, U, out,
                       ctrl, mode);
  else
    hipLaunchKernelGGL((k_stage_learned<STAGE, false, TH>), grid, dim3(TileShape<TH>::NT), 0, s, g, run, in, U, out,
                       ctrl, mode);
  return hipGetLastError();
}

hipError_t launch_stage_learned(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U,
                                const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                                hipStream_t s) {
  switch (stage) {
    case 1: return stage_learned_t<1>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return stage_learned_t<2>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return stage_learned_t<3>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

template <int STAGE>
static hipError_t recon_t(const GridDev &g, const LearnRun &run, int mode, const double *U, const double *R,
                          const double *W, Ctrl *ctrl, hipStream_t s) {
  const double *X = (STAGE == 3) ? W : U;
  dim3 grid((unsigned)((g.n[0] + RTX - 1) / RTX), (unsigned)((g.n[1] + RTY - 1) / RTY));
  hipLaunchKernelGGL((k_recon_learned<STAGE>), grid, dim3(256), 0, s, g, run, U, R, X, ctrl, mode);
  return hipGetLastError();
}

hipError_t launch_recon_learned(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U,
                                const double *R, const double *W, Ctrl *ctrl, hipStream_t s) {
  switch (stage) {
    case 1: return recon_t<1>(g, run, mode, U, R, W, ctrl, s);
    case 2: return recon_t<2>(g, run, mode, U, R, W, ctrl, s);
    default: return recon_t<3>(g, run, mode, U, R, W, ctrl, s);
  }
}

template <int STAGE>
static hipError_t flux_t(const GridDev &g, const LearnRun &run, int mode, const double *U, const double *R,
                         const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  in.X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  hipLaunchKernelGGL((k_flux_fs<STAGE>), dim3((unsigned)((g.N + 63) / 64)), dim3(256), 0, s, g, in, run.fs, U, out,
                     ctrl, mode);
  return hipGetLastError();
}

hipError_t launch_flux_fs(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U, const double *R,
                          const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  switch (stage) {
    case 1: return flux_t<1>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return flux_t<2>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return flux_t<3>(g, run, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

}  // namespace mmf
