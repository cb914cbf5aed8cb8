// mmf_learn_api.cpp -- C-ABI of the learned face reconstruction
// (include/mmf.h, "learned face reconstruction").
//
// mmf_learn owns the network (fp32 parameters, Adam moments, packed bf16
// fragments for inference) and a device-resident training set built from
// fine runs by mmf_learn_collect.  A 2D order-2 mmf_ctx switches to it with
// mmf_set_learned_reconstruction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mmf.h"
#include "mmf_ctx.hpp"
#include "mmf_learn.hpp"

namespace L = mmf::learn;

struct mmf_learn {
  hipStream_t stream = nullptr;
  float *dP = nullptr, *dM = nullptr, *dV = nullptr, *dG = nullptr;
  short *dF = nullptr;
  float *dX = nullptr, *dT = nullptr, *dW = nullptr;
  int64_t capacity = 0, nsamples = 0;
  float *dpartial = nullptr, *dploss = nullptr;
  int64_t partial_blocks = 0;
  int64_t adam_t = 0;
  std::string err;
};

namespace {

int lfail(mmf_learn *n, int code, const std::string &m) {
  if (n) n->err = m;
  return code;
}

#define LCHK(n, expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return lfail(n, MMF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CCHK(ctx, expr)                                                             \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
      return MMF_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform(double lo, double hi) { return lo + (hi - lo) * ((double)(next() >> 11) * (1.0 / 9007199254740992.0)); }
};

int ensure_partial(mmf_learn *n, int64_t rows) {
  const int64_t nb = mmf::learn_train_blocks(rows);
  if (nb <= n->partial_blocks) return MMF_OK;
  if (n->dpartial) (void)hipFree(n->dpartial);
  if (n->dploss) (void)hipFree(n->dploss);
  n->dpartial = nullptr;
  n->dploss = nullptr;
  n->partial_blocks = 0;
  LCHK(n, hipMalloc(&n->dpartial, sizeof(float) * (size_t)nb * L::NPARAMS));
  LCHK(n, hipMalloc(&n->dploss, sizeof(float) * (size_t)nb));
  n->partial_blocks = nb;
  return MMF_OK;
}

int repack(mmf_learn *n) {
  LCHK(n, mmf::launch_learn_pack(n->dP, n->dF, n->stream));
  return MMF_OK;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)); }
};

}  // namespace

extern "C" {

int mmf_learn_create(uint64_t seed, int64_t capacity, mmf_learn **out) {
  if (!out || capacity < 0) return MMF_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;
  mmf_learn *n = new (std::nothrow) mmf_learn();
  if (!n) return MMF_ERR_OUT_OF_MEMORY;
  auto bad = [&](int code) {
    mmf_learn_destroy(n);
    return code;
  };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return bad(MMF_ERR_HIP);
  const size_t P = L::NPARAMS;
  if (hipMalloc(&n->dP, sizeof(float) * P) != hipSuccess || hipMalloc(&n->dM, sizeof(float) * P) != hipSuccess ||
      hipMalloc(&n->dV, sizeof(float) * P) != hipSuccess || hipMalloc(&n->dG, sizeof(float) * P) != hipSuccess ||
      hipMalloc(&n->dF, sizeof(short) * L::FRAG_SHORTS) != hipSuccess)
    return bad(MMF_ERR_OUT_OF_MEMORY);
  n->capacity = capacity;
  if (capacity > 0 &&
      (hipMalloc(&n->dX, sizeof(float) * (size_t)capacity * L::NIN) != hipSuccess ||
       hipMalloc(&n->dT, sizeof(float) * (size_t)capacity * L::NOUT) != hipSuccess ||
       hipMalloc(&n->dW, sizeof(float) * (size_t)capacity) != hipSuccess))
    return bad(MMF_ERR_OUT_OF_MEMORY);
  // weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)) in parameter order, zero biases
  std::vector<float> p(P, 0.0f);
  SplitMix64 rng{seed};
  const int fan[3] = {L::NIN, L::NH, L::NH};
  const int off[3] = {L::OFF_W1, L::OFF_W2, L::OFF_W3};
  const int cnt[3] = {L::NH * L::NIN, L::NH * L::NH, L::NOUT * L::NH};
  for (int l = 0; l < 3; ++l) {
    const double b = 1.0 / std::sqrt((double)fan[l]);
    for (int e = 0; e < cnt[l]; ++e) p[(size_t)(off[l] + e)] = (float)rng.uniform(-b, b);
  }
  if (mmf_learn_set_params(n, p.data()) != MMF_OK) return bad(MMF_ERR_HIP);
  *out = n;
  return MMF_OK;
}

int mmf_learn_destroy(mmf_learn *n) {
  if (!n) return MMF_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *q : {(void *)n->dP, (void *)n->dM, (void *)n->dV, (void *)n->dG, (void *)n->dF, (void *)n->dX,
                  (void *)n->dT, (void *)n->dW, (void *)n->dpartial, (void *)n->dploss})
    if (q) (void)hipFree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return MMF_OK;
}

const char *mmf_learn_last_error(const mmf_learn *n) { return n ? n->err.c_str() : "null network"; }

int mmf_learn_get_params(const mmf_learn *n, float *params) {
  if (!n || !params) return MMF_ERR_INVALID;
  mmf_learn *m = const_cast<mmf_learn *>(n);
  LCHK(m, hipStreamSynchronize(n->stream));
  LCHK(m, hipMemcpy(params, n->dP, sizeof(float) * L::NPARAMS, hipMemcpyDeviceToHost));
  return MMF_OK;
}

int mmf_learn_set_params(mmf_learn *n, const float *params) {
  if (!n || !params) return MMF_ERR_INVALID;
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(n->dP, params, sizeof(float) * L::NPARAMS, hipMemcpyHostToDevice));
  LCHK(n, hipMemset(n->dM, 0, sizeof(float) * L::NPARAMS));
  LCHK(n, hipMemset(n->dV, 0, sizeof(float) * L::NPARAMS));
  n->adam_t = 0;
  int rc = repack(n);
  if (rc != MMF_OK) return rc;
  LCHK(n, hipStreamSynchronize(n->stream));
  return MMF_OK;
}

int mmf_learn_collect(mmf_learn *n, mmf_ctx *fine, int32_t factor, const int64_t *coarse_cells, int32_t ncells,
                      double sensor_threshold) {
  if (!n || !fine || !coarse_cells || ncells < 0) return MMF_ERR_INVALID;
  const GridDev &g = fine->g;
  if (g.dim != 2 || g.flags) return lfail(n, MMF_ERR_INVALID, "collect: the fine grid must be 2D without bodies");
  if (factor < 1 || factor > 8 || g.n[0] % factor || g.n[1] % factor)
    return lfail(n, MMF_ERR_INVALID, "collect: factor in 1..8 dividing the fine grid");
  if (n->nsamples + 8 * (int64_t)ncells > n->capacity) return lfail(n, MMF_ERR_INVALID, "collect: capacity exceeded");
  const int64_t nc = (g.n[0] / factor) * (g.n[1] / factor);
  for (int32_t i = 0; i < ncells; ++i)
    if (coarse_cells[i] < 0 || coarse_cells[i] >= nc) return lfail(n, MMF_ERR_INVALID, "collect: coarse cell index");
  if (ncells == 0) return MMF_OK;
  DevBuf<int64_t> dcells;
  LCHK(n, dcells.alloc((size_t)ncells));
  LCHK(n, hipStreamSynchronize(fine->stream));
  LCHK(n, hipMemcpy(dcells.p, coarse_cells, sizeof(int64_t) * (size_t)ncells, hipMemcpyHostToDevice));
  const int64_t base = n->nsamples;
  LCHK(n, mmf::launch_learn_collect(g, fine->dU, factor, dcells.p, ncells, sensor_threshold, n->dX + base * L::NIN,
                                    n->dT + base * L::NOUT, n->dW + base, n->stream));
  LCHK(n, hipStreamSynchronize(n->stream));
  n->nsamples += 8 * (int64_t)ncells;
  return MMF_OK;
}

int mmf_learn_num_samples(const mmf_learn *n, int64_t *count) {
  if (!n || !count) return MMF_ERR_INVALID;
  *count = n->nsamples;
  return MMF_OK;
}

int mmf_learn_get_samples(const mmf_learn *n, int64_t first, int64_t count, float *x, float *t, float *w) {
  if (!n || first < 0 || count < 0 || first + count > n->nsamples) return MMF_ERR_INVALID;
  mmf_learn *m = const_cast<mmf_learn *>(n);
  LCHK(m, hipStreamSynchronize(n->stream));
  if (count == 0) return MMF_OK;
  if (x) LCHK(m, hipMemcpy(x, n->dX + first * L::NIN, sizeof(float) * count * L::NIN, hipMemcpyDeviceToHost));
  if (t) LCHK(m, hipMemcpy(t, n->dT + first * L::NOUT, sizeof(float) * count * L::NOUT, hipMemcpyDeviceToHost));
  if (w) LCHK(m, hipMemcpy(w, n->dW + first, sizeof(float) * count, hipMemcpyDeviceToHost));
  return MMF_OK;
}

int mmf_learn_set_samples(mmf_learn *n, const float *x, const float *t, const float *w, int64_t count) {
  if (!n || count < 0 || count > n->capacity || (count && (!x || !t || !w))) return MMF_ERR_INVALID;
  LCHK(n, hipStreamSynchronize(n->stream));
  if (count) {
    LCHK(n, hipMemcpy(n->dX, x, sizeof(float) * count * L::NIN, hipMemcpyHostToDevice));
    LCHK(n, hipMemcpy(n->dT, t, sizeof(float) * count * L::NOUT, hipMemcpyHostToDevice));
    LCHK(n, hipMemcpy(n->dW, w, sizeof(float) * count, hipMemcpyHostToDevice));
  }
  n->nsamples = count;
  return MMF_OK;
}

int mmf_learn_loss_grad(mmf_learn *n, const float *x, const float *t, const float *w, int64_t count, double *loss,
                        float *grad) {
  if (!n || !x || !t || !w || count < 1) return MMF_ERR_INVALID;
  int rc = ensure_partial(n, count);
  if (rc != MMF_OK) return rc;
  DevBuf<float> dx, dt, dw;
  DevBuf<double> dl;
  LCHK(n, dx.alloc((size_t)count * L::NIN));
  LCHK(n, dt.alloc((size_t)count * L::NOUT));
  LCHK(n, dw.alloc((size_t)count));
  LCHK(n, dl.alloc(1));
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(dx.p, x, sizeof(float) * count * L::NIN, hipMemcpyHostToDevice));
  LCHK(n, hipMemcpy(dt.p, t, sizeof(float) * count * L::NOUT, hipMemcpyHostToDevice));
  LCHK(n, hipMemcpy(dw.p, w, sizeof(float) * count, hipMemcpyHostToDevice));
  LCHK(n, mmf::launch_learn_train(n->dP, dx.p, dt.p, dw.p, count, 0, 0, n->dpartial, n->dploss, n->stream));
  LCHK(n, mmf::launch_learn_reduce(n->dpartial, n->dploss, mmf::learn_train_blocks(count), count, n->dG, dl.p,
                                   n->stream));
  LCHK(n, hipStreamSynchronize(n->stream));
  if (grad) LCHK(n, hipMemcpy(grad, n->dG, sizeof(float) * L::NPARAMS, hipMemcpyDeviceToHost));
  if (loss) LCHK(n, hipMemcpy(loss, dl.p, sizeof(double), hipMemcpyDeviceToHost));
  return MMF_OK;
}

int mmf_learn_adam_step(mmf_learn *n, const float *grad, double lr) {
  if (!n || !grad || !(lr >= 0.0)) return MMF_ERR_INVALID;
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(n->dG, grad, sizeof(float) * L::NPARAMS, hipMemcpyHostToDevice));
  n->adam_t += 1;
  const float bc1 = (float)(1.0 - std::pow(0.9, (double)n->adam_t));
  const float bc2 = (float)(1.0 - std::pow(0.999, (double)n->adam_t));
  LCHK(n, mmf::launch_learn_adam(n->dP, n->dG, n->dM, n->dV, (float)lr, bc1, bc2, n->stream));
  int rc = repack(n);
  if (rc != MMF_OK) return rc;
  LCHK(n, hipStreamSynchronize(n->stream));
  return MMF_OK;
}

int mmf_learn_train(mmf_learn *n, int32_t steps, int32_t batch, double lr0, double lr1, uint64_t seed,
                    double *final_loss) {
  if (!n || steps < 0 || batch < 1 || !(lr0 >= 0.0) || !(lr1 >= 0.0)) return MMF_ERR_INVALID;
  if (n->nsamples < 1) return lfail(n, MMF_ERR_STATE, "train: no training samples");
  if (steps == 0) return MMF_OK;
  int rc = ensure_partial(n, batch);
  if (rc != MMF_OK) return rc;
  DevBuf<double> dl;
  LCHK(n, dl.alloc((size_t)steps));
  const int64_t nb = mmf::learn_train_blocks(batch);
  for (int32_t s = 0; s < steps; ++s) {
    const double lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(M_PI * (double)s / (double)steps));
    n->adam_t += 1;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)n->adam_t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)n->adam_t));
    const uint64_t sseed = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(s + 1));
    LCHK(n, mmf::launch_learn_train(n->dP, n->dX, n->dT, n->dW, batch, n->nsamples, sseed, n->dpartial, n->dploss,
                                    n->stream));
    LCHK(n, mmf::launch_learn_reduce(n->dpartial, n->dploss, nb, batch, n->dG, dl.p + s, n->stream));
    LCHK(n, mmf::launch_learn_adam(n->dP, n->dG, n->dM, n->dV, (float)lr, bc1, bc2, n->stream));
  }
  rc = repack(n);
  if (rc != MMF_OK) return rc;
  std::vector<double> losses((size_t)steps);
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(losses.data(), dl.p, sizeof(double) * (size_t)steps, hipMemcpyDeviceToHost));
  if (final_loss) {
    const int32_t k = std::min<int32_t>(steps, 50);
    double a = 0.0;
    for (int32_t i = steps - k; i < steps; ++i) a += losses[(size_t)i];
    *final_loss = a / k;
  }
  return MMF_OK;
}

int mmf_learn_forward(mmf_learn *n, const float *x, int64_t count, int32_t use_mfma, float *y) {
  if (!n || count < 0 || (count && (!x || !y))) return MMF_ERR_INVALID;
  if (count == 0) return MMF_OK;
  DevBuf<float> dx, dy;
  LCHK(n, dx.alloc((size_t)count * L::NIN));
  LCHK(n, dy.alloc((size_t)count * L::NOUT));
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(dx.p, x, sizeof(float) * count * L::NIN, hipMemcpyHostToDevice));
  LCHK(n, mmf::launch_learn_forward(n->dP, n->dF, dx.p, count, use_mfma ? 1 : 0, dy.p, n->stream));
  LCHK(n, hipStreamSynchronize(n->stream));
  LCHK(n, hipMemcpy(y, dy.p, sizeof(float) * count * L::NOUT, hipMemcpyDeviceToHost));
  return MMF_OK;
}

int mmf_set_learned_reconstruction(mmf_ctx *ctx, const mmf_learn *n, double sensor_threshold, int32_t enable) {
  if (!ctx) return MMF_ERR_INVALID;
  if (!enable) {
    ctx->learn_on = false;
    return MMF_OK;
  }
  if (ctx->g.dim != 2 || ctx->g.order != 2 || ctx->g.flags || ctx->g.bslot) {
    ctx->err = "learned reconstruction: 2D order-2 grids without bodies or per-face border conditions";
    return MMF_ERR_INVALID;
  }
  CCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (!ctx->dFS) CCHK(ctx, hipMalloc(&ctx->dFS, sizeof(double) * 4 * 5 * (size_t)ctx->g.fs));
  if (!ctx->dlcount) {
    const size_t bytes = sizeof(unsigned long long) * (size_t)mmf::learn_count_slots(ctx->g);
    CCHK(ctx, hipMalloc(&ctx->dlcount, bytes));
    CCHK(ctx, hipMemset(ctx->dlcount, 0, bytes));
  }
  if (n) CCHK(ctx, hipStreamSynchronize(n->stream));
  ctx->learn.net.frags = n ? n->dF : nullptr;
  ctx->learn.net.params = n ? n->dP : nullptr;
  ctx->learn.theta = sensor_threshold;
  if (!ctx->learn_opts_set) {
    ctx->learn.clip = 0;
    ctx->learn.blend = 1.0;
  }
  ctx->learn.fs = ctx->dFS;
  ctx->learn.count = ctx->dlcount;
  ctx->learn_on = true;
  return MMF_OK;
}

int mmf_set_learned_options(mmf_ctx *ctx, int32_t clip, double blend) {
  if (!ctx) return MMF_ERR_INVALID;
  if (!(blend >= 0.0 && blend <= 1.0)) {
    ctx->err = "learned options: blend in [0, 1]";
    return MMF_ERR_INVALID;
  }
  ctx->learn.clip = clip ? 1 : 0;
  ctx->learn.blend = blend;
  ctx->learn_opts_set = true;
  return MMF_OK;
}

int mmf_get_learned_counts(mmf_ctx *ctx, int64_t *learned_cells, int64_t *fallback_cells) {
  if (!ctx) return MMF_ERR_INVALID;
  unsigned long long h[2] = {0, 0};
  if (ctx->dlcount) {
    std::vector<unsigned long long> slots((size_t)mmf::learn_count_slots(ctx->g));
    CCHK(ctx, hipStreamSynchronize(ctx->stream));
    CCHK(ctx, hipMemcpy(slots.data(), ctx->dlcount, sizeof(unsigned long long) * slots.size(),
                        hipMemcpyDeviceToHost));
    for (size_t i = 0; i < slots.size(); i += 2) {
      h[0] += slots[i];
      h[1] += slots[i + 1];
    }
  }
  if (learned_cells) *learned_cells = (int64_t)h[0];
  if (fallback_cells) *fallback_cells = (int64_t)h[1];
  return MMF_OK;
}

int mmf_reset_learned_counts(mmf_ctx *ctx) {
  if (!ctx) return MMF_ERR_INVALID;
  if (ctx->dlcount) {
    CCHK(ctx, hipStreamSynchronize(ctx->stream));
    CCHK(ctx, hipMemset(ctx->dlcount, 0, sizeof(unsigned long long) * (size_t)mmf::learn_count_slots(ctx->g)));
  }
  return MMF_OK;
}

int mmf_get_face_states(mmf_ctx *ctx, double *faces) {
  if (!ctx || !faces) return MMF_ERR_INVALID;
  if (!ctx->learn_on) {
    ctx->err = "face states: the learned reconstruction is not active";
    return MMF_ERR_STATE;
  }
  const hipError_t e = mmf::launch_recon_learned(ctx->g, ctx->learn, 1, mmf::MODE_RHS_ONLY, ctx->dU, ctx->dR,
                                                  ctx->dW, ctx->dctrl, ctx->stream);
  CCHK(ctx, e);
  const int64_t N = ctx->g.N, fs = ctx->g.fs;
  std::vector<double> soa((size_t)(20 * fs));
  CCHK(ctx, hipMemcpyAsync(soa.data(), ctx->dFS, sizeof(double) * soa.size(), hipMemcpyDeviceToHost, ctx->stream));
  CCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int64_t c = 0; c < N; ++c)
    for (int f = 0; f < 4; ++f)
      for (int k = 0; k < 5; ++k) faces[(c * 4 + f) * 5 + k] = soa[(size_t)((f * 5 + k) * fs + c)];
  return MMF_OK;
}

int mmf_restrict_state(mmf_ctx *ctx, int32_t factor, double *coarse) {
  if (!ctx || !coarse) return MMF_ERR_INVALID;
  const GridDev &g = ctx->g;
  if (g.dim != 2 || factor < 1 || g.n[0] % factor || g.n[1] % factor) {
    ctx->err = "restrict: 2D grid and a factor dividing it";
    return MMF_ERR_INVALID;
  }
  const int64_t n = (g.n[0] / factor) * (g.n[1] / factor) * 5;
  DevBuf<double> d;
  CCHK(ctx, d.alloc((size_t)n));
  CCHK(ctx, mmf::launch_learn_restrict(g, ctx->dU, factor, d.p, ctx->stream));
  CCHK(ctx, hipStreamSynchronize(ctx->stream));
  CCHK(ctx, hipMemcpy(coarse, d.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  return MMF_OK;
}

}  // extern "C"
