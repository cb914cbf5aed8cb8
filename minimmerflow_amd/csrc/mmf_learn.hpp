// mmf_learn.hpp -- learned face reconstruction: network layout and launchers.
//
// Network 100 -> 64 -> 64 -> 16 with GELU on the hidden layers.  Inputs are
// the normalised 5x5 stencil of rho, rho u, rho v, rho E around a cell;
// outputs the normalised face states of its four faces (x-, x+, y-, y+).
// Parameters (fp32 master copy) are laid out W1[64][100], b1, W2[64][64],
// b2, W3[16][64], b3.  Inference runs on the matrix cores in bf16 with fp32
// accumulation from 48 pre-packed 16x16 weight fragments.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmf_internal.hpp"

namespace mmf {
namespace learn {

constexpr int NIN = 100, NH = 64, NOUT = 16;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NOUT * NH;
constexpr int NPARAMS = OFF_B3 + NOUT;  // 11664
constexpr int NFRAG = 4 * 7 + 4 * 4 + 4;  // 16x16 bf16 fragments of W1, W2, W3
constexpr int FRAG_SHORTS = NFRAG * 64 * 4;
constexpr int TRAIN_ROWS = 32;  // rows per training block
// floor of the per-cell normalisation scale, relative to |rho| + |rho E|:
// variations far below the solution's scale are not magnified
constexpr double TAU_REL = 1e-3;

}  // namespace learn

// Network as the reconstruction kernel sees it.
struct LearnDev {
  const short *frags;    // packed bf16 fragments (FRAG_SHORTS)
  const float *params;   // fp32 parameters (biases read from here)
};

// per-context learned-reconstruction state
struct LearnRun {
  LearnDev net;               // frags == nullptr: plain face values everywhere
  double theta;               // shock-sensor threshold
  int clip;                   // clip learned face values between the means of the two cells of the face
  double blend;               // face value = plain + blend (learned - plain)
  double *fs;                 // face states, [4 faces][5 comps][fs]
  // per reconstruction block b: [2b] cells reconstructed by the network,
  // [2b+1] cells on the fallback (learn_count_slots(g) entries)
  unsigned long long *count;
};

// Number of counter entries the reconstruction kernel uses on grid g.
int64_t learn_count_slots(const GridDev &g);

hipError_t launch_learn_pack(const float *params, short *frags, hipStream_t s);
hipError_t launch_learn_train(const float *params, const float *X, const float *T, const float *W, int64_t rows,
                              int64_t nsamples, uint64_t seed, float *partial, float *ploss, hipStream_t s);
int64_t learn_train_blocks(int64_t rows);
hipError_t launch_learn_reduce(const float *partial, const float *ploss, int64_t nblocks, int64_t rows, float *grad,
                               double *loss, hipStream_t s);
hipError_t launch_learn_adam(float *params, const float *grad, float *m, float *v, float lr, float bc1, float bc2,
                             hipStream_t s);
hipError_t launch_learn_forward(const float *params, const short *frags, const float *X, int64_t n, int use_mfma,
                                float *Y, hipStream_t s);
hipError_t launch_learn_collect(const GridDev &fine, const double *Uf, int factor, const int64_t *cells, int ncells,
                                double theta, float *X, float *T, float *W, hipStream_t s);
hipError_t launch_learn_restrict(const GridDev &fine, const double *Uf, int factor, double *coarse, hipStream_t s);
hipError_t launch_recon_learned(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U,
                                const double *R, const double *W, Ctrl *ctrl, hipStream_t s);
// Fused learned stage (reconstruction + flux + RK update in one launch).
hipError_t launch_stage_learned(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U,
                                const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                                hipStream_t s);
hipError_t launch_flux_fs(const GridDev &g, const LearnRun &run, int stage, int mode, const double *U, const double *R,
                          const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s);

}  // namespace mmf
