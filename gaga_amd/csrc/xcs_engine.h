// Host/device-shared descriptors of one recovery problem and its solver state.
#pragma once
#include <stdint.h>

namespace xcs {

enum Alg { ALG_NIHT = 0, ALG_HTP = 1, ALG_CSMPSP = 2, ALG_CGIHT = 3 };
enum Ens { ENS_GEN = 0, ENS_SMV = 1, ENS_DCT = 2 };

// Scalar state of one solve; lives in device memory between launches so a
// solve can be split over several launches (the "plain" host-stepped form).
struct XState {
  int initialized, done, stop_reason, iter;
  int k_bin, k_bin_grad, max_bin, sum, supp_count;
  int beta_to_zero, supp_change, fail, cycle, pad0;
  float alpha_bin, min_val, max_change, max_value, slope, mu, err, err_start, diff, pad1;
  float prev[16], evo[16];
};

// Everything a team needs to solve one instance.  Device pointers only.
struct XProblem {
  int alg, ens, m, n, k, lda, num_bins, maxiter, restart, p, nz, log2n;
  float tol;
  const float *A;  // gen: column-major, leading dimension lda (rows m..lda-1 are 0)
  const int *col_ptr, *rows;  // smv CSC (n+1, nz)
  const float *vals;
  const int *row_ptr, *cols;  // smv CSR (m+1, nz)
  const float *rvals;
  const int *dct_rows;        // dct: m sampled rows of the DCT-II
  const float *y;             // length lda, zero padded
  float *x;                   // n
  float *thres, *grad, *gprev, *gprev_thres;  // n
  int *bins, *bins2;                          // n
  float *resid, *resid_update, *resid2;       // lda each, zero padded
  int *counters;                              // >= num_bins
  int *supp;                                  // n: ascending support list
  int *wg_counts;                             // team size
  float *partials;                            // MAXCH * lda (gen gather partials)
  float2 *cwork;                              // dct: 3 * n complex scratch
  const float2 *twiddle;                      // dct: n roots of unity, then n quarter-wave
  XState *st;
  float *trace_err;                           // optional, maxiter + 2
  int *trace_kbin;                            // optional, maxiter + 2
};

constexpr int MAXCH = 64;  // gen gather: max column chunks (partials rows)

}  // namespace xcs
