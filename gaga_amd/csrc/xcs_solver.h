// Team-level greedy recovery: ensemble operators and the four algorithms.
//
// Every routine here is executed by all threads of all workgroups of a team
// (see xcs_device.h).  Scalars live in registers and are computed redundantly
// and identically by every thread, so control flow is uniform.  Vectors live
// in global memory; "strided ownership" means index i of a length-L vector is
// handled by team thread i mod gsize, so chained elementwise passes need no
// barrier.  Barriers (t.sync) are placed where another thread's writes are read.
//
// Reference semantics followed (file:line in the reference):
//   NIHT    GAGA/gaga/algorithms.cu:1274-1441, RestrictedSD functions.cu:659-745
//   HTP     algorithms.cu:2491-2716, RestrictedCG functions.cu:922-1000
//   CSMPSP  algorithms.cu:3372-3674
//   CGIHT   algorithms.cu:3689-3942, functions.cu:1004-1102 (restricted step),
//           functions.cu:1107-1220 (unrestricted step)
//   support selection functions.cu:374-436, kernels.cu:64-81, 135-193
// Inner products: per-thread partials in FP32 (residual norms in FP64), the
// cross-thread combination in FP64, rounded to FP32 for the step logic; the
// residual-norm history is FP32 like the reference's.
#pragma once
#include "xcs_device.h"
#include "xcs_engine.h"
#include "xcs_fft.h"

namespace xcs {

// ------------------------------------------------------------ gen ----------
// out[i] = sum_{s < cnt} A[i, supp[s]] * v[supp[s]]   (gen_matvec_sparse,
// GAGA/Matrices/gen_gpu.cu:79-132, as a deterministic two-phase gather).
// out is written with strided ownership.  Reads v: caller makes it visible.
__device__ void gen_A(Team &t, const XProblem &P, const float *v, float *out, int cnt) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int nq = P.lda >> 2;
  const int row_tiles = (nq + WAVE - 1) / WAVE;
  int cs = (cnt + MAXCH - 1) / MAXCH;
  if (cs < 8) cs = 8;
  const int nch = (cnt + cs - 1) / cs;
  const int items = row_tiles * nch;
  for (int it = t.gwave(); it < items; it += t.nwaves()) {
    const int rt = it % row_tiles, ch = it / row_tiles;
    const int q = rt * WAVE + lane;
    if (q < nq) {
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      const int s1 = min(cnt, ch * cs + cs);
      for (int s = ch * cs; s < s1; s++) {
        const int j = P.supp[s];
        const float xv = v[j];
        const float4 a = reinterpret_cast<const float4 *>(P.A + (size_t)j * P.lda)[q];
        acc.x = fmaf(a.x, xv, acc.x);
        acc.y = fmaf(a.y, xv, acc.y);
        acc.z = fmaf(a.z, xv, acc.z);
        acc.w = fmaf(a.w, xv, acc.w);
      }
      reinterpret_cast<float4 *>(P.partials + (size_t)ch * P.lda)[q] = acc;
    }
  }
  t.sync();
  for (int i = t.gtid(); i < P.m; i += t.gsize()) {
    float s = 0.f;
    for (int c = 0; c < nch; c++) s += P.partials[(size_t)c * P.lda + i];
    out[i] = s;
  }
}

// g = A^T r, dense (gen_matvec on the transpose, gen_gpu.cu:33-74, 161-171),
// one wave per column, reading the single column-major copy of A.
// epi(j, g_j) is called once per column by lane 0 of the owning wave.
template <class Epi>
__device__ void gen_AT(Team &t, Lds &L, const XProblem &P, const float *r, Epi &&epi) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int nq = P.lda >> 2;
  const float4 *r4 = reinterpret_cast<const float4 *>(r);
  if (P.lda <= RLDS) {
    for (int q = threadIdx.x; q < nq; q += blockDim.x) L.s.r[q] = r4[q];
    __syncthreads();
    for (int j = t.gwave(); j < P.n; j += t.nwaves()) {
      const float4 *col = reinterpret_cast<const float4 *>(P.A + (size_t)j * P.lda);
      float acc = 0.f;
      int q = lane;
      for (; q + 3 * WAVE < nq; q += 4 * WAVE) {
        const float4 a0 = col[q], a1 = col[q + WAVE], a2 = col[q + 2 * WAVE], a3 = col[q + 3 * WAVE];
        const float4 b0 = L.s.r[q], b1 = L.s.r[q + WAVE], b2 = L.s.r[q + 2 * WAVE], b3 = L.s.r[q + 3 * WAVE];
        acc = fmaf(a0.x, b0.x, acc); acc = fmaf(a0.y, b0.y, acc);
        acc = fmaf(a0.z, b0.z, acc); acc = fmaf(a0.w, b0.w, acc);
        acc = fmaf(a1.x, b1.x, acc); acc = fmaf(a1.y, b1.y, acc);
        acc = fmaf(a1.z, b1.z, acc); acc = fmaf(a1.w, b1.w, acc);
        acc = fmaf(a2.x, b2.x, acc); acc = fmaf(a2.y, b2.y, acc);
        acc = fmaf(a2.z, b2.z, acc); acc = fmaf(a2.w, b2.w, acc);
        acc = fmaf(a3.x, b3.x, acc); acc = fmaf(a3.y, b3.y, acc);
        acc = fmaf(a3.z, b3.z, acc); acc = fmaf(a3.w, b3.w, acc);
      }
      for (; q < nq; q += WAVE) {
        const float4 a = col[q], b = L.s.r[q];
        acc = fmaf(a.x, b.x, acc); acc = fmaf(a.y, b.y, acc);
        acc = fmaf(a.z, b.z, acc); acc = fmaf(a.w, b.w, acc);
      }
      acc = wave_sum(acc);
      if (lane == 0) epi(j, acc);
    }
    __syncthreads();
  } else {
    for (int j = t.gwave(); j < P.n; j += t.nwaves()) {
      const float4 *col = reinterpret_cast<const float4 *>(P.A + (size_t)j * P.lda);
      float acc = 0.f;
      for (int q = lane; q < nq; q += WAVE) {
        const float4 a = col[q], b = r4[q];
        acc = fmaf(a.x, b.x, acc); acc = fmaf(a.y, b.y, acc);
        acc = fmaf(a.z, b.z, acc); acc = fmaf(a.w, b.w, acc);
      }
      acc = wave_sum(acc);
      if (lane == 0) epi(j, acc);
    }
  }
}

// ------------------------------------------------------------ smv ----------
// y = A v over the CSR copy (sparse_matvec, GAGA/Matrices/smv_gpu.cu:29-37,
// as a deterministic row gather instead of atomics).  Strided ownership.
__device__ void smv_A(Team &t, const XProblem &P, const float *v, float *out) {
  for (int i = t.gtid(); i < P.m; i += t.gsize()) {
    float s = 0.f;
    const int e1 = P.row_ptr[i + 1];
    for (int e = P.row_ptr[i]; e < e1; e++) s = fmaf(P.rvals[e], v[P.cols[e]], s);
    out[i] = s;
  }
}
// g = A^T r over CSC (sparse_matvecT, smv_gpu.cu:40-47): column gather.
template <class Epi>
__device__ void smv_AT(Team &t, const XProblem &P, const float *r, Epi &&epi) {
  for (int j = t.gtid(); j < P.n; j += t.gsize()) {
    float s = 0.f;
    const int e1 = P.col_ptr[j + 1];
    for (int e = P.col_ptr[j]; e < e1; e++) s = fmaf(P.vals[e], r[P.rows[e]], s);
    epi(j, s);
  }
}

// --------------------------------------------------------- dispatch --------
// out = A v_S.  v must be zero off the support for smv/dct (it always is:
// callers pass thresholded vectors).  Strided ownership of out.
__device__ void op_A(Team &t, Lds &L, const XProblem &P, const float *v, float *out, int cnt) {
  if (P.ens == ENS_GEN)
    gen_A(t, P, v, out, cnt);
  else if (P.ens == ENS_SMV)
    smv_A(t, P, v, out);
  else
    dct_A(t, L, P, v, out);
}
template <class Epi>
__device__ void op_AT(Team &t, Lds &L, const XProblem &P, const float *r, Epi &&epi) {
  if (P.ens == ENS_GEN)
    gen_AT(t, L, P, r, epi);
  else if (P.ens == ENS_SMV)
    smv_AT(t, P, r, epi);
  else
    dct_AT(t, L, P, r, epi);
}

// ------------------------------------------------------ support state ------
// LinearBinning over v (kernels.cu:135-146); counters must be zero and
// visible.  Ends with a team barrier (counters complete).
__device__ void team_binning(Team &t, Lds &L, const XProblem &P, const float *v, int *bins,
                             int max_bin, float slope, float intercept) {
  const bool lds = max_bin <= HLDS;
  if (lds) {
    for (int b = threadIdx.x; b < max_bin; b += blockDim.x) L.s.hist[b] = 0;
    __syncthreads();
  }
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float temp = fabsf(v[i]);
    int b;
    if ((double)temp > (double)intercept * .000001) {
      b = (int)fmaxf(0.0f, __fmul_rn(slope, __fsub_rn(intercept, temp)));
      if (b < max_bin) {
        if (lds)
          atomicAdd(&L.s.hist[b], 1);
        else
          atomicAdd(&P.counters[b], 1);
      }
    } else {
      b = (int)__fmaf_rn(slope, intercept, 1.0f);
    }
    bins[i] = b;
  }
  if (lds) {
    __syncthreads();
    for (int b = threadIdx.x; b < max_bin; b += blockDim.x)
      if (L.s.hist[b]) atomicAdd(&P.counters[b], L.s.hist[b]);
  }
  t.sync();
}

// Zero the histogram counters (strided); caller syncs before binning.
__device__ void zero_counters(Team &t, const XProblem &P) {
  for (int b = t.gtid(); b < P.num_bins; b += t.gsize()) P.counters[b] = 0;
}

// FindSupportSet's binning branch (functions.cu:404-428) on vector v whose
// bins go to `bins`; scalars in S.  Caller zeroed counters and synced.
__device__ void find_support_bin(Team &t, Lds &L, const XProblem &P, const float *v, int *bins,
                                 XState &S, float slope, float intercept, int *k_bin_out) {
  team_binning(t, L, P, v, bins, S.max_bin, slope, intercept);
  int kb, sum;
  kbin_select(P.counters, S.max_bin, P.k, &kb, &sum);
  if (sum < P.k) {
    S.alpha_bin = S.alpha_bin * 0.5f;
    S.max_bin = (int)((float)P.num_bins * (1.0f - S.alpha_bin));
  }
  S.min_val = __fsub_rn(intercept, __fdiv_rn((float)(kb + 1), slope));
  S.sum = sum;
  *k_bin_out = kb;
}

// Ascending support list {i : bins[i] <= k_bin} into P.supp; returns its
// length (uniform).  Requires bins visible; ends with supp visible.
__device__ int build_support(Team &t, const XProblem &P, int k_bin) {
  const int n = P.n;
  const int per = (n + t.size - 1) / t.size;
  const int lo = min(n, t.rank * per), hi = min(n, lo + per);
  int c = 0;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) c += (P.bins[i] <= k_bin);
  int wg_total;
  wg_exclusive_scan(c, &wg_total);
  int offset = 0, total = wg_total;
  if (t.size > 1) {
    if (threadIdx.x == 0) P.wg_counts[t.rank] = wg_total;
    t.sync();
    int before = 0, all = 0;
    for (int w = threadIdx.x; w < t.size; w += blockDim.x) {
      const int v = P.wg_counts[w];
      all += v;
      if (w < t.rank) before += v;
    }
    int vv[2] = {before, all};
    // workgroup sums (integer: exact, order free)
    __shared__ int sb[2][MAX_WAVES];
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    vv[0] = wave_sum_i(vv[0]);
    vv[1] = wave_sum_i(vv[1]);
    if (lane == 0) {
      sb[0][wave] = vv[0];
      sb[1][wave] = vv[1];
    }
    __syncthreads();
    offset = 0;
    total = 0;
    for (int w = 0; w < (int)blockDim.x / WAVE; w++) {
      offset += sb[0][w];
      total += sb[1][w];
    }
    __syncthreads();
  }
  for (int base = lo; base < hi; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int f = (i < hi) && (P.bins[i] <= k_bin);
    int tot;
    const int pos = wg_exclusive_scan(f, &tot);
    if (f) P.supp[offset + pos] = i;
    offset += tot;
  }
  t.sync();
  return total;
}

// ------------------------------------------------------------ helpers ------
__device__ __forceinline__ void push16(float *h, float v) {
#pragma unroll
  for (int j = 0; j < 15; j++) h[j] = h[j + 1];
  h[15] = v;
}
__device__ __forceinline__ float max16(const float *h) {
  float mx = h[0];
#pragma unroll
  for (int j = 1; j < 16; j++)
    if (h[j] > mx) mx = h[j];
  return mx;
}

// resid = y - A x_S ; returns the FP64 partial of ||resid||^2 (strided).
__device__ double residual(Team &t, Lds &L, const XProblem &P, const XState &S, float *resid,
                           float *tmp) {
  op_A(t, L, P, P.x, tmp, S.supp_count);
  double e2 = 0.0;
  for (int i = t.gtid(); i < P.m; i += t.gsize()) {
    const float r = __fsub_rn(P.y[i], tmp[i]);
    resid[i] = r;
    e2 += (double)r * (double)r;
  }
  return e2;
}

template <int K>
__device__ __forceinline__ void reduce(Team &t, double (&v)[K], const int (&op)[K]) {
  team_reduce<double, K>(t, v, op);
}

// ------------------------------------------------------------ steps --------
// RestrictedSD_S_* (functions.cu:473-745): returns maxChange; *xmax gets the
// thread partial of max|x_new|.
__device__ float step_restricted_sd(Team &t, Lds &L, const XProblem &P, XState &S, float *xmax) {
  const int kb = S.k_bin;
  double e2 = residual(t, L, P, S, P.resid, P.resid_update);
  {
    double v[1] = {e2};
    const int op[1] = {0};
    reduce(t, v, op);
    push16(S.prev, (float)sqrt(v[0]));
  }
  float gmax = 0.f, th2 = 0.f;
  op_AT(t, L, P, P.resid, [&](int j, float g) {
    P.grad[j] = g;
    const float th = (P.bins[j] <= kb) ? g : 0.f;
    P.thres[j] = th;
    th2 = fmaf(th, th, th2);
    gmax = fmaxf(gmax, fabsf(g));
  });
  t.sync();
  op_A(t, L, P, P.thres, P.resid_update, S.supp_count);
  float a2 = 0.f;
  for (int i = t.gtid(); i < P.m; i += t.gsize()) a2 = fmaf(P.resid_update[i], P.resid_update[i], a2);
  double v[3] = {a2, th2, gmax};
  const int op[3] = {0, 0, 1};
  reduce(t, v, op);
  const float err = (float)sqrt(v[0]);
  float mu = (float)sqrt(v[1]);
  if (mu < 400.0f * err) {
    mu = mu / err;
    mu = mu * mu;
  } else {
    mu = 0.f;
  }
  float xm = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float xn = fmaf(mu, P.grad[i], P.x[i]);
    P.x[i] = xn;
    xm = fmaxf(xm, fabsf(xn));
  }
  *xmax = xm;
  return 2.0f * mu * (float)v[2];
}

// RestrictedCG_S_* (functions.cu:752-1000): one CG step on the support.
// grad = search direction, gprev = CG residual, S.mu carries <r,r>.
// Pushes the squared residual onto hist.
__device__ void step_restricted_cg(Team &t, Lds &L, const XProblem &P, XState &S, float *hist) {
  const int kb = S.k_bin;
  op_A(t, L, P, P.grad, P.resid, S.supp_count);
  float a2 = 0.f;
  for (int i = t.gtid(); i < P.m; i += t.gsize()) a2 = fmaf(P.resid[i], P.resid[i], a2);
  float alpha;
  {
    double v[1] = {a2};
    const int op[1] = {0};
    reduce(t, v, op);
    alpha = (float)v[0];
  }
  if (S.mu < 1000.0f * alpha)
    alpha = S.mu / alpha;
  else
    alpha = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) P.x[i] = fmaf(alpha, P.grad[i], P.x[i]);
  float gp2 = 0.f;
  op_AT(t, L, P, P.resid, [&](int j, float g) {
    const float th = (P.bins[j] <= kb) ? g : 0.f;
    const float gp = fmaf(-alpha, th, P.gprev[j]);
    P.gprev[j] = gp;
    gp2 = fmaf(gp, gp, gp2);
  });
  float beta = S.mu;
  {
    double v[1] = {gp2};
    const int op[1] = {0};
    reduce(t, v, op);
    S.mu = (float)v[0];
  }
  if (S.mu < 1000.0f * beta)
    beta = S.mu / beta;
  else
    beta = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize())
    P.grad[i] = __fadd_rn(__fmul_rn(beta, P.grad[i]), P.gprev[i]);
  t.sync();
  double e2 = residual(t, L, P, S, P.resid, P.resid_update);
  double v[1] = {e2};
  const int op[1] = {0};
  reduce(t, v, op);
  push16(hist, (float)v[0]);
}

// CG projection on the current support shared by HTP and CSMPSP
// (algorithms.cu:2607-2669, 3431-3488, 3548-3604).  Returns errCG.
__device__ float cg_projection(Team &t, Lds &L, const XProblem &P, XState &S, bool fixed_diff) {
  const int kb = S.k_bin;
  float hist[16], evo[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    hist[j] = 0.f;
    evo[j] = 1.f;
  }
  hist[15] = S.prev[15];
  float diffCG = 1.f;
  const float tolCG = P.tol * P.tol;
  const float tol2 = 10.0f * tolCG;
  double e2 = residual(t, L, P, S, P.resid, P.resid_update);
  t.sync();
  float g2 = 0.f;
  op_AT(t, L, P, P.resid, [&](int j, float g) {
    const float th = (P.bins[j] <= kb) ? g : 0.f;
    P.grad[j] = th;
    P.gprev[j] = th;
    g2 = fmaf(th, th, g2);
  });
  double v[2] = {g2, e2};
  const int op[2] = {0, 0};
  reduce(t, v, op);
  S.mu = (float)v[0];
  float errCG = (float)v[1];
  hist[15] = errCG;
  int iterCG = 0;
  while ((errCG > tolCG) & (iterCG < 15) & (diffCG > tol2)) {
    step_restricted_cg(t, L, P, S, hist);
    errCG = hist[15];
    iterCG++;
#pragma unroll
    for (int j = 0; j < 15; j++) evo[j] = evo[j + 1];
    evo[15] = hist[14] - hist[15];
    diffCG = fixed_diff ? 10.0f : max16(evo);
  }
  return errCG;
}

// RestrictedCGwithSupportEvolution_S_* (functions.cu:1004-1102).
__device__ float step_cg_restricted(Team &t, Lds &L, const XProblem &P, XState &S, float *xmax) {
  const int kb = S.k_bin;
  double e2 = residual(t, L, P, S, P.resid, P.resid_update);
  {
    double v[1] = {e2};
    const int op[1] = {0};
    reduce(t, v, op);
    push16(S.prev, (float)sqrt(v[0]));
  }
  float th2 = 0.f;
  op_AT(t, L, P, P.resid, [&](int j, float g) {
    P.grad[j] = g;
    const float th = (P.bins[j] <= kb) ? g : 0.f;
    P.thres[j] = th;
    th2 = fmaf(th, th, th2);
  });
  float beta;
  {
    double v[1] = {th2};
    const int op[1] = {0};
    reduce(t, v, op);
    const float mu_new = (float)v[0];
    if (S.beta_to_zero == 1) {
      beta = 0.f;
      S.mu = mu_new;
    } else {
      beta = S.mu;
      S.mu = mu_new;
      if (S.mu < 1000.0f * beta)
        beta = S.mu / beta;
      else
        beta = 0.f;
    }
  }
  float pmax = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float gp = __fadd_rn(__fmul_rn(beta, P.gprev[i]), P.grad[i]);
    P.gprev[i] = gp;
    P.gprev_thres[i] = (P.bins[i] <= kb) ? gp : 0.f;
    pmax = fmaxf(pmax, fabsf(gp));
  }
  t.sync();
  op_A(t, L, P, P.gprev_thres, P.resid_update, S.supp_count);
  float a2 = 0.f;
  for (int i = t.gtid(); i < P.m; i += t.gsize()) a2 = fmaf(P.resid_update[i], P.resid_update[i], a2);
  double v[2] = {a2, pmax};
  const int op[2] = {0, 1};
  reduce(t, v, op);
  float alpha = (float)v[0];
  if (S.mu < 1000.0f * alpha)
    alpha = S.mu / alpha;
  else
    alpha = 0.f;
  float xm = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float xn = fmaf(alpha, P.gprev[i], P.x[i]);
    P.x[i] = xn;
    xm = fmaxf(xm, fabsf(xn));
  }
  *xmax = xm;
  return 2.0f * alpha * (float)v[1];
}

// UnrestrictedCG_S_* (functions.cu:1107-1220).
__device__ float step_cg_unrestricted(Team &t, Lds &L, const XProblem &P, XState &S, float *xmax) {
  const int kb = S.k_bin;
  // resid_update = -(A x - y)
  op_A(t, L, P, P.x, P.resid_update, S.supp_count);
  double e2 = 0.0;
  for (int i = t.gtid(); i < P.m; i += t.gsize()) {
    const float r = -__fsub_rn(P.resid_update[i], P.y[i]);
    P.resid_update[i] = r;
    e2 += (double)r * (double)r;
  }
  {
    double v[1] = {e2};
    const int op[1] = {0};
    reduce(t, v, op);
    push16(S.prev, (float)sqrt(v[0]));
  }
  op_AT(t, L, P, P.resid_update, [&](int j, float g) {
    P.grad[j] = g;
    P.thres[j] = (P.bins[j] <= kb) ? g : 0.f;
  });
  t.sync();
  op_A(t, L, P, P.gprev_thres, P.resid, S.supp_count);
  t.sync();
  op_A(t, L, P, P.thres, P.resid_update, S.supp_count);
  float beta = 0.f;
  if (S.beta_to_zero != 1) {
    float num = 0.f, den = 0.f;
    for (int i = t.gtid(); i < P.m; i += t.gsize()) {
      num = fmaf(P.resid[i], P.resid_update[i], num);
      den = fmaf(P.resid[i], P.resid[i], den);
    }
    double v[2] = {num, den};
    const int op[2] = {0, 0};
    reduce(t, v, op);
    beta = -((float)v[0]) / (float)v[1];
  }
  float pmax = 0.f, an = 0.f, mu2 = 0.f, ad = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float gp = __fadd_rn(__fmul_rn(beta, P.gprev[i]), P.grad[i]);
    P.gprev[i] = gp;
    const float gpt = (P.bins[i] <= kb) ? gp : 0.f;
    P.gprev_thres[i] = gpt;
    pmax = fmaxf(pmax, fabsf(gp));
    const float th = P.thres[i];
    an = fmaf(gpt, th, an);
    mu2 = fmaf(th, th, mu2);
  }
  for (int i = t.gtid(); i < P.m; i += t.gsize()) {
    const float r = __fadd_rn(__fmul_rn(beta, P.resid[i]), P.resid_update[i]);
    P.resid[i] = r;
    ad = fmaf(r, r, ad);
  }
  double v[4] = {an, ad, pmax, mu2};
  const int op[4] = {0, 0, 1, 0};
  reduce(t, v, op);
  float alpha = (float)v[0] / (float)v[1];
  float xm = 0.f;
  for (int i = t.gtid(); i < P.n; i += t.gsize()) {
    const float xn = fmaf(alpha, P.gprev[i], P.x[i]);
    P.x[i] = xn;
    xm = fmaxf(xm, fabsf(xn));
  }
  S.mu = (float)v[3];
  *xmax = xm;
  return 2.0f * alpha * (float)v[2];
}

// ------------------------------------------------------------ driver -------
__device__ __forceinline__ bool should_continue(const XState &S, const XProblem &P, int *reason) {
  if (!(S.err > P.tol)) { *reason = 0; return false; }
  if (!(S.iter < P.maxiter)) { *reason = 1; return false; }
  if (!(S.err < 100.0f * S.err_start)) { *reason = 2; return false; }
  if (!((double)S.diff > .01 * (double)P.tol)) { *reason = 3; return false; }
  if (S.fail != 0) { *reason = 4; return false; }
  if (!S.cycle) { *reason = 5; return false; }
  return true;
}

// residNorm_evolution / cycle / stagnation bookkeeping (algorithms.cu:1398-1421).
__device__ void stop_update(XState &S, int fail_after) {
#pragma unroll
  for (int j = 0; j < 15; j++) S.evo[j] = S.evo[j + 1];
  S.evo[15] = S.prev[14] - S.prev[15];
  S.diff = max16(S.evo);
  for (int j = 0; j < 15; j++)
    if ((S.prev[15] == S.prev[j]) && (S.iter > 4)) S.cycle = 0;
  if (S.iter > fail_after) {
    const float root = 1 / 15.0f;
    float rate = S.prev[15] / S.prev[0];
    rate = powf(rate, root);
    if (rate > 0.999f) S.fail = 1;
  }
}

__device__ void trace_write(const Team &t, const XProblem &P, int idx, int kb, float e) {
  if (t.rank == 0 && threadIdx.x == 0) {
    if (P.trace_kbin) P.trace_kbin[idx] = kb;
    if (P.trace_err) P.trace_err[idx] = e;
  }
}

// Threshold x by the current bins (kernels.cu:74-81); strided.
__device__ void threshold_x(Team &t, const XProblem &P, int kb) {
  for (int i = t.gtid(); i < P.n; i += t.gsize())
    if (P.bins[i] > kb) P.x[i] = 0.f;
}

// Rebin x with a fresh max (the `minVal <= maxChange` branch shared by NIHT,
// HTP and CGIHT).  xmax_partial: thread partial of max|x|.
__device__ void rebin_x(Team &t, Lds &L, const XProblem &P, XState &S, float xmax_partial,
                        bool save_prev) {
  {
    double v[1] = {xmax_partial};
    const int op[1] = {1};
    zero_counters(t, P);
    if (save_prev)
      for (int i = t.gtid(); i < P.n; i += t.gsize()) P.bins2[i] = P.bins[i];
    reduce(t, v, op);  // barrier: counters zeroed everywhere
    S.max_value = (float)v[0];
    S.slope = (float)(P.num_bins - 1) / S.max_value;
  }
  find_support_bin(t, L, P, P.x, P.bins, S, S.slope, S.max_value, &S.k_bin);
}

// Solve (or continue) the instance for at most `budget` outer iterations.
__device__ void solve_instance(Team &t, Lds &L, const XProblem &P, int budget) {
  XState S = *P.st;
  const int n = P.n, m = P.m;
  if (S.done) return;
  if (!S.initialized) {
    // common initialisation (algorithms.cu:1284-1358 and equivalents)
    S.iter = 0;
    S.k_bin = 0;
    S.k_bin_grad = 0;
    S.sum = 0;
    S.alpha_bin = 0.25f;
    S.max_bin = (int)((float)P.num_bins * (1.0f - S.alpha_bin));
    S.min_val = 0.f;
    S.max_change = (P.alg == ALG_NIHT || P.alg == ALG_CGIHT) ? 1.0f : 0.0f;
    S.mu = 0.f;
    S.diff = 1.f;
    S.fail = 0;
    S.cycle = 1;
    S.beta_to_zero = 1;
    S.supp_change = 1;
    S.stop_reason = 0;
    for (int j = 0; j < 16; j++) {
      S.prev[j] = 0.f;
      S.evo[j] = 1.f;
    }
    double e2 = 0.0;
    for (int i = t.gtid(); i < m; i += t.gsize()) {
      const float r = P.y[i];
      P.resid[i] = r;
      e2 += (double)r * (double)r;
    }
    zero_counters(t, P);
    t.sync();
    float xm = 0.f;
    op_AT(t, L, P, P.y, [&](int j, float g) {
      P.x[j] = g;
      xm = fmaxf(xm, fabsf(g));
    });
    double v[2] = {e2, xm};
    const int op[2] = {0, 1};
    reduce(t, v, op);
    S.err = (float)sqrt(v[0]);
    S.err_start = S.err;
    S.max_value = (float)v[1];
    if (S.max_value == 0.f) {  // A^T y == 0: x = 0 is returned
      S.initialized = 1;
      S.done = 1;
      S.supp_count = 0;
      if (t.rank == 0 && threadIdx.x == 0) *P.st = S;
      return;
    }
    S.slope = (float)(P.num_bins - 1) / S.max_value;
    // FindSupportSet with maxChange 1 (NIHT/CGIHT) or max_value (HTP/CSMPSP):
    // minVal = 0 is never greater, so this always bins.
    find_support_bin(t, L, P, P.x, P.bins, S, S.slope, S.max_value, &S.k_bin);
    threshold_x(t, P, S.k_bin);
    S.supp_count = build_support(t, P, S.k_bin);
    trace_write(t, P, 0, S.k_bin, S.err);
    if (P.alg == ALG_CSMPSP) {
      // the initial projection counts as iteration 1 (algorithms.cu:3486-3490)
      const float errCG = cg_projection(t, L, P, S, false);
      S.err = sqrtf(errCG);
      S.prev[15] = S.err;
      S.iter = 1;
      trace_write(t, P, 1, S.k_bin, S.err);
    } else if (P.alg == ALG_CGIHT) {
      // r^0 and its restriction (algorithms.cu:3762-3808)
      double e = residual(t, L, P, S, P.resid, P.resid_update);
      double w[1] = {e};
      const int o1[1] = {0};
      reduce(t, w, o1);
      S.err = (float)sqrt(w[0]);
      S.prev[15] = S.err;
      op_AT(t, L, P, P.resid, [&](int j, float g) {
        P.gprev[j] = g;
        P.gprev_thres[j] = (P.bins[j] <= S.k_bin) ? g : 0.f;
      });
      for (int j = 0; j < 15; j++) S.evo[j] = S.evo[j + 1];
      S.evo[15] = S.prev[14] - S.prev[15];
      S.beta_to_zero = 1;
      S.supp_change = 1;
    }
    S.initialized = 1;
    t.sync();
  }

  int reason = 0;
  while (budget > 0 && should_continue(S, P, &reason)) {
    budget--;
    if (P.alg == ALG_NIHT || P.alg == ALG_HTP) {
      float xm;
      S.max_change = step_restricted_sd(t, L, P, S, &xm);
      if (S.min_val <= S.max_change) {
        rebin_x(t, L, P, S, xm, false);
      } else {
        S.min_val = S.min_val - S.max_change;
      }
      threshold_x(t, P, S.k_bin);
      t.sync();
      S.supp_count = build_support(t, P, S.k_bin);
      if (P.alg == ALG_NIHT) {
        S.err = S.prev[15];
        stop_update(S, 749);
      } else {
        const float errCG = cg_projection(t, L, P, S, false);
        S.err = sqrtf(errCG);
        push16(S.prev, S.err);
        stop_update(S, 125);
      }
    } else if (P.alg == ALG_CSMPSP) {
      // gradient support, joined with the current one (algorithms.cu:3524-3542)
      float gm = 0.f;
      zero_counters(t, P);
      op_AT(t, L, P, P.resid, [&](int j, float g) {
        P.grad[j] = g;
        gm = fmaxf(gm, fabsf(g));
      });
      {
        double v[1] = {gm};
        const int op[1] = {1};
        reduce(t, v, op);
        S.max_value = (float)v[0];
      }
      S.min_val = 0.f;
      S.slope = (float)(P.num_bins - 1) / S.max_value;
      find_support_bin(t, L, P, P.grad, P.bins2, S, S.slope, S.max_value, &S.k_bin_grad);
      for (int i = t.gtid(); i < n; i += t.gsize())
        if (P.bins2[i] <= S.k_bin_grad && P.bins[i] > S.k_bin) P.bins[i] = S.k_bin;
      t.sync();
      S.supp_count = build_support(t, P, S.k_bin);
      cg_projection(t, L, P, S, true);
      // forced rebin of x (algorithms.cu:3607-3620)
      float xm = 0.f;
      for (int i = t.gtid(); i < n; i += t.gsize()) xm = fmaxf(xm, fabsf(P.x[i]));
      S.min_val = 0.f;
      rebin_x(t, L, P, S, xm, false);
      threshold_x(t, P, S.k_bin);
      t.sync();
      S.supp_count = build_support(t, P, S.k_bin);
      double e2 = residual(t, L, P, S, P.resid, P.resid_update);
      double v[1] = {e2};
      const int op[1] = {0};
      reduce(t, v, op);
      S.err = (float)sqrt(v[0]);
      push16(S.prev, S.err);
      stop_update(S, 125);
    } else {  // CGIHT
      float xm;
      if (P.restart || !S.supp_change)
        S.max_change = step_cg_restricted(t, L, P, S, &xm);
      else
        S.max_change = step_cg_unrestricted(t, L, P, S, &xm);
      S.beta_to_zero = 0;
      if (S.min_val <= S.max_change) {
        const int kb_prev = S.k_bin;
        rebin_x(t, L, P, S, xm, true);
        // checkSupportEvolution (kernels.cu:178-193), evaluated race-free
        float ch = 0.f;
        for (int i = t.gtid(); i < n; i += t.gsize())
          if (P.bins2[i] <= kb_prev && P.bins[i] > S.k_bin) ch = 1.f;
        double v[1] = {ch};
        const int op[1] = {1};
        reduce(t, v, op);
        S.supp_change = v[0] > 0.0 ? 1 : 0;
        if (S.supp_change) {
          if (P.restart) {
            S.beta_to_zero = 1;
          } else {
            for (int i = t.gtid(); i < n; i += t.gsize())
              P.gprev_thres[i] = (P.bins[i] <= S.k_bin) ? P.gprev[i] : 0.f;
          }
        }
        threshold_x(t, P, S.k_bin);
        t.sync();
        S.supp_count = build_support(t, P, S.k_bin);
      } else {
        S.supp_change = 0;
        S.min_val = S.min_val - S.max_change;
        threshold_x(t, P, S.k_bin);
        t.sync();
      }
      S.err = S.prev[15];
      stop_update(S, 749);
    }
    S.iter++;
    trace_write(t, P, S.iter, S.k_bin, S.prev[15]);
  }
  if (!should_continue(S, P, &reason)) {
    S.done = 1;
    S.stop_reason = reason;
  }
  if (t.rank == 0 && threadIdx.x == 0) *P.st = S;
}

}  // namespace xcs
