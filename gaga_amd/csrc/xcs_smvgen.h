// Device smv generator: columns, zero-row repair, CSR copy.
#pragma once
#include <climits>
#include "xcs_device.h"

namespace xcs {

// ------------------------------------------------------------ smv generator --
// createProblem_smv.cu:31-167 with createMeasurements_smv.cu:129-427, built on
// the device from counter-based draws so that every column is independent.
constexpr int SMV_P_MAX = 64;          // nonzeros per column the generator supports
constexpr uint32_t SMV_DRAWS_MAX = 1u << 20;  // rejection draws per column before giving up

// One thread per column: p distinct rows by rejection (Philox key (seed,
// instance), counter (j, t, STREAM_SMV_ROWS, 0), row = bounded(word 0, m)),
// ascending; values +-1 from (j, t, STREAM_SMV_VALS) for ensemble 2, else 1,
// times `scale`; per-row counts into cnt (zeroed by the caller).
__global__ void smv_columns_kernel(int m, int n, int p, int ens, float scale, uint32_t seed,
                                   uint32_t inst, int *rows, float *vals, int *cnt, int *fail) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int r[SMV_P_MAX];
  int got = 0;
  for (uint32_t t = 0; got < p; t++) {
    if (t >= SMV_DRAWS_MAX) {  // p close to m: report instead of looping on
      atomicOr(fail, 1);
      return;
    }
    uint32_t o[4];
    philox4x32_10((uint32_t)j, t, STREAM_SMV_ROWS, 0u, seed, inst, o);
    const int cand = (int)bounded_u32(o[0], (uint32_t)m);
    bool dup = false;
    for (int q = 0; q < got; q++) dup |= (r[q] == cand);
    if (!dup) r[got++] = cand;
  }
  for (int a = 1; a < p; a++) {
    const int v = r[a];
    int b = a - 1;
    while (b >= 0 && r[b] > v) {
      r[b + 1] = r[b];
      b--;
    }
    r[b + 1] = v;
  }
  for (int t = 0; t < p; t++) {
    float v = 1.0f;
    if (ens == 2) {
      uint32_t o[4];
      philox4x32_10((uint32_t)j, (uint32_t)t, STREAM_SMV_VALS, 0u, seed, inst, o);
      v = (o[0] & 1u) ? 1.0f : -1.0f;
    }
    if (ens == 1 || ens == 2) v = v * scale;
    const size_t e = (size_t)j * p + t;
    rows[e] = r[t];
    vals[e] = v;
    atomicAdd(&cnt[r[t]], 1);
  }
}

// Zero-row repair (createMeasurements_smv.cu:263-383) and the CSR row offsets;
// one workgroup.  Every empty row takes one entry from a row that occurs more
// than once: the donors are the repeated occurrences in entry order (entry
// j*p + t; the first occurrence of each row stays), the i-th donor goes to the
// i-th empty row in ascending order, each donor's column is sorted again with
// its values.  Then row_ptr = exclusive scan of the counts.  *empty = number of
// empty rows before the repair.  Scratch: first, zero_list, donors (m each).
__global__ void __launch_bounds__(1024) smv_repair_kernel(int m, int n, int p, int *rows,
                                                          float *vals, int *cnt, int *row_ptr,
                                                          int *first, int *zero_list,
                                                          int *donors, int *empty) {
  const long np = (long)n * p;
  int z = 0;
  {
    int c = 0;
    for (int r = threadIdx.x; r < m; r += blockDim.x) c += cnt[r] == 0 ? 1 : 0;
    (void)wg_exclusive_scan(c, &z);
  }
  if (threadIdx.x == 0) *empty = z;
  if (z > 0) {
    // the empty rows, ascending
    int base = 0;
    for (int r0 = 0; r0 < m; r0 += blockDim.x) {
      const int r = r0 + threadIdx.x;
      const int f = (r < m && cnt[r] == 0) ? 1 : 0;
      int tot;
      const int pos = wg_exclusive_scan(f, &tot);
      if (f) zero_list[base + pos] = r;
      base += tot;
    }
    // first occurrence of every row in entry order
    for (int r = threadIdx.x; r < m; r += blockDim.x) first[r] = INT_MAX;
    __syncthreads();
    for (long e = threadIdx.x; e < np; e += blockDim.x) atomicMin(&first[rows[e]], (int)e);
    __syncthreads();
    // the first z repeated occurrences, in entry order
    base = 0;
    for (long e0 = 0; e0 < np && base < z; e0 += blockDim.x) {
      const long e = e0 + threadIdx.x;
      const int f = (e < np && first[rows[e]] != (int)e) ? 1 : 0;
      int tot;
      const int pos = wg_exclusive_scan(f, &tot);
      if (f && base + pos < z) donors[base + pos] = (int)e;
      base += tot;
    }
    __syncthreads();  // donors written by other threads after the scan's last barrier
    const int nd = min(base, z);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      const int e = donors[i], to = zero_list[i];
      atomicSub(&cnt[rows[e]], 1);
      atomicAdd(&cnt[to], 1);
      rows[e] = to;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      const int j = donors[i] / p;
      if (i > 0 && donors[i - 1] / p == j) continue;  // one thread per column
      int *r = rows + (size_t)j * p;
      float *v = vals + (size_t)j * p;
      for (int a = 1; a < p; a++) {
        const int rv = r[a];
        const float vv = v[a];
        int b = a - 1;
        while (b >= 0 && r[b] > rv) {
          r[b + 1] = r[b];
          v[b + 1] = v[b];
          b--;
        }
        r[b + 1] = rv;
        v[b + 1] = vv;
      }
    }
    __syncthreads();
  }
  int base = 0;
  for (int r0 = 0; r0 < m; r0 += blockDim.x) {
    const int r = r0 + threadIdx.x;
    const int c = r < m ? cnt[r] : 0;
    int tot;
    const int pos = wg_exclusive_scan(c, &tot);
    if (r < m) row_ptr[r] = base + pos;
    base += tot;
  }
  if (threadIdx.x == 0) row_ptr[m] = base;
}

// CSC column offsets j*p and the CSR fill: one thread per column places its
// entries with a per-row cursor (fill, zeroed by the caller); the order within
// a row is then fixed by smv_row_sort_kernel.
__global__ void smv_fill_kernel(int m, int n, int p, const int *rows, const float *vals,
                                const int *row_ptr, int *fill, int *col_ptr, int *cols,
                                float *rvals) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  col_ptr[j] = j * p;
  if (j == n) return;
  for (int t = 0; t < p; t++) {
    const size_t e = (size_t)j * p + t;
    const int r = rows[e];
    const int pos = row_ptr[r] + atomicAdd(&fill[r], 1);
    if (pos < row_ptr[r + 1]) {  // always, unless a column gave up (reported as an error)
      cols[pos] = j;
      rvals[pos] = vals[e];
    }
  }
}

// Columns ascending within each row.  The columns of a row are distinct, so an
// entry's place is the number of smaller columns in its row: one 64-thread
// workgroup per row stages the row in LDS, ranks every entry against it and
// writes it to its place.  Rows longer than SMV_SORT_LDS (m tiny against n p)
// fall back to an insertion sort by one thread.
constexpr int SMV_SORT_LDS = 4096;
__global__ void __launch_bounds__(64) smv_row_sort_kernel(int m, const int *row_ptr, int *cols,
                                                          float *rvals) {
  __shared__ int sc[SMV_SORT_LDS];
  __shared__ float sv[SMV_SORT_LDS];
  const int r = blockIdx.x;
  if (r >= m) return;
  const int lo = row_ptr[r], len = row_ptr[r + 1] - lo;
  if (len <= SMV_SORT_LDS) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      sc[i] = cols[lo + i];
      sv[i] = rvals[lo + i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      const int c = sc[i];
      int rank = 0;
      for (int q = 0; q < len; q++) rank += sc[q] < c ? 1 : 0;
      cols[lo + rank] = c;
      rvals[lo + rank] = sv[i];
    }
  } else if (threadIdx.x == 0) {
    for (int a = lo + 1; a < lo + len; a++) {
      const int c = cols[a];
      const float v = rvals[a];
      int b = a - 1;
      while (b >= lo && cols[b] > c) {
        cols[b + 1] = cols[b];
        rvals[b + 1] = rvals[b];
        b--;
      }
      cols[b + 1] = c;
      rvals[b + 1] = v;
    }
  }
}

}  // namespace xcs
