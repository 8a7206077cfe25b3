// Job-table kernels: many t-SNE jobs of uneven size advance together.
//
// A launch covers several jobs.  Each job owns a contiguous range of blocks
// (`first[e] .. first[e+1]`, the same block count its single-job launch
// would use), finds its descriptor by a binary search of blockIdx.x over the
// prefix table, and runs the single-job stage body with its job-local block
// index.  Jobs therefore give bit-identical results alone and in a batch,
// and a batch of S jobs costs 6 + 3 * (#grid sizes) + (#attraction launches)
// launches per iteration instead of 10 * S.  The descriptor's iteration
// counter, the phase constants and the layout-buffer parity live on the
// device: batch_finalize advances them after the job's update, and a job
// whose schedule is complete goes inactive, so captured sequences of
// iterations can run jobs with different schedules.
#pragma once

#include "tsne_device.hpp"
#include "tsne_pairs.hpp"

namespace fe {
namespace {

struct JobDesc {
  float2* y[2];
  float2* gains;
  float2* upd;
  const int* row_ptr;
  const int* col;
  const float* val;
  BlockTables blk;  // segment-blocked attraction tables
  int groups;
  PairTables pairs;  // pair-once attraction tables
  int pair_ranges;
  SplatBuffers splat;
  float2* z1;
  float2* z2;
  float2* kt;
  float2* o1;
  float2* o2;
  float4* field;
  const float2* tw;
  TsneStats* st;
  LayoutPartial* lpart;
  double* zpart;
  float2* attr;
  int pair_attr;  // the update reads the attraction from the pair tables' slots
  int n, g;
  int upd_blocks;  // blocks of the job's update launch (layout partials)
  int cur, iter, total_iter, early_iter, active;
  float early_exag, exag, lr_early, lr_normal, mom0, mom1, min_gain, max_step;
};

// Entry e of a launch: descriptor jobs[slot[e]], blocks first[e] .. first[e+1].
__device__ __forceinline__ int find_entry(const int* __restrict__ first, int m) {
  const int bid = static_cast<int>(blockIdx.x);
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kSplatThreads) void batch_splat_count(JobDesc* jobs,
                                                              const int* __restrict__ slot,
                                                              const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  splat_count_body(J->y[J->cur], J->st, J->splat, J->n, J->g, blockIdx.x - first[e],
                   first[e + 1] - first[e]);
}

__global__ __launch_bounds__(kSplatThreads) void batch_splat_colscan(JobDesc* jobs,
                                                                    const int* __restrict__ slot,
                                                                    const int* __restrict__ first,
                                                                    int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  splat_colscan_body(J->splat, J->g, splat_bin_blocks(J->n), blockIdx.x - first[e]);
}

__global__ __launch_bounds__(kSplatThreads) void batch_splat_scatter(JobDesc* jobs,
                                                                const int* __restrict__ slot,
                                                                const int* __restrict__ first,
                                                                int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  splat_scatter_body(J->y[J->cur], J->st, J->splat, J->n, J->g, blockIdx.x - first[e],
                     first[e + 1] - first[e]);
}

__global__ __launch_bounds__(kThreads) void batch_splat_tiles(JobDesc* jobs,
                                                              const int* __restrict__ slot,
                                                              const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  splat_tiles_body(J->st, J->splat, J->g, blockIdx.x - first[e], first[e + 1] - first[e]);
}

template <int F>
__global__ __launch_bounds__(kFftThreads) void batch_rows_fwd(JobDesc* jobs, const int* __restrict__ slot,
                                                           const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  conv_rows_fwd_body<F>(J->splat, J->st, J->tw, J->z1, J->z2, J->kt, blockIdx.x - first[e]);
}

// blocks of a job's range follow the XCD-aware column order of conv_cols;
// every entry of this launch has a multiple of 8 blocks, so job-local block
// ids keep the hardware's XCD assignment
template <int F>
__global__ __launch_bounds__(kFftThreads) void batch_cols(JobDesc* jobs, const int* __restrict__ slot,
                                                       const int* __restrict__ first, int m) {
  extern __shared__ float2 lds[];
  constexpr int G = F / 2;
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const int c = xcd_column(blockIdx.x - first[e], G);
  if (c > G) return;
  conv_cols_body<F>(J->z1, J->z2, J->kt, J->tw, J->o1, J->o2, J->zpart, lds, c);
}

template <int F>
__global__ __launch_bounds__(kFftThreads) void batch_rows_inv(JobDesc* jobs, const int* __restrict__ slot,
                                                           const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  conv_rows_inv_body<F>(J->o1, J->o2, J->tw, J->zpart, J->st, J->field, blockIdx.x - first[e]);
}

// segment-blocked attraction: one workgroup per group of a job
template <bool STAGE>
__global__ __launch_bounds__(kStagedThreads) void batch_attract_blocked(JobDesc* jobs,
                                                                        const int* __restrict__ slot,
                                                                        const int* __restrict__ first,
                                                                        int m) {
  extern __shared__ float2 sy[];
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const float exag = J->iter < J->early_iter ? J->early_exag : J->exag;
  const int g = xcd_remap(static_cast<int>(blockIdx.x) - first[e], first[e + 1] - first[e]);
  attract_group<STAGE>(J->y[J->cur], J->blk, J->st->mean_xf, J->st->mean_yf, exag, J->attr, g,
                       sy);
}

// pair-once attraction: one workgroup per chunk range of a job; the update
// reads its partial slots (integer sums: the block order within a job is free)
__global__ __launch_bounds__(kPairThreads) void batch_attract_pairs(JobDesc* jobs,
                                                                    const int* __restrict__ slot,
                                                                    const int* __restrict__ first,
                                                                    int m) {
  extern __shared__ float2 lds[];
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const int w = xcd_remap(static_cast<int>(blockIdx.x) - first[e], first[e + 1] - first[e]);
  attract_pairs_range(J->y[J->cur], J->st, J->pairs, J->n, w, lds);
}

template <int TPR>
__global__ __launch_bounds__(kThreads) void batch_attract(JobDesc* jobs, const int* __restrict__ slot,
                                                          const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const float exag = J->iter < J->early_iter ? J->early_exag : J->exag;
  attract_body<TPR>(J->y[J->cur], J->row_ptr, J->col, J->val, J->st, exag, J->attr, J->n,
                    blockIdx.x - first[e], first[e + 1] - first[e]);
}

__global__ __launch_bounds__(kThreads) void batch_update(JobDesc* jobs, const int* __restrict__ slot,
                                                         const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const int it = J->iter, cur = J->cur;
  const bool early = it < J->early_iter;
  const StepParams sp{early ? J->early_exag : J->exag, early ? J->lr_early : J->lr_normal,
                      early ? J->mom0 : J->mom1, J->min_gain, J->max_step};
  const int blk = static_cast<int>(blockIdx.x) - first[e], nblk = first[e + 1] - first[e];
  if (J->pair_attr)
    update_body<0>(J->y[cur], J->y[cur ^ 1], J->gains, J->upd,
                   attr_pairs(J->pairs, J->st, sp.exaggeration), J->field, J->st, J->lpart, J->n,
                   J->g, sp, nullptr, blk, nblk);
  else
    update_body<0>(J->y[cur], J->y[cur ^ 1], J->gains, J->upd, AttrArray{J->attr}, J->field,
                   J->st, J->lpart, J->n, J->g, sp, nullptr, blk, nblk);
}

// one block per job: the next iteration's geometry from the update's layout
// partials, then the job's iteration counter, buffer parity and activity
__global__ __launch_bounds__(kThreads) void batch_finalize(JobDesc* jobs,
                                                           const int* __restrict__ slot,
                                                           const int* __restrict__ first, int m) {
  const int e = find_entry(first, m);
  JobDesc* J = jobs + slot[e];
  if (!J->active) return;
  fold_layout(J->lpart, J->upd_blocks, J->st, J->n, J->g, true);
  if (threadIdx.x == 0) {
    const int it = J->iter + 1;
    J->cur ^= 1;
    J->iter = it;
    if (it >= J->total_iter) J->active = 0;
  }
}

}  // namespace
}  // namespace fe
