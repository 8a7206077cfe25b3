// Host engine and C-ABI for the t-SNE FFT-grid optimizer on gfx950.
//
// The driver follows fastembedr_cuda_opentsne_fft_from_knn_impl
// (src/embedding_cuda_kernels.cpp:5629-6176): validate, build P on the device,
// initialise the layout, run the early-exaggeration and normal phases
// (captured into graphs), copy the layout back.  The per-iteration work is
// the five kernels of tsne_device.hpp.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"
#include "serve_iface.hpp"
#include "tsne_device.hpp"
#include "tsne_affinity.hpp"
#include "tsne_batch.hpp"

namespace fe {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& message) { g_last_error = message; }
const std::string& last_error() { return g_last_error; }
void clear_error() { g_last_error.clear(); }

int resolve_grid_size(int n, int requested) {
  if (requested > 0) return requested;
  int g = 512;
  if (n < 20000) g = 256;
  if (n >= 100000) g = 1024;
  const char* env = std::getenv("FASTEMBEDR_TSNE_FFT_GRID");
  if (env && env[0]) {
    const int r = std::atoi(env);
    if (r >= 16 && r <= 1024 && (r & (r - 1)) == 0) g = r;
  }
  return g;
}

// compute units of the current device (256 on MI355X)
static int compute_units() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    return prop.multiProcessorCount;
  return 256;
}

static bool graph_capture_enabled() {
  const char* env = std::getenv("FASTEMBEDR_HIP_GRAPH_CAPTURE");
  return !(env && env[0] == '0');
}

static bool overlap_enabled() {
  const char* env = std::getenv("FASTEMBEDR_HIP_OVERLAP");
  return !(env && env[0] == '0');
}

// fork/join events of one plan: two per iteration, enough that a captured
// chunk of iterations never reuses one
constexpr int kForkEvents = 128;

// ---------------------------------------------------------------------------
// P on the device
// ---------------------------------------------------------------------------
struct DeviceCsr {
  int* row_ptr = nullptr;
  int* col = nullptr;
  float* val = nullptr;
  long long nnz = 0;
};

static int exclusive_scan(const int* in, int n, int* out, int* sums, int* total, hipStream_t s) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  scan_blocks<<<nb, 256, 0, s>>>(in, n, out, sums);
  FE_LAUNCH_CHECK("scan_blocks");
  scan_sums<<<1, 256, 0, s>>>(sums, nb, total);
  FE_LAUNCH_CHECK("scan_sums");
  scan_add<<<(n + 255) / 256, 256, 0, s>>>(out, n, sums);
  FE_LAUNCH_CHECK("scan_add");
  return 0;
}

template <typename DistT>
static int launch_perplexity(const int* idx, const DistT* dist, int n, int k, int offset,
                             double perplexity, int* nbr, float* pdir, unsigned* flags,
                             hipStream_t s) {
  constexpr int rpb = aff_rows_per_block<DistT>();
  const size_t lds = static_cast<size_t>(rpb) * k * (sizeof(DistT) + sizeof(int));
  if (lds > 65536) {
    FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&perplexity_rows<DistT>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(lds)));
  }
  perplexity_rows<DistT><<<(n + rpb - 1) / rpb, 256, lds, s>>>(idx, dist, n, n, k, offset,
                                                                perplexity, nbr, pdir, flags);
  FE_LAUNCH_CHECK("perplexity_rows");
  return 0;
}

// Builds P into `state` (col/val allocated there after nnz is known through
// `alloc_csr`).  idx/dist are device pointers in the column-major layout.
template <typename DistT, typename AllocCsr>
static int build_affinities(const int* idx, const DistT* dist, int n, int k, int offset,
                            double perplexity, hipStream_t s, DeviceCsr* csr, AllocCsr alloc_csr) {
  const long long total = static_cast<long long>(n) * k;
  const int count_blocks = 1024;
  const int nb_scan = (n + 1 + kScanBlock - 1) / kScanBlock;
  ArenaPlan plan;
  plan.add<int>(total);                  // nbr
  plan.add<float>(total);                // pdir
  plan.add<int>(n + 1);                  // deg
  plan.add<int>(n + 1);                  // raw_ptr
  plan.add<int>(n + 1);                  // cursor
  plan.add<int>(n + 1);                  // merged_len
  plan.add<unsigned long long>(2 * total);  // raw entries
  plan.add<int>(nb_scan + 1);            // scan sums
  plan.add<int>(4);                      // scan total, flags
  plan.add<double>(count_blocks);        // mass partials
  Arena tmp;
  if (tmp.init(plan.bytes, "t-SNE affinity workspace")) return 1;
  int* nbr = tmp.take<int>(total);
  float* pdir = tmp.take<float>(total);
  int* deg = tmp.take<int>(n + 1);
  int* raw_ptr = tmp.take<int>(n + 1);
  int* cursor = tmp.take<int>(n + 1);
  int* merged_len = tmp.take<int>(n + 1);
  unsigned long long* raw = tmp.take<unsigned long long>(2 * total);
  int* sums = tmp.take<int>(nb_scan + 1);
  int* misc = tmp.take<int>(4);
  double* mass_part = tmp.take<double>(count_blocks);
  unsigned* flags = reinterpret_cast<unsigned*>(misc + 1);

  FE_CHECK(hipMemsetAsync(misc, 0, 4 * sizeof(int), s));
  if (launch_perplexity<DistT>(idx, dist, n, k, offset, perplexity, nbr, pdir, flags, s)) return 1;
  unsigned h_flags = 0;
  FE_CHECK(hipMemcpyAsync(&h_flags, flags, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  if (h_flags & kErrIndexRange) return fail("KNN indices are out of range.");
  if (h_flags & kErrDistance) return fail("KNN distances must be finite and non-negative.");

  FE_CHECK(hipMemsetAsync(deg, 0, (n + 1) * sizeof(int), s));
  const int row_blocks = (n + 3) / 4;  // one wave per row
  sym_row_counts<<<row_blocks, 256, 0, s>>>(nbr, n, k, deg, cursor);
  FE_LAUNCH_CHECK("sym_row_counts");
  sym_count<<<count_blocks, 256, 0, s>>>(nbr, pdir, total, k, deg, mass_part);
  FE_LAUNCH_CHECK("sym_count");
  if (exclusive_scan(deg, n + 1, raw_ptr, sums, misc, s)) return 1;
  // cursor[i] holds row i's forward count: its reverse entries follow them
  sym_fill<<<row_blocks, 256, 0, s>>>(nbr, pdir, n, k, raw_ptr, cursor, raw);
  FE_LAUNCH_CHECK("sym_fill");
  // rows of up to 512 entries by one wave each, one launch per length class;
  // the cursors are no longer needed and hold the queue of longer rows
  int* long_count = misc + 3;  // cleared with misc
  const int wblocks = (n + 3) / 4;
  sym_sort_merge_wave<1><<<wblocks, 256, 0, s>>>(raw, raw_ptr, n, merged_len, cursor, long_count);
  sym_sort_merge_wave<2><<<wblocks, 256, 0, s>>>(raw, raw_ptr, n, merged_len, cursor, long_count);
  sym_sort_merge_wave<4><<<wblocks, 256, 0, s>>>(raw, raw_ptr, n, merged_len, cursor, long_count);
  sym_sort_merge_wave<kWaveSortMaxR><<<wblocks, 256, 0, s>>>(raw, raw_ptr, n, merged_len, cursor,
                                                             long_count);
  FE_LAUNCH_CHECK("sym_sort_merge_wave");
  sym_sort_merge<<<std::min(n, 2048), 256, 0, s>>>(raw, raw_ptr, cursor, long_count, merged_len,
                                                   flags);
  FE_LAUNCH_CHECK("sym_sort_merge");
  FE_CHECK(hipMemsetAsync(merged_len + n, 0, sizeof(int), s));
  std::vector<double> mp(count_blocks);
  FE_CHECK(hipMemcpyAsync(mp.data(), mass_part, count_blocks * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(&h_flags, flags, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  Arena long_scratch;
  if (h_flags & kLongRow) {
    // rows longer than kSortCap: sorted in global memory, one block each
    ArenaPlan lp;
    lp.add<unsigned long long>(2 * total);
    if (long_scratch.init(lp.bytes, "t-SNE affinity long-row sort")) return 1;
    unsigned long long* scratch = long_scratch.take<unsigned long long>(2 * total);
    sym_sort_merge_global<<<64, kLongThreads, 0, s>>>(raw, scratch, raw_ptr, cursor, long_count,
                                                     merged_len);
    FE_LAUNCH_CHECK("sym_sort_merge_global");
    FE_CHECK(hipStreamSynchronize(s));
  }
  double mass = 0.0;
  for (double v : mp) mass += v;
  if (!std::isfinite(mass) || mass <= 0.0) {
    bool any = false;
    for (double v : mp) any = any || v != 0.0;
    return fail(any ? "t-SNE probability normalization failed."
                    : "KNN graph produced no non-self t-SNE edges.");
  }
  // final row pointers (scanned into `deg`, which is free again)
  if (exclusive_scan(merged_len, n + 1, deg, sums, misc, s)) return 1;
  int h_nnz = 0;
  FE_CHECK(hipMemcpyAsync(&h_nnz, deg + n, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  int* row_ptr = nullptr;
  int* col = nullptr;
  float* val = nullptr;
  if (alloc_csr(h_nnz, &row_ptr, &col, &val)) return 1;
  FE_CHECK(hipMemcpyAsync(row_ptr, deg, (n + 1) * sizeof(int), hipMemcpyDeviceToDevice, s));
  const float norm = static_cast<float>(0.5 / mass);
  sym_compact<<<(n + 3) / 4, 256, 0, s>>>(raw, raw_ptr, row_ptr, n, norm, col, val);
  FE_LAUNCH_CHECK("sym_compact");
  FE_CHECK(hipStreamSynchronize(s));
  csr->row_ptr = row_ptr;
  csr->col = col;
  csr->val = val;
  csr->nnz = h_nnz;
  return 0;
}

// ---------------------------------------------------------------------------
// The resident optimizer state
// ---------------------------------------------------------------------------
class TsnePlan {
 public:
  int n = 0, k = 0, g = 0, f = 0, tpr = 32;
  long long iters_run = 0;
  double setup_us = 0.0;
  DeviceCsr csr;
  Arena state;
  Arena csr_store;
  // segment-blocked attraction: groups of rows, their windows, the padded
  // segments and their step and wave tables
  Arena group_store;
  Arena attr_store;
  BlockTables blk{};
  long long seg_entries = 0;
  int groups = 0;
  int attr_kind = 0;  // AttractionKind
  // the caller's choices: attraction form (kAttrRows, kAttrStaged or
  // kAttrPairs; -1 built in) and workgroups of the pair form (0 built in)
  int attr_request = -1, pair_wg_request = 0;
  // pair-once attraction: the staged and rest streams, chunks, slots
  Arena pair_store;
  PairTables pairs{};
  int pair_ranges = 0, pair_nchunks = 0, pair_nslots = 0;
  long long pair_entries = 0, rest_entries = 0;
  float2* y[2] = {nullptr, nullptr};
  float2* gains = nullptr;
  float2* upd = nullptr;
  // splat: bins of points by grid tile, work items, the float grid of the
  // row FFTs
  SplatBuffers splat{};
  float2 *z1 = nullptr, *z2 = nullptr, *kt = nullptr, *o1 = nullptr, *o2 = nullptr;
  float4* field = nullptr;
  float2* tw = nullptr;
  TsneStats* st = nullptr;
  LayoutPartial* lpart = nullptr;
  double* zpart = nullptr;
  float2* diag = nullptr;  // 2n scratch for diagnostics / output
  // exact route (grid_size < 0) and exact Z / KL: per-slice force and Z
  // partials of exact_rep_kernel, KL block partials (allocated on first use)
  bool exact = false;
  Arena exact_store;
  double2* ex_part = nullptr;
  double* ex_zpart = nullptr;
  double* kl_part = nullptr;
  int ex_bx = 0, ex_splits = 0, ex_chunk = 0;
  int cur = 0;
  // blocks per launch: attraction (TPR lanes per point), per-point kernels,
  // splat binning (point ranges) and splat tiles
  int upd_blocks = 0, splat_blocks = 0, stat_blocks = 0, item_blocks = 0;
  Stream stream;
  // the attraction runs on `side`, beside the splat and the convolution on
  // `stream` (both read only the layout and its statistics); the update waits
  // for both.  FASTEMBEDR_HIP_OVERLAP=0 keeps every launch on `stream`.
  bool overlap = false;
  bool allow_overlap = true;  // false for a batch's jobs: the batch forks for all of them
  Stream side;
  EventRing forks;
  struct CachedGraph {
    StepParams sp;
    int iters;
    int parity;
    std::unique_ptr<GraphExec> exec;
  };
  std::vector<CachedGraph> graphs;

  template <typename DistT>
  int create(const int* idx, const DistT* dist, bool device_inputs, int n_, int k_, int offset,
             float perplexity, int grid_size, const float* init, unsigned seed, bool centre);
  int setup_attraction(hipStream_t s);
  int setup_pairs(hipStream_t s, bool* built);
  int ensure_exact();
  int launch_exact(bool forces);
  int kl(double* kl_out, double* sum_q_out);
  int enqueue_iteration(const StepParams& sp);
  int enqueue_diag(const StepParams& sp);
  GraphExec* chunk_graph(int len, const StepParams& sp);
  int run(int iters, const StepParams& sp);
  int prepare(int iters, const StepParams& sp, long long after);
  long long device_bytes() const {
    return static_cast<long long>(state.capacity() + csr_store.capacity() +
                                  group_store.capacity() + attr_store.capacity() +
                                  pair_store.capacity() + exact_store.capacity());
  }
  int layout(float* out_host);
  int profile(int iters, const StepParams& sp, double* ms);
};

static int choose_tpr(double avg_row) {
  if (avg_row < 12) return 4;
  if (avg_row < 24) return 8;
  if (avg_row < 48) return 16;
  if (avg_row < 160) return 32;
  return 64;
}

// FASTEMBEDR_HIP_ATTRACTION selects the attraction kernel: rows (one lane
// group per CSR row), lanes (segment-blocked, points read from the layout),
// staged (segment-blocked, window passes read the LDS windows), pairs (every
// pair once, tiles staged in LDS; the default where most pairs fall in staged
// tiles, else staged).  lanes and staged sum the same terms in the same order
// and agree bit for bit.
static int attraction_kind() {
  const char* v = std::getenv("FASTEMBEDR_HIP_ATTRACTION");
  if (!v || !*v) return kAttrPairs;
  if (std::strcmp(v, "rows") == 0) return kAttrRows;
  if (std::strcmp(v, "lanes") == 0) return kAttrLanes;
  if (std::strcmp(v, "staged") == 0) return kAttrStaged;
  return kAttrPairs;
}

// The pair-once tables (tsne_pairs.hpp).  *built stays false when P's tiles
// do not suit them (too few pairs in staged tiles, or sizes beyond the
// tables' 32-bit offsets): the caller then sets up the segment-blocked kernel.
int TsnePlan::setup_pairs(hipStream_t s, bool* built) {
  *built = false;
  const int nb = (n + kPairBlock - 1) / kPairBlock;
  if (nb > kPairMaxBlocks || csr.nnz <= 0) return 0;
  const size_t nt = static_cast<size_t>(nb) * nb;
  Arena scratch;
  ArenaPlan sp;
  sp.add<unsigned>(nt);   // pairs per tile
  sp.add<unsigned>(nt);   // padded pairs per tile
  sp.add<int>(nt);        // dense tile ids
  sp.add<unsigned>(4);    // largest row sum
  sp.add<int>(n);         // rest entries per row
  sp.add<unsigned>(nb);   // padded rest entries per row block
  sp.add<unsigned>(nt);   // tile cursors
  sp.add<unsigned>(nb);   // rest cursors
  if (scratch.init(sp.bytes, "t-SNE pair tables set-up")) return 1;
  unsigned* tile_cnt = scratch.take<unsigned>(nt);
  unsigned* tile_pad = scratch.take<unsigned>(nt);
  int* tile_id = scratch.take<int>(nt);
  unsigned* maxbits = scratch.take<unsigned>(4);
  int* rest_len = scratch.take<int>(n);
  unsigned* rest_pad = scratch.take<unsigned>(nb);
  unsigned* tile_cursor = scratch.take<unsigned>(nt);
  unsigned* rest_cursor = scratch.take<unsigned>(nb);
  FE_CHECK(hipMemsetAsync(tile_cnt, 0, nt * sizeof(unsigned), s));
  FE_CHECK(hipMemsetAsync(tile_pad, 0, nt * sizeof(unsigned), s));
  FE_CHECK(hipMemsetAsync(maxbits, 0, 4 * sizeof(unsigned), s));
  FE_CHECK(hipMemsetAsync(rest_pad, 0, nb * sizeof(unsigned), s));
  const int rb = (n + kThreads - 1) / kThreads;
  pair_count_kernel<<<rb, kThreads, 0, s>>>(csr.row_ptr, csr.col, csr.val, n, nb, tile_cnt, tile_pad,
                                            maxbits);
  FE_LAUNCH_CHECK("pair_count_kernel");
  std::vector<unsigned> hcnt(nt), hpad(nt);
  unsigned hmax = 0;
  FE_CHECK(hipMemcpyAsync(hcnt.data(), tile_cnt, nt * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(hpad.data(), tile_pad, nt * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(&hmax, maxbits, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  // staged tiles: every diagonal tile with pairs, off-diagonal ones with
  // enough pairs to pay for staging their second block; their stream regions
  // in (I, J) order
  struct Tile {
    int I, J;
    long long base, len;
  };
  std::vector<Tile> tiles;
  std::vector<int> hid(nt, -1);
  long long total = 0, dense_pairs = 0, all_pairs = 0;
  // FASTEMBEDR_HIP_PAIR_DENSE_MIN overrides the staging threshold
  long long dense_min = kPairDenseMin;
  if (const char* v = std::getenv("FASTEMBEDR_HIP_PAIR_DENSE_MIN"))
    if (std::atoll(v) > 0) dense_min = std::atoll(v);
  for (int I = 0; I < nb; ++I)
    for (int J = I; J < nb; ++J) {
      const size_t t = static_cast<size_t>(I) * nb + J;
      all_pairs += hcnt[t];
      if (hcnt[t] == 0 || (I != J && hcnt[t] < dense_min)) continue;
      hid[t] = static_cast<int>(tiles.size());
      tiles.push_back(Tile{I, J, total, static_cast<long long>(hpad[t])});
      total += hpad[t];
      dense_pairs += hcnt[t];
    }
  // the rest in both directions costs about 2.6 times a staged pair: below
  // half the pairs staged the segment-blocked kernel is as good
  if (2 * dense_pairs < all_pairs || total + kPairPad >= (1ll << 31)) return 0;
  FE_CHECK(hipMemcpyAsync(tile_id, hid.data(), nt * sizeof(int), hipMemcpyHostToDevice, s));
  pair_rest_count_kernel<<<rb, kThreads, 0, s>>>(csr.row_ptr, csr.col, n, nb, tile_id, rest_len,
                                                 rest_pad);
  FE_LAUNCH_CHECK("pair_rest_count_kernel");
  std::vector<unsigned> hrest(nb);
  FE_CHECK(hipMemcpyAsync(hrest.data(), rest_pad, nb * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  std::vector<long long> rbase(nb + 1, 0);
  for (int I = 0; I < nb; ++I) rbase[I + 1] = rbase[I] + hrest[I];
  const long long rtotal = rbase[nb];
  if (rtotal + kPairPad >= (1ll << 31)) return 0;
  // chunks: each region cut into pieces of at most kPairChunkEntries, row
  // block by row block (its tiles in J order, then its rest)
  std::vector<PairChunk> chunks;
  auto cut = [&](int I, int J, long long b0, long long len, int rest) {
    for (long long o = 0; o < len; o += kPairChunkEntries) {
      PairChunk c{};
      c.I = I;
      c.J = J;
      c.e0 = static_cast<int>(b0 + o);
      c.e1 = static_cast<int>(b0 + std::min<long long>(len, o + kPairChunkEntries));
      c.rest = rest;
      chunks.push_back(c);
    }
  };
  for (size_t t = 0; t < tiles.size();) {
    const int I = tiles[t].I;
    for (; t < tiles.size() && tiles[t].I == I; ++t) cut(I, tiles[t].J, tiles[t].base, tiles[t].len, 0);
  }
  // rest regions go right after their row block's tiles
  {
    std::vector<PairChunk> ordered;
    size_t k = 0;
    for (int I = 0; I < nb; ++I) {
      for (; k < chunks.size() && chunks[k].I == I; ++k) ordered.push_back(chunks[k]);
      const size_t before = chunks.size();
      if (hrest[I]) cut(I, I, rbase[I], hrest[I], 1);
      for (size_t r = before; r < chunks.size(); ++r) ordered.push_back(chunks[r]);
      chunks.resize(before);
    }
    chunks.swap(ordered);
  }
  // workgroups: one per CU when the attraction has the GPU to itself (1M
  // points: 145 us alone against 151 with two per CU); beside the splat and
  // the FFT passes (the overlapped iteration) five for every eight CUs, so
  // those kernels keep CUs of their own (3830 it/s against 3785 with 3/4 of
  // the CUs and 3370 with one per CU); never more than there are chunks
  const int cus = compute_units();
  const int want = pair_wg_request > 0 ? pair_wg_request : (overlap ? 5 * cus / 8 : cus);
  const int ranges = std::max(1, std::min(static_cast<int>(chunks.size()), want));
  // Workgroups take contiguous ranges of the chunk list of about equal cost:
  // a rest entry (gathers from the layout in memory) counts 1.3 staged
  // entries, and staging a block of the layout (a new row block, a new
  // off-diagonal tile) 8192 entries.  A chunk that a range boundary falls
  // inside is cut there (at a multiple of kLaneEntries), so the ranges meet
  // their shares whatever the chunk size (1M points: the attraction alone
  // 188 us with ranges of equal entries in whole chunks, 170 us so).
  constexpr long long kStageCost = 8192, kRestPct = 130;
  std::vector<int> range_first(ranges + 1, 0);
  {
    const int nc0 = static_cast<int>(chunks.size());
    auto stage_cost = [&](const PairChunk& ch, const PairChunk* prev) {
      long long w = 0;
      if (!prev || prev->I != ch.I) w += kStageCost;
      if (!ch.rest && ch.J != ch.I &&
          (!prev || prev->rest || prev->I != ch.I || prev->J != ch.J))
        w += kStageCost;
      return w;
    };
    long long all_cost = 0;
    for (int c = 0; c < nc0; ++c) {
      const PairChunk& ch = chunks[c];
      all_cost += stage_cost(ch, c ? &chunks[c - 1] : nullptr) +
                  static_cast<long long>(ch.e1 - ch.e0) * (ch.rest ? kRestPct : 100) / 100;
    }
    // a cut adds a staging to the range after it
    all_cost += static_cast<long long>(ranges - 1) * kStageCost;
    std::vector<PairChunk> cutc;
    cutc.reserve(nc0 + ranges);
    long long acc = 0;  // cost of the chunks (and pieces) before the current one
    int r = 1;
    for (int c = 0; c < nc0; ++c) {
      PairChunk ch = chunks[c];
      const long long pct = ch.rest ? kRestPct : 100;
      acc += stage_cost(ch, c ? &chunks[c - 1] : nullptr);
      // boundaries inside this chunk: the entry where the running cost
      // reaches r * all_cost / ranges
      while (r < ranges) {
        const long long target = (all_cost * r + ranges - 1) / ranges;
        const long long len = ch.e1 - ch.e0;
        if (acc + len * pct / 100 < target) break;
        const long long x = std::max(0ll, (target - acc) * 100 / pct) / kLaneEntries * kLaneEntries;
        if (x >= len) break;  // at the chunk's end: the next chunk starts range r
        if (x > 0) {
          PairChunk head = ch;
          head.e1 = static_cast<int>(ch.e0 + x);
          cutc.push_back(head);
          ch.e0 = head.e1;
          acc += x * pct / 100 + kStageCost;  // the next range stages again
        }
        range_first[r++] = static_cast<int>(cutc.size());
      }
      acc += static_cast<long long>(ch.e1 - ch.e0) * pct / 100;
      cutc.push_back(ch);
    }
    for (; r < ranges; ++r) range_first[r] = static_cast<int>(cutc.size());
    range_first[ranges] = static_cast<int>(cutc.size());
    chunks.swap(cutc);
  }
  const int nc = static_cast<int>(chunks.size());
  // slots: one per run of chunks of a range sharing a row block (row side)
  // and per run sharing a staged off-diagonal tile (column side)
  std::vector<std::vector<int>> block_slots(nb);
  int nslots = 0;
  for (int r = 0; r < ranges; ++r)
    for (int c = range_first[r]; c < range_first[r + 1]; ++c) {
      PairChunk& ch = chunks[c];
      const bool first = c == range_first[r];
      const PairChunk* prev = first ? nullptr : &chunks[c - 1];
      if (first || prev->I != ch.I) {
        ch.slot_i = nslots++;
        block_slots[ch.I].push_back(ch.slot_i);
      } else {
        ch.slot_i = prev->slot_i;
      }
      const bool cross = !ch.rest && ch.J != ch.I;
      ch.slot_j = -1;
      if (cross) {
        if (!first && !prev->rest && prev->I == ch.I && prev->J == ch.J) {
          ch.slot_j = prev->slot_j;
        } else {
          ch.slot_j = nslots++;
          block_slots[ch.J].push_back(ch.slot_j);
        }
      }
      const bool last = c + 1 == range_first[r + 1];
      const PairChunk* next = last ? nullptr : &chunks[c + 1];
      ch.flush = 0;
      if (last || next->I != ch.I) ch.flush |= 1;
      if (cross && (last || next->rest || next->I != ch.I || next->J != ch.J)) ch.flush |= 2;
    }
  std::vector<int> hslot_ptr(nb + 1, 0), hslot_list;
  for (int b = 0; b < nb; ++b) {
    hslot_ptr[b + 1] = hslot_ptr[b] + static_cast<int>(block_slots[b].size());
    hslot_list.insert(hslot_list.end(), block_slots[b].begin(), block_slots[b].end());
  }
  const size_t ent = static_cast<size_t>(total) + kPairPad;
  const size_t rent = static_cast<size_t>(rtotal) + kPairPad;
  ArenaPlan pp;
  pp.add<unsigned short>(ent);
  pp.add<float>(ent);
  pp.add<unsigned short>(ent / kLaneEntries);
  pp.add<int>(rent);
  pp.add<float>(rent);
  pp.add<unsigned short>(rent / kLaneEntries);
  pp.add<PairChunk>(chunks.size());
  pp.add<int>(ranges + 1);
  pp.add<int>(nb + 1);
  pp.add<int>(std::max<size_t>(1, hslot_list.size()));
  pp.add<unsigned long long>(static_cast<size_t>(std::max(nslots, 1)) * kPairBlock);
  if (pair_store.init(pp.bytes, "t-SNE pair attraction tables")) return 1;
  unsigned short* col16 = pair_store.take<unsigned short>(ent);
  float* pval = pair_store.take<float>(ent);
  unsigned short* row16 = pair_store.take<unsigned short>(ent / kLaneEntries);
  int* rcol = pair_store.take<int>(rent);
  float* rval = pair_store.take<float>(rent);
  unsigned short* rrow16 = pair_store.take<unsigned short>(rent / kLaneEntries);
  PairChunk* dchunk = pair_store.take<PairChunk>(chunks.size());
  int* drange = pair_store.take<int>(ranges + 1);
  int* dslot_ptr = pair_store.take<int>(nb + 1);
  int* dslot_list = pair_store.take<int>(std::max<size_t>(1, hslot_list.size()));
  unsigned long long* part =
      pair_store.take<unsigned long long>(static_cast<size_t>(std::max(nslots, 1)) * kPairBlock);
  // whole streams zeroed: padding past every run and past the ends reads as
  // value 0 at in-range columns
  FE_CHECK(hipMemsetAsync(col16, 0, ent * sizeof(unsigned short), s));
  FE_CHECK(hipMemsetAsync(pval, 0, ent * sizeof(float), s));
  FE_CHECK(hipMemsetAsync(row16, 0, ent / kLaneEntries * sizeof(unsigned short), s));
  FE_CHECK(hipMemsetAsync(rcol, 0, rent * sizeof(int), s));
  FE_CHECK(hipMemsetAsync(rval, 0, rent * sizeof(float), s));
  FE_CHECK(hipMemsetAsync(rrow16, 0, rent / kLaneEntries * sizeof(unsigned short), s));
  std::vector<unsigned> hcur(std::max<size_t>(1, tiles.size())), hrcur(nb);
  for (size_t t = 0; t < tiles.size(); ++t) hcur[t] = static_cast<unsigned>(tiles[t].base);
  for (int I = 0; I < nb; ++I) hrcur[I] = static_cast<unsigned>(rbase[I]);
  if (!tiles.empty())
    FE_CHECK(hipMemcpyAsync(tile_cursor, hcur.data(), tiles.size() * sizeof(unsigned),
                            hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(rest_cursor, hrcur.data(), nb * sizeof(unsigned), hipMemcpyHostToDevice, s));
  pair_fill_kernel<<<rb, kThreads, 0, s>>>(csr.row_ptr, csr.col, csr.val, n, nb, tile_id, rest_len,
                                           tile_cursor, rest_cursor, col16, pval, row16, rcol, rval,
                                           rrow16);
  FE_LAUNCH_CHECK("pair_fill_kernel");
  FE_CHECK(hipMemcpyAsync(dchunk, chunks.data(), chunks.size() * sizeof(PairChunk),
                          hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(drange, range_first.data(), (ranges + 1) * sizeof(int),
                          hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(dslot_ptr, hslot_ptr.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if (!hslot_list.empty())
    FE_CHECK(hipMemcpyAsync(dslot_list, hslot_list.data(), hslot_list.size() * sizeof(int),
                            hipMemcpyHostToDevice, s));
  FE_CHECK(hipStreamSynchronize(s));
  // the largest row sum of P (float sums; 1% headroom) bounds the sums
  float maxrow = 0.0f;
  std::memcpy(&maxrow, &hmax, sizeof(float));
  if (!std::isfinite(maxrow) || maxrow <= 0.0f) return 0;
  pairs = PairTables{col16, pval, row16, rcol, rval, rrow16, dchunk, drange, part, dslot_ptr,
                     dslot_list, 1.01 * static_cast<double>(maxrow)};
  pair_ranges = ranges;
  pair_nchunks = nc;
  pair_nslots = nslots;
  pair_entries = total;
  rest_entries = rtotal;
  static bool attr_set = false;
  if (!attr_set) {
    FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&attract_pairs),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kPairLds)));
    attr_set = true;
  }
  *built = true;
  return 0;
}

int TsnePlan::setup_attraction(hipStream_t s) {
  const long long nnz = csr.nnz;
  attr_kind = kAttrRows;
  int kind = attr_request >= 0 ? attr_request : attraction_kind();
  if (kind == kAttrPairs) {
    bool built = false;
    if (setup_pairs(s, &built)) return 1;
    if (built) {
      attr_kind = kAttrPairs;
      return 0;
    }
    kind = kAttrStaged;
  }
  // the padded segments must keep 32-bit offsets (at most kPasses segments
  // per row, each padded by fewer than 8 entries)
  const long long bound = nnz + static_cast<long long>(kPasses) * kLaneEntries * n + kLanePad;
  if (nnz <= 0 || kind == kAttrRows || bound >= (1ll << 31) ||
      static_cast<long long>(kPasses) * n + 1 >= (1ll << 31))
    return 0;
  const long long slots = static_cast<long long>(kPasses) * n;  // + 1 for the scans
  // (small P: at least 32 groups, so a job alone still spreads over the
  // GPU); the cut changes no sum
  Arena scratch;
  ArenaPlan sp;
  const int nb_rows = (n + 1 + kScanBlock - 1) / kScanBlock;
  const int nb_slots = static_cast<int>((slots + 1 + kScanBlock - 1) / kScanBlock);
  sp.add<int>(n + 1);
  sp.add<int>(n + 1);
  sp.add<int>(std::max(nb_rows, nb_slots));
  sp.add<int>(1);
  sp.add<int>(slots + 1);
  sp.add<int>(slots + 1);
  if (scratch.init(sp.bytes, "t-SNE attraction set-up")) return 1;
  int* len = scratch.take<int>(n + 1);
  int* lrow = scratch.take<int>(n + 1);
  int* sums = scratch.take<int>(std::max(nb_rows, nb_slots));
  int* total = scratch.take<int>(1);
  int* seg_len = scratch.take<int>(slots + 1);
  int* off = scratch.take<int>(slots + 1);
  lane_lengths<<<(n + kThreads) / kThreads, kThreads, 0, s>>>(csr.row_ptr, n, len);
  FE_LAUNCH_CHECK("lane_lengths");
  if (exclusive_scan(len, n + 1, lrow, sums, total, s)) return 1;
  int h_lnnz = 0;
  FE_CHECK(hipMemcpyAsync(&h_lnnz, lrow + n, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  // groups of about 16 * 8192 padded entries; when that gives more groups
  // than compute units, groups of about twice that (fewer window copies) in
  // a whole number of rounds of the compute units (one workgroup fits a CU)
  long long gcount = (h_lnnz + 16ll * 8192 - 1) / (16ll * 8192);
  const int cus = compute_units();
  if (gcount > cus) {
    gcount = (h_lnnz + 16ll * 16384 - 1) / (16ll * 16384);
    gcount = std::max<long long>(cus, (gcount + cus / 2) / cus * cus);
  }
  gcount = std::max<long long>(gcount, std::min<long long>(32, std::max(1, n / 64)));
  groups = static_cast<int>(std::min<long long>(gcount, 1 << 16));
  // group rows, windows and wave ranges; then the segments, whose size the
  // windows decide
  ArenaPlan wp;
  wp.add<int>(groups + 1);
  wp.add<int2>(static_cast<size_t>(groups) * kMaxWindows);
  wp.add<int2>(static_cast<size_t>(groups) * kPasses);
  if (group_store.init(wp.bytes, "t-SNE attraction groups")) return 1;
  int* group_row = group_store.take<int>(groups + 1);
  int2* window = group_store.take<int2>(static_cast<size_t>(groups) * kMaxWindows);
  int2* pass_range = group_store.take<int2>(static_cast<size_t>(groups) * kPasses);
  cut_rows<<<(groups + kThreads) / kThreads, kThreads, 0, s>>>(lrow, n, h_lnnz, groups, group_row);
  FE_LAUNCH_CHECK("cut_rows");
  blk_windows<<<groups, 256, 0, s>>>(csr.row_ptr, csr.col, group_row, n, window);
  FE_LAUNCH_CHECK("blk_windows");
  blk_seg_lengths<<<(n + kThreads) / kThreads, kThreads, 0, s>>>(
      csr.row_ptr, csr.col, group_row, groups, window, n, seg_len);
  FE_LAUNCH_CHECK("blk_seg_lengths");
  if (exclusive_scan(seg_len, static_cast<int>(slots + 1), off, sums, total, s)) return 1;
  int h_ent = 0;
  FE_CHECK(hipMemcpyAsync(&h_ent, off + slots, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  seg_entries = h_ent;
  const size_t ent = static_cast<size_t>(h_ent) + kLanePad;
  const int steps = h_ent / kLaneStep + 1 + kLanePadSteps;
  ArenaPlan ap;
  ap.add<int>(ent);
  ap.add<unsigned short>(ent);
  ap.add<float>(ent);
  ap.add<int>(ent / kLaneEntries);
  ap.add<unsigned long long>(steps);
  if (attr_store.init(ap.bytes, "t-SNE attraction tables")) return 1;
  int* pcol = attr_store.take<int>(ent);
  unsigned short* pcol16 = attr_store.take<unsigned short>(ent);
  float* pval = attr_store.take<float>(ent);
  int* lane_row = attr_store.take<int>(ent / kLaneEntries);
  unsigned long long* starts = attr_store.take<unsigned long long>(steps);
  FE_CHECK(hipMemsetAsync(pcol + h_ent, 0, kLanePad * sizeof(int), s));
  FE_CHECK(hipMemsetAsync(pcol16 + h_ent, 0, kLanePad * sizeof(unsigned short), s));
  FE_CHECK(hipMemsetAsync(pval + h_ent, 0, kLanePad * sizeof(float), s));
  FE_CHECK(hipMemsetAsync(lane_row + h_ent / kLaneEntries, 0,
                          kLanePad / kLaneEntries * sizeof(int), s));
  FE_CHECK(hipMemsetAsync(starts, 0, steps * sizeof(unsigned long long), s));
  blk_fill<<<(n + kThreads / 64 - 1) / (kThreads / 64), kThreads, 0, s>>>(
      csr.row_ptr, csr.col, csr.val, group_row, groups, window, n, off, lane_row,
      pcol, pcol16, pval, starts);
  FE_LAUNCH_CHECK("blk_fill");
  blk_pass_ranges<<<(groups * kPasses + kThreads - 1) / kThreads, kThreads, 0, s>>>(
      group_row, groups, off, pass_range);
  FE_LAUNCH_CHECK("blk_pass_ranges");
  FE_CHECK(hipStreamSynchronize(s));
  blk = BlockTables{pcol, pcol16, pval, lane_row, starts, group_row, window, pass_range};
  attr_kind = kind;
  static bool attr_set = false;
  if (!attr_set) {
    FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&attract_blocked<true>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kStagedLds)));
    attr_set = true;
  }
  return 0;
}

template <typename DistT>
int TsnePlan::create(const int* idx, const DistT* dist, bool device_inputs, int n_, int k_,
                     int offset, float perplexity, int grid_size, const float* init, unsigned seed,
                     bool centre) {
  n = n_;
  k = k_;
  // grid_size < 0: exact repulsion; the grid buffers keep their smallest size
  if (grid_size < -1) return fail("grid_size must be -1 (exact), 0 (the grid rule) or a power of two");
  exact = grid_size < 0;
  g = exact ? 16 : resolve_grid_size(n, grid_size);
  f = 2 * g;
  if (g < 16 || g > 1024 || (g & (g - 1)) != 0)
    return fail("grid_size must be a power of two between 16 and 1024");
  if (stream.init()) return 1;
  overlap = !exact && allow_overlap && overlap_enabled();
  if (overlap && (side.init() || forks.init(kForkEvents))) return 1;
  hipStream_t s = stream.get();
  const auto t0 = std::chrono::steady_clock::now();

  // inputs on the device
  Arena upload;
  const int* d_idx = idx;
  const DistT* d_dist = dist;
  const size_t total = static_cast<size_t>(n) * k;
  if (!device_inputs) {
    ArenaPlan up;
    up.add<int>(total);
    up.add<DistT>(total);
    if (upload.init(up.bytes, "t-SNE KNN upload")) return 1;
    int* ui = upload.take<int>(total);
    DistT* ud = upload.take<DistT>(total);
    FE_CHECK(hipMemcpyAsync(ui, idx, total * sizeof(int), hipMemcpyHostToDevice, s));
    FE_CHECK(hipMemcpyAsync(ud, dist, total * sizeof(DistT), hipMemcpyHostToDevice, s));
    d_idx = ui;
    d_dist = ud;
  }
  auto alloc_csr = [&](int nnz, int** row_ptr, int** col, float** val) -> int {
    const size_t m = static_cast<size_t>(nnz > 0 ? nnz : 1);
    ArenaPlan cp;
    cp.add<int>(n + 1);
    cp.add<int>(m);
    cp.add<float>(m);
    if (csr_store.init(cp.bytes, "t-SNE affinity CSR")) return 1;
    *row_ptr = csr_store.take<int>(n + 1);
    *col = csr_store.take<int>(m);
    *val = csr_store.take<float>(m);
    return 0;
  };
  if (build_affinities<DistT>(d_idx, d_dist, n, k, offset, perplexity, s, &csr, alloc_csr)) return 1;
  upload.release();
  tpr = choose_tpr(static_cast<double>(csr.nnz) / n);
  if (setup_attraction(s)) return 1;

  // optimizer state
  const size_t gg = static_cast<size_t>(g) * g;
  const size_t gf = static_cast<size_t>(g) * f;
  upd_blocks = static_cast<int>(std::min<long long>(
      (static_cast<long long>(n) * tpr + kThreads - 1) / kThreads, 2048));
  splat_blocks = splat_bin_blocks(n);
  item_blocks = splat_item_blocks(n, g);
  const int nbins = splat_bins(g), ntiles = nbins / 4;
  stat_blocks = static_cast<int>(std::min<long long>((n + kThreads - 1) / kThreads, 2048));
  ArenaPlan sp;
  sp.add<float2>(n);  // y0
  sp.add<float2>(n);  // y1
  sp.add<float2>(n);  // gains
  sp.add<float2>(n);  // upd
  sp.add<float2>(2 * static_cast<size_t>(n));  // diag
  sp.add<int>(static_cast<size_t>(splat_blocks) * nbins);  // rows
  sp.add<int>(nbins);                     // bin_count
  sp.add<int>(nbins + 1);                 // bin_start
  sp.add<int>(ntiles + 1);                // tile_items
  sp.add<float2>(n);                      // binned
  sp.add<unsigned long long>(3 * gg);     // wide
  sp.add<float4>(gg);                     // sgrid
  sp.add<float2>(gf);  // z1
  sp.add<float2>(gf);  // z2
  sp.add<float2>(static_cast<size_t>(g + 1) * g);  // kt
  sp.add<float2>(gf);  // o1
  sp.add<float2>(gf);  // o2
  sp.add<float4>(gg);  // field
  sp.add<float2>(fft_tw_size(f));  // twiddles
  sp.add<TsneStats>(1);
  sp.add<LayoutPartial>(stat_blocks);
  sp.add<double>(g + 1);
  if (state.init(sp.bytes, "t-SNE optimizer state")) return 1;
  y[0] = state.take<float2>(n);
  y[1] = state.take<float2>(n);
  gains = state.take<float2>(n);
  upd = state.take<float2>(n);
  diag = state.take<float2>(2 * static_cast<size_t>(n));
  splat.rows = state.take<int>(static_cast<size_t>(splat_blocks) * nbins);
  splat.bin_count = state.take<int>(nbins);
  splat.bin_start = state.take<int>(nbins + 1);
  splat.tile_items = state.take<int>(ntiles + 1);
  splat.binned = state.take<float2>(n);
  splat.wide = state.take<unsigned long long>(3 * gg);
  splat.sgrid = state.take<float4>(gg);
  z1 = state.take<float2>(gf);
  z2 = state.take<float2>(gf);
  kt = state.take<float2>(static_cast<size_t>(g + 1) * g);
  o1 = state.take<float2>(gf);
  o2 = state.take<float2>(gf);
  field = state.take<float4>(gg);
  tw = state.take<float2>(fft_tw_size(f));
  st = state.take<TsneStats>(1);
  lpart = state.take<LayoutPartial>(stat_blocks);
  zpart = state.take<double>(g + 1);

  std::vector<float2> htw(fft_tw_size(f));
  fft_twiddles(f, htw.data());
  FE_CHECK(hipMemcpyAsync(tw, htw.data(), htw.size() * sizeof(float2), hipMemcpyHostToDevice, s));
  std::vector<float> y0(2 * static_cast<size_t>(n));
  if (init) {
    std::memcpy(y0.data(), init, y0.size() * sizeof(float));
  } else {
    // tsne_neighbors.cpp:1944-1950
    std::mt19937 rng(seed);
    std::normal_distribution<float> normal(0.0f, 1.0e-4f);
    for (float& v : y0) v = normal(rng);
  }
  for (float v : y0)
    if (!std::isfinite(v)) return fail("`Y_init` must contain only finite coordinates.");
  FE_CHECK(hipMemcpyAsync(y[0], y0.data(), y0.size() * sizeof(float), hipMemcpyHostToDevice, s));
  std::vector<float2> ones(n, make_float2(1.0f, 1.0f));
  FE_CHECK(hipMemcpyAsync(gains, ones.data(), n * sizeof(float2), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemsetAsync(upd, 0, n * sizeof(float2), s));
  FE_CHECK(hipMemsetAsync(splat.wide, 0, 3 * gg * sizeof(unsigned long long), s));
  FE_CHECK(hipMemsetAsync(st, 0, sizeof(TsneStats), s));
  layout_stats_kernel<<<stat_blocks, kThreads, 0, s>>>(y[0], lpart, n);
  FE_LAUNCH_CHECK("layout_stats_kernel");
  layout_finalize<<<1, kThreads, 0, s>>>(lpart, stat_blocks, st, n, g, centre ? 1 : 0);
  FE_LAUNCH_CHECK("layout_finalize");
  if (f > 1024) {
    // conv_cols uses four padded arrays of dynamic LDS
    const int bytes = static_cast<int>(conv_cols_lds<2048>());
    hipError_t rc = hipSuccess;
    if (f == 2048)
      rc = hipFuncSetAttribute(reinterpret_cast<const void*>(&conv_cols<2048>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (check(rc, "hipFuncSetAttribute(conv_cols)")) return 1;
  }
  if (exact && ensure_exact()) return 1;
  FE_CHECK(hipStreamSynchronize(s));
  setup_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  cur = 0;
  return 0;
}

template <int F>
static int launch_conv(TsnePlan& p, hipStream_t s) {
  constexpr int G = F / 2;
  conv_rows_fwd<F><<<G + G / 2, kFftThreads, 0, s>>>(p.splat, p.st, p.tw, p.z1, p.z2, p.kt);
  FE_LAUNCH_CHECK("conv_rows_fwd");
  conv_cols<F><<<conv_cols_blocks(G), kFftThreads, conv_cols_lds<F>(), s>>>(p.z1, p.z2, p.kt, p.tw,
                                                                             p.o1, p.o2,
                                                                             p.zpart);
  FE_LAUNCH_CHECK("conv_cols");
  conv_rows_inv<F><<<G, kFftThreads, 0, s>>>(p.o1, p.o2, p.tw, p.zpart, p.st, p.field);
  FE_LAUNCH_CHECK("conv_rows_inv");
  return 0;
}

static int launch_splat(TsnePlan& p, hipStream_t s) {
  const float2* yin = p.y[p.cur];
  splat_count<<<p.splat_blocks, kSplatThreads, 0, s>>>(yin, p.st, p.splat, p.n, p.g);
  FE_LAUNCH_CHECK("splat_count");
  splat_colscan<<<splat_colscan_blocks(p.g), kSplatThreads, 0, s>>>(p.splat, p.g, p.splat_blocks);
  FE_LAUNCH_CHECK("splat_colscan");
  splat_scatter<<<p.splat_blocks, kSplatThreads, 0, s>>>(yin, p.st, p.splat, p.n, p.g);
  FE_LAUNCH_CHECK("splat_scatter");
  splat_tiles<<<p.item_blocks, kThreads, 0, s>>>(p.st, p.splat, p.g);
  FE_LAUNCH_CHECK("splat_tiles");
  return 0;
}

static int launch_conv_any(TsnePlan& p, hipStream_t s) {
  switch (p.f) {
    case 32: return launch_conv<32>(p, s);
    case 64: return launch_conv<64>(p, s);
    case 128: return launch_conv<128>(p, s);
    case 256: return launch_conv<256>(p, s);
    case 512: return launch_conv<512>(p, s);
    case 1024: return launch_conv<1024>(p, s);
    case 2048: return launch_conv<2048>(p, s);
    default: return fail("unsupported FFT size");
  }
}

// attraction into diag[0, n), then the per-point gather/update (MODE 0) or
// the repulsive force into diag[n, 2n) (MODE 1).  The pair form leaves its
// sums in partial slots; with to_array = false they stay there for the
// optimiser update (launch_update<0>), which reads them itself.
static int launch_attract(TsnePlan& p, float exaggeration, hipStream_t s, bool to_array = true) {
  const float2* yin = p.y[p.cur];
  if (p.attr_kind == kAttrPairs) {
    attract_pairs<<<p.pair_ranges, kPairThreads, kPairLds, s>>>(yin, p.pairs, p.st, p.n);
    FE_LAUNCH_CHECK("attract_pairs");
    if (to_array) {
      pairs_reduce<<<p.stat_blocks, kThreads, 0, s>>>(p.pairs, p.st, exaggeration, p.diag, p.n);
      FE_LAUNCH_CHECK("pairs_reduce");
    }
    return 0;
  }
  if (p.attr_kind == kAttrStaged) {
    attract_blocked<true><<<p.groups, kStagedThreads, kStagedLds, s>>>(yin, p.blk, p.st,
                                                                       exaggeration, p.diag, p.n);
    FE_LAUNCH_CHECK("attract_blocked");
    return 0;
  }
  if (p.attr_kind == kAttrLanes) {
    attract_blocked<false><<<p.groups, kStagedThreads, 0, s>>>(yin, p.blk, p.st, exaggeration,
                                                               p.diag, p.n);
    FE_LAUNCH_CHECK("attract_blocked");
    return 0;
  }
#define FE_ATTR(T)                                                                     \
  attract_kernel<T><<<p.upd_blocks, kThreads, 0, s>>>(yin, p.csr.row_ptr, p.csr.col,   \
                                                      p.csr.val, p.st, exaggeration,   \
                                                      p.diag, p.n)
  switch (p.tpr) {
    case 4: FE_ATTR(4); break;
    case 8: FE_ATTR(8); break;
    case 16: FE_ATTR(16); break;
    case 32: FE_ATTR(32); break;
    default: FE_ATTR(64); break;
  }
#undef FE_ATTR
  FE_LAUNCH_CHECK("attract_kernel");
  return 0;
}

template <int MODE>
static int launch_update(TsnePlan& p, const StepParams& sp, hipStream_t s) {
  if (MODE == 0 && p.attr_kind == kAttrPairs) {
    update_pairs_kernel<<<p.stat_blocks, kThreads, 0, s>>>(p.y[p.cur], p.y[p.cur ^ 1], p.gains,
                                                           p.upd, p.pairs, p.field, p.st, p.lpart,
                                                           p.n, p.g, sp);
    FE_LAUNCH_CHECK("update_pairs_kernel");
  } else {
    update_kernel<MODE><<<p.stat_blocks, kThreads, 0, s>>>(p.y[p.cur], p.y[p.cur ^ 1], p.gains,
                                                           p.upd, p.diag, p.field, p.st, p.lpart,
                                                           p.n, p.g, sp, p.diag + p.n);
    FE_LAUNCH_CHECK("update_kernel");
  }
  if constexpr (MODE == 0) {
    layout_finalize<<<1, kThreads, 0, s>>>(p.lpart, p.stat_blocks, p.st, p.n, p.g, 1);
    FE_LAUNCH_CHECK("layout_finalize");
  }
  return 0;
}

// Exact-route buffers: the j range of every point is cut into `splits`
// slices of `chunk` points (enough blocks to fill the GPU at small n).
int TsnePlan::ensure_exact() {
  if (ex_part) return 0;
  ex_bx = (n + kThreads - 1) / kThreads;
  const int want = std::max(1, (4 * compute_units() + ex_bx - 1) / ex_bx);
  ex_splits = std::min({want, 64, std::max(1, n / 256)});
  ex_chunk = (n + ex_splits - 1) / ex_splits;
  ArenaPlan ep;
  ep.add<double2>(static_cast<size_t>(ex_splits) * n);
  ep.add<double>(static_cast<size_t>(ex_splits) * ex_bx);
  ep.add<double>(ex_bx);
  if (exact_store.init(ep.bytes, "t-SNE exact repulsion partials")) return 1;
  ex_part = exact_store.take<double2>(static_cast<size_t>(ex_splits) * n);
  ex_zpart = exact_store.take<double>(static_cast<size_t>(ex_splits) * ex_bx);
  kl_part = exact_store.take<double>(ex_bx);
  return 0;
}

int TsnePlan::launch_exact(bool forces) {
  hipStream_t s = stream.get();
  const dim3 grid(ex_bx, ex_splits);
  if (forces)
    exact_rep_kernel<true><<<grid, kThreads, 0, s>>>(y[cur], st, ex_part, ex_zpart, n, ex_chunk);
  else
    exact_rep_kernel<false><<<grid, kThreads, 0, s>>>(y[cur], st, ex_part, ex_zpart, n, ex_chunk);
  FE_LAUNCH_CHECK("exact_rep_kernel");
  return 0;
}

// KL of the current layout with the exact Z (evaluate_kl_f): the Z partials,
// folded in their slice-major block order on the host and rounded to float
// as compute_sum_q_f returns Z, then the row sums.
int TsnePlan::kl(double* kl_out, double* sum_q_out) {
  if (ensure_exact() || launch_exact(false)) return 1;
  hipStream_t s = stream.get();
  const size_t nz = static_cast<size_t>(ex_splits) * ex_bx;
  std::vector<double> zp(nz), kp(ex_bx);
  FE_CHECK(hipMemcpyAsync(zp.data(), ex_zpart, nz * sizeof(double), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  double z = 0.0;
  for (double v : zp) z += v;
  const double zf = static_cast<double>(static_cast<float>(std::fmax(z, static_cast<double>(FLT_MIN))));
  kl_rows_kernel<<<ex_bx, kThreads, 0, s>>>(y[cur], st, csr.row_ptr, csr.col, csr.val, n, zf,
                                            kl_part);
  FE_LAUNCH_CHECK("kl_rows_kernel");
  FE_CHECK(hipMemcpyAsync(kp.data(), kl_part, ex_bx * sizeof(double), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  double total = 0.0;
  for (double v : kp) total += v;
  if (kl_out) *kl_out = total;
  if (sum_q_out) *sum_q_out = zf;
  return 0;
}

int TsnePlan::enqueue_iteration(const StepParams& sp) {
  hipStream_t s = stream.get();
  if (exact) {
    if (ensure_exact() || launch_exact(true)) return 1;
    if (launch_attract(*this, sp.exaggeration, s)) return 1;
    update_exact_kernel<0><<<stat_blocks, kThreads, 0, s>>>(
        y[cur], y[cur ^ 1], gains, upd, diag, ex_part, ex_splits, ex_zpart, ex_splits * ex_bx, st,
        lpart, n, sp, nullptr);
    FE_LAUNCH_CHECK("update_exact_kernel");
    layout_finalize<<<1, kThreads, 0, s>>>(lpart, stat_blocks, st, n, g, 1);
    FE_LAUNCH_CHECK("layout_finalize");
    cur ^= 1;
    return 0;
  }
  if (overlap) {
    hipStream_t a = side.get();
    hipEvent_t fork = forks.next(), join = forks.next();
    FE_CHECK(hipEventRecord(fork, s));
    // the splat is enqueued (in a capture: its nodes created) before the
    // attraction, so a replayed graph keeps the splat and FFT chain, the
    // critical path, on the first stream and the cross-stream start latency
    // falls on the attraction, which has slack (1M points: finalize-to-splat
    // gap 10.5 -> 4.9 us, join 9.5 -> 6.2 us per iteration)
    if (launch_splat(*this, s)) return 1;
    FE_CHECK(hipStreamWaitEvent(a, fork, 0));
    if (launch_attract(*this, sp.exaggeration, a, false)) return 1;
    FE_CHECK(hipEventRecord(join, a));
    if (launch_conv_any(*this, s)) return 1;
    FE_CHECK(hipStreamWaitEvent(s, join, 0));
  } else {
    if (launch_splat(*this, s)) return 1;
    if (launch_conv_any(*this, s)) return 1;
    if (launch_attract(*this, sp.exaggeration, s, false)) return 1;
  }
  if (launch_update<0>(*this, sp, s)) return 1;
  cur ^= 1;
  return 0;
}

int TsnePlan::enqueue_diag(const StepParams& sp) {
  hipStream_t s = stream.get();
  if (launch_splat(*this, s)) return 1;
  if (launch_conv_any(*this, s)) return 1;
  if (launch_attract(*this, sp.exaggeration, s)) return 1;
  return launch_update<1>(*this, sp, s);
}

static bool same_params(const StepParams& a, const StepParams& b) {
  return std::memcmp(&a, &b, sizeof(StepParams)) == 0;
}

// Runs are cut into chunks of at most kGraphChunk iterations; each chunk of
// at least kGraphMin iterations replays a captured graph of that length,
// constants and starting buffer parity (an odd chunk flips the parity).
constexpr int kGraphChunk = 50;
static_assert(2 * kGraphChunk <= kForkEvents, "a captured chunk reuses no fork event");
constexpr int kGraphMin = 4;
constexpr size_t kGraphCache = 16;

GraphExec* TsnePlan::chunk_graph(int len, const StepParams& sp) {
  for (auto& c : graphs)
    if (c.iters == len && c.parity == cur && same_params(c.sp, sp)) return c.exec.get();
  hipStream_t s = stream.get();
  hipGraph_t graph = nullptr;
  bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
  const int cur0 = cur;
  for (int it = 0; ok && it < len; ++it) ok = enqueue_iteration(sp) == 0;
  const hipError_t end = hipStreamEndCapture(s, &graph);
  cur = cur0;
  GraphExec* exec = nullptr;
  if (ok && end == hipSuccess && graph) {
    CachedGraph c{sp, len, cur0, std::make_unique<GraphExec>()};
    if (c.exec->instantiate(graph) == 0) {
      if (graphs.size() >= kGraphCache) graphs.erase(graphs.begin());
      exec = c.exec.get();
      graphs.push_back(std::move(c));
    }
  } else if (graph) {
    (void)hipGraphDestroy(graph);
  }
  (void)hipGetLastError();
  clear_error();
  return exec;
}

int TsnePlan::run(int iters, const StepParams& sp) {
  hipStream_t s = stream.get();
  const bool capture = graph_capture_enabled();
  for (int left = iters; left > 0;) {
    const int len = std::min(left, kGraphChunk);
    GraphExec* exec = capture && len >= kGraphMin ? chunk_graph(len, sp) : nullptr;
    if (exec) {
      FE_CHECK(hipGraphLaunch(exec->get(), s));
      cur ^= len & 1;
    } else {
      for (int it = 0; it < len; ++it)
        if (enqueue_iteration(sp)) return 1;
    }
    left -= len;
  }
  iters_run += iters;
  return 0;
}

// Captures the graphs that run(iters, sp) will replay when it starts after
// `after` more iterations, without running anything.
int TsnePlan::prepare(int iters, const StepParams& sp, long long after) {
  if (!graph_capture_enabled()) return 0;
  const int cur0 = cur;
  cur ^= static_cast<int>(after & 1);
  for (int left = iters; left > 0;) {
    const int len = std::min(left, kGraphChunk);
    if (len >= kGraphMin && !chunk_graph(len, sp)) {
      cur = cur0;
      return fail("iteration graph capture failed");
    }
    cur ^= len & 1;
    left -= len;
  }
  cur = cur0;
  return 0;
}

int TsnePlan::layout(float* out_host) {
  hipStream_t s = stream.get();
  centred_copy_kernel<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(y[cur], st, diag, n);
  FE_LAUNCH_CHECK("centred_copy_kernel");
  FE_CHECK(hipMemcpyAsync(out_host, diag, 2 * static_cast<size_t>(n) * sizeof(float),
                          hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  return 0;
}

// kernel_ms[0..5]: splat, conv_rows_fwd, conv_cols, conv_rows_inv,
// attract, update; kernel_ms[6]: whole iteration
int TsnePlan::profile(int iters, const StepParams& sp, double* ms) {
  constexpr int K = 6;
  if (exact) return fail("kernel profile covers the FFT-grid route only");
  hipStream_t s = stream.get();
  std::vector<hipEvent_t> ev((K + 1) * iters);
  for (auto& e : ev) FE_CHECK(hipEventCreate(&e));
  for (int it = 0; it < iters; ++it) {
    hipEvent_t* e = &ev[(K + 1) * it];
    FE_CHECK(hipEventRecord(e[0], s));
    if (launch_splat(*this, s)) return 1;
    FE_CHECK(hipEventRecord(e[1], s));
    int rc = 0;
#define FE_PROF_CONV(FF)                                                                      \
  case FF: {                                                                                  \
    constexpr int G = FF / 2;                                                                 \
    conv_rows_fwd<FF><<<G + G / 2, kFftThreads, 0, s>>>(splat, st, tw, z1, z2, kt);              \
    (void)hipEventRecord(e[2], s);                                                            \
    conv_cols<FF><<<conv_cols_blocks(G), kFftThreads, conv_cols_lds<FF>(), s>>>(            \
        z1, z2, kt, tw, o1, o2, zpart);                                                       \
    (void)hipEventRecord(e[3], s);                                                            \
    conv_rows_inv<FF><<<G, kFftThreads, 0, s>>>(o1, o2, tw, zpart, st, field);                   \
    break;                                                                                    \
  }
    switch (f) {
      FE_PROF_CONV(32)
      FE_PROF_CONV(64)
      FE_PROF_CONV(128)
      FE_PROF_CONV(256)
      FE_PROF_CONV(512)
      FE_PROF_CONV(1024)
      FE_PROF_CONV(2048)
      default: rc = 1;
    }
#undef FE_PROF_CONV
    if (rc) return fail("unsupported FFT size");
    FE_CHECK(hipEventRecord(e[4], s));
    if (launch_attract(*this, sp.exaggeration, s, false)) return 1;
    FE_CHECK(hipEventRecord(e[5], s));
    if (launch_update<0>(*this, sp, s)) return 1;
    FE_CHECK(hipEventRecord(e[6], s));
    cur ^= 1;
  }
  FE_CHECK(hipStreamSynchronize(s));
  for (int q = 0; q <= K; ++q) ms[q] = 0.0;
  for (int it = 0; it < iters; ++it) {
    hipEvent_t* e = &ev[(K + 1) * it];
    for (int q = 0; q < K; ++q) {
      float t = 0.f;
      FE_CHECK(hipEventElapsedTime(&t, e[q], e[q + 1]));
      ms[q] += t;
    }
    float t = 0.f;
    FE_CHECK(hipEventElapsedTime(&t, e[0], e[K]));
    ms[K] += t;
  }
  for (int q = 0; q <= K; ++q) ms[q] /= iters;
  for (auto& e : ev) (void)hipEventDestroy(e);
  iters_run += iters;
  return 0;
}

// ---------------------------------------------------------------------------
// shared validation (embedding_cuda_kernels.cpp:5656-5672,
// embedding_cuda_impl.cpp:2265-2301)
// ---------------------------------------------------------------------------
static int validate_tsne(const void* idx, const void* dist, int n, int k, int n_components,
                         float perplexity, int early_iter, int n_iter, float early_exag,
                         float exag, float lr, int lr_auto, float mom0, float mom1,
                         float min_gain, int index_offset) {
  if (!idx || !dist) return fail("null HIP openTSNE input pointer");
  if (n < 2 || k < 1) return fail("KNN input must have at least two rows and one neighbor column.");
  if (k > kMaxK) return fail("HIP openTSNE FFT-grid currently supports at most 256 neighbors.");
  if (n_components != 2)
    return fail("HIP openTSNE FFT-grid currently supports exactly two output components.");
  if (static_cast<long long>(n) * k * 2 > 2147483647ll)
    return fail("n * k is too large for 32-bit CSR offsets.");
  if (!(perplexity > 0.0f)) return fail("perplexity must be positive");
  if (early_iter < 0 || n_iter < 0 || early_iter + n_iter < 1)
    return fail("HIP openTSNE FFT-grid requires at least one optimization iteration.");
  if (!(early_exag > 0.0f) || !(exag > 0.0f))
    return fail("HIP openTSNE FFT-grid exaggeration values must be positive.");
  if (!lr_auto && !(lr > 0.0f)) return fail("HIP openTSNE FFT-grid learning rate must be positive.");
  if (mom0 < 0.0f || mom1 < 0.0f)
    return fail("HIP openTSNE FFT-grid momentum values must be non-negative.");
  if (!(min_gain > 0.0f)) return fail("HIP openTSNE FFT-grid min_gain must be positive.");
  if (index_offset != 0 && index_offset != 1) return fail("index_offset must be 0 or 1");
  return 0;
}

// The C-ABI's attraction choice (0 built in, 1 row groups, 2 segment-blocked
// LDS windows, 3 every pair once) as a plan request.
static int attraction_request(int attraction, int* request) {
  switch (attraction) {
    case 0: *request = -1; return 0;
    case 1: *request = kAttrRows; return 0;
    case 2: *request = kAttrStaged; return 0;
    case 3: *request = kAttrPairs; return 0;
    default: return fail("attraction must be 0 (built in), 1 (rows), 2 (staged) or 3 (pairs)");
  }
}

int tsne_job_validate(const fastembedr_hip_tsne_job& j) {
  int request = -1;
  if (attraction_request(j.attraction, &request)) return 1;
  if (j.pair_workgroups < 0) return fail("pair_workgroups must be non-negative");
  if (j.grid_size < 0) return fail("batched jobs run the FFT-grid route (grid_size >= 0)");
  if (j.grid_size > 0 && (j.grid_size < 16 || j.grid_size > 1024 || (j.grid_size & (j.grid_size - 1))))
    return fail("grid_size must be a power of two between 16 and 1024");
  if (validate_tsne(j.indices, j.distances, j.n, j.k, 2, j.perplexity, j.early_exaggeration_iter,
                    j.n_iter, j.early_exaggeration, j.exaggeration, j.learning_rate_early, 0,
                    j.initial_momentum, j.final_momentum, j.min_gain, j.index_offset))
    return 1;
  if (!(j.learning_rate_normal > 0.0f)) return fail("HIP openTSNE FFT-grid learning rate must be positive.");
  return 0;
}

template <typename DistT>
static int tsne_one_shot(const int* indices, const DistT* distances, bool device_inputs,
                         const float* init, int has_init, int n, int k, int n_components,
                         float perplexity, int early_iter, int n_iter, float early_exag,
                         float exag, float lr, int lr_auto, float mom0, float mom1, float min_gain,
                         float max_step, unsigned seed, int index_offset, int grid_size,
                         float* out) {
  clear_error();
  if (validate_tsne(indices, distances, n, k, n_components, perplexity, early_iter, n_iter,
                    early_exag, exag, lr, lr_auto, mom0, mom1, min_gain, index_offset))
    return 1;
  if (!out || (has_init && !init)) return fail("null HIP openTSNE input pointer");
  TsnePlan plan;
  if (plan.create<DistT>(indices, distances, device_inputs, n, k, index_offset, perplexity,
                         grid_size, has_init ? init : nullptr, seed, true))
    return 1;
  const float lr_early = lr_auto ? static_cast<float>(n) / early_exag : lr;
  const float lr_normal = lr_auto ? static_cast<float>(n) / exag : lr;
  if (early_iter > 0 &&
      plan.run(early_iter, StepParams{early_exag, lr_early, mom0, min_gain, max_step}))
    return 1;
  if (n_iter > 0 && plan.run(n_iter, StepParams{exag, lr_normal, mom1, min_gain, max_step}))
    return 1;
  return plan.layout(out);
}

// ---------------------------------------------------------------------------
// Batched jobs (job-table kernels of tsne_batch.hpp) with HBM-budget admission
// ---------------------------------------------------------------------------

// Device bytes a job holds while resident (P at its worst case of 2*n*k
// entries) and the transient peak while its P is built.
static long long job_resident_bytes(int n, int k, int g) {
  const long long f = 2ll * g, gg = 1ll * g * g, nk = 1ll * n * k;
  // P, its padded segments (at most kPasses per row) with their tables, the
  // layout state, the splat bins and grids, the FFT buffers
  const long long segs = std::min<long long>(2 * nk, 1ll * kPasses * n);
  const long long seg_nnz = 2 * nk + 1ll * kLaneEntries * segs + kLanePad;
  const long long groups = std::max<long long>(32, seg_nnz / (16 * 4096) + 1);
  const long long blocked = 10ll * seg_nnz + seg_nnz / 2 + 8ll * (seg_nnz / kLaneStep + 4) +
                            4ll * groups * (1 + 2 * kMaxWindows + 2 * kPasses);
  // or the pair streams: at most nk pairs staged in runs padded per row and
  // tile, a rest of at most nk entries both ways (else the blocked form is
  // chosen), the chunk and slot tables and 32 KB per partial slot
  const long long nb = (n + kPairBlock - 1) / kPairBlock;
  const long long tiles = nb * (nb + 1) / 2;
  const long long staged = nk + 1ll * kLaneEntries * (n + tiles) + kPairPad;
  const long long rest = nk + 1ll * kLaneEntries * n + kPairPad;
  // (the cuts at workgroup range boundaries at most double the chunks)
  const long long chunks = 2 * (staged / kPairChunkEntries + rest / kPairChunkEntries + tiles + nb);
  const long long pairs = 6ll * staged + staged / 4 + 8ll * rest + rest / 4 +
                          static_cast<long long>(sizeof(PairChunk)) * chunks + 8ll * chunks +
                          4ll * (nb + 1) + 2ll * chunks * kPairBlock * 8;
  return 8ll * 2 * nk + 4ll * (n + 1) + std::max(blocked, pairs) +
         8ll * 7 * n + 16ll * (splat_bins(g) + 1) + 4ll * splat_bin_blocks(n) * splat_bins(g) +
         40ll * gg +
         8ll * 4 * g * f + 8ll * (g + 1) * g + 16ll * gg + 8ll * fft_tw_size(2 * g) + 32ll * 2048 +
         8ll * (g + 1) +
         4096;
}

static long long job_setup_bytes(int n, int k) {
  const long long nk = 1ll * n * k;
  return 8ll * nk /* upload */ + 8ll * nk /* nbr, pdir */ + 16ll * nk /* raw */ + 16ll * n +
         8ll * kPasses * n /* segment scan */;
}

// launch-table parameters of the segment-blocked attraction kernels (the
// row-group kernel's launches carry their lanes per row)
constexpr int kStagedLaunch = -1000;
constexpr int kLanesLaunch = -1001;
constexpr int kPairsLaunch = -1002;

class TsneBatch {
 public:
  int max_jobs = 0;
  long long budget = 0, used = 0;
  Stream stream;
  std::vector<std::unique_ptr<TsnePlan>> slots;
  std::vector<long long> slot_bytes;
  std::vector<JobDesc> desc;
  Arena dev;
  JobDesc* d_desc = nullptr;
  int* d_tab = nullptr;
  int tab_cap = 0;
  bool dirty = true;
  struct Launch {
    int kind;   // 0 splat_count, 11 splat_colscan, 6 splat_scatter, 7 splat_tiles,
                // 1 rows_fwd, 2 cols, 3 rows_inv, 4 attract, 5 update, 9 finalize
    int param;  // F or TPR
    int off;    // offset of slot[] in d_tab; first[] follows it
    int m;      // entries
    int blocks;
  };
  std::vector<Launch> launches;
  std::unique_ptr<GraphExec> graph;  // captured iterations of the current table
  // the attraction launches run on `side`, beside the splat and FFT launches
  // of every job on `stream`, as in a single plan; FASTEMBEDR_HIP_OVERLAP=0
  // keeps every launch on `stream`
  bool overlap = false;
  Stream side;
  EventRing forks;

  int init(int max_jobs_, long long budget_) {
    max_jobs = max_jobs_;
    budget = budget_;
    if (stream.init()) return 1;
    overlap = overlap_enabled();
    if (overlap && (side.init() || forks.init(kForkEvents))) return 1;
    slots.resize(max_jobs);
    slot_bytes.assign(max_jobs, 0);
    desc.resize(max_jobs);
    std::memset(desc.data(), 0, desc.size() * sizeof(JobDesc));
    tab_cap = 2 * (11 * max_jobs + 64) + 64;  // every job is in at most 11 launches
    ArenaPlan p;
    p.add<JobDesc>(max_jobs);
    p.add<int>(tab_cap);
    if (dev.init(p.bytes, "t-SNE batch tables")) return 1;
    d_desc = dev.take<JobDesc>(max_jobs);
    d_tab = dev.take<int>(tab_cap);
    FE_CHECK(hipMemsetAsync(d_desc, 0, max_jobs * sizeof(JobDesc), stream.get()));
    return check(hipStreamSynchronize(stream.get()), "batch init");
  }

  // returns 0 admitted, 1 error, 2 not admitted (budget or no free slot)
  int add(const fastembedr_hip_tsne_job& j, bool device_inputs, int* slot_out) {
    int slot = -1;
    for (int s = 0; s < max_jobs; ++s)
      if (!slots[s]) {
        slot = s;
        break;
      }
    if (slot < 0) {
      set_error("no free batch slot");
      return 2;
    }
    if (j.grid_size < 0) return fail("batched jobs run the FFT-grid route (grid_size >= 0)");
    const int g = resolve_grid_size(j.n, j.grid_size);
    const long long need = job_resident_bytes(j.n, j.k, g);
    if (budget > 0 && used + need + job_setup_bytes(j.n, j.k) > budget) {
      set_error("HBM budget does not admit this job yet");
      return 2;
    }
    if (tsne_job_validate(j)) return 1;
    auto plan = std::make_unique<TsnePlan>();
    if (attraction_request(j.attraction, &plan->attr_request)) return 1;
    plan->pair_wg_request = j.pair_workgroups;
    plan->allow_overlap = false;
    if (plan->create<float>(j.indices, j.distances, device_inputs, j.n, j.k, j.index_offset,
                            j.perplexity, j.grid_size, j.init, j.seed, true))
      return 1;
    // the estimate admitted it; what it actually holds decides
    if (budget > 0 && used + plan->device_bytes() > budget) {
      set_error("HBM budget does not admit this job yet");
      return 2;
    }
    JobDesc d;
    std::memset(&d, 0, sizeof(d));
    TsnePlan& p = *plan;
    d.y[0] = p.y[0];
    d.y[1] = p.y[1];
    d.gains = p.gains;
    d.upd = p.upd;
    d.row_ptr = p.csr.row_ptr;
    d.col = p.csr.col;
    d.val = p.csr.val;
    d.blk = p.blk;
    d.groups = p.groups;
    d.pairs = p.pairs;
    d.pair_ranges = p.pair_ranges;
    d.splat = p.splat;
    d.z1 = p.z1;
    d.z2 = p.z2;
    d.kt = p.kt;
    d.o1 = p.o1;
    d.o2 = p.o2;
    d.field = p.field;
    d.tw = p.tw;
    d.st = p.st;
    d.lpart = p.lpart;
    d.zpart = p.zpart;
    d.attr = p.diag;
    d.pair_attr = p.attr_kind == kAttrPairs ? 1 : 0;
    d.n = p.n;
    d.g = p.g;
    d.upd_blocks = p.stat_blocks;
    d.cur = p.cur;
    d.iter = 0;
    d.early_iter = j.early_exaggeration_iter;
    d.total_iter = j.early_exaggeration_iter + j.n_iter;
    d.active = 1;
    d.early_exag = j.early_exaggeration;
    d.exag = j.exaggeration;
    d.lr_early = j.learning_rate_early;
    d.lr_normal = j.learning_rate_normal;
    d.mom0 = j.initial_momentum;
    d.mom1 = j.final_momentum;
    d.min_gain = j.min_gain;
    d.max_step = j.max_step_norm;
    desc[slot] = d;
    FE_CHECK(hipMemcpyAsync(d_desc + slot, &desc[slot], sizeof(JobDesc), hipMemcpyHostToDevice,
                            stream.get()));
    FE_CHECK(hipStreamSynchronize(stream.get()));
    slot_bytes[slot] = p.device_bytes();
    used += slot_bytes[slot];
    slots[slot] = std::move(plan);
    dirty = true;
    *slot_out = slot;
    return 0;
  }

  int remove(int slot) {
    if (slot < 0 || slot >= max_jobs || !slots[slot]) return fail("invalid batch slot");
    FE_CHECK(hipStreamSynchronize(stream.get()));
    JobDesc z;
    std::memset(&z, 0, sizeof(z));
    desc[slot] = z;
    FE_CHECK(hipMemcpy(d_desc + slot, &z, sizeof(JobDesc), hipMemcpyHostToDevice));
    slots[slot].reset();
    used -= slot_bytes[slot];
    slot_bytes[slot] = 0;
    dirty = true;
    return 0;
  }

  int rebuild() {
    launches.clear();
    std::vector<int> tab;
    auto add_launch = [&](int kind, int param, const std::vector<std::pair<int, int>>& ent) {
      if (ent.empty()) return;
      Launch L{kind, param, static_cast<int>(tab.size()), static_cast<int>(ent.size()), 0};
      for (auto& e : ent) tab.push_back(e.first);
      int acc = 0;
      for (auto& e : ent) {
        tab.push_back(acc);
        acc += e.second;
      }
      tab.push_back(acc);
      L.blocks = acc;
      launches.push_back(L);
    };
    std::vector<std::pair<int, int>> splat, tiles, single;
    for (int s = 0; s < max_jobs; ++s)
      if (slots[s]) {
        splat.push_back({s, slots[s]->splat_blocks});
        tiles.push_back({s, slots[s]->item_blocks});
        single.push_back({s, 1});
      }
    std::vector<std::pair<int, int>> colscan;
    for (int s = 0; s < max_jobs; ++s)
      if (slots[s]) colscan.push_back({s, splat_colscan_blocks(slots[s]->g)});
    add_launch(0, 0, splat);
    add_launch(11, 0, colscan);
    add_launch(6, 0, splat);
    add_launch(7, 0, tiles);
    for (int kind = 1; kind <= 3; ++kind) {
      for (int F = 32; F <= 2048; F *= 2) {
        std::vector<std::pair<int, int>> ent;
        for (int s = 0; s < max_jobs; ++s) {
          if (!slots[s] || slots[s]->f != F) continue;
          const int G = F / 2;
          ent.push_back({s, kind == 1 ? G + G / 2 : (kind == 2 ? conv_cols_blocks(G) : G)});
        }
        add_launch(kind, F, ent);
      }
    }
    {
      std::vector<std::pair<int, int>> ent;
      for (int s = 0; s < max_jobs; ++s)
        if (slots[s] && slots[s]->attr_kind == kAttrStaged) ent.push_back({s, slots[s]->groups});
      add_launch(4, kStagedLaunch, ent);
    }
    {
      std::vector<std::pair<int, int>> ent;
      for (int s = 0; s < max_jobs; ++s)
        if (slots[s] && slots[s]->attr_kind == kAttrLanes) ent.push_back({s, slots[s]->groups});
      add_launch(4, kLanesLaunch, ent);
    }
    {
      std::vector<std::pair<int, int>> ent;
      for (int s = 0; s < max_jobs; ++s)
        if (slots[s] && slots[s]->attr_kind == kAttrPairs) ent.push_back({s, slots[s]->pair_ranges});
      add_launch(4, kPairsLaunch, ent);
    }
    for (int tpr = 4; tpr <= 64; tpr *= 2) {
      std::vector<std::pair<int, int>> ent;
      for (int s = 0; s < max_jobs; ++s)
        if (slots[s] && slots[s]->attr_kind == kAttrRows && slots[s]->tpr == tpr)
          ent.push_back({s, slots[s]->upd_blocks});
      add_launch(4, tpr, ent);
    }
    std::vector<std::pair<int, int>> upd;
    for (int s = 0; s < max_jobs; ++s)
      if (slots[s]) upd.push_back({s, slots[s]->stat_blocks});
    add_launch(5, 0, upd);
    add_launch(9, 0, single);
    if (static_cast<int>(tab.size()) > tab_cap) return fail("batch table overflow");
    if (!tab.empty())
      FE_CHECK(hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    dirty = false;
    return 0;
  }

  // One iteration of every job: the splat launches, then (forked onto the
  // side stream) the attraction launches, the FFT launches, and after the
  // join the update and finalize launches.  The splat is enqueued before the
  // fork, as in a single plan, so a replayed graph keeps the splat and FFT
  // chain on the first stream.
  int enqueue_iteration() {
    hipStream_t s = stream.get();
    auto stage = [](int kind) {
      return kind == 0 || kind == 11 || kind == 6 || kind == 7 ? 0 : (kind == 4 ? 1 : (kind <= 3 ? 2 : 3));
    };
    if (!overlap) {
      for (const Launch& L : launches)
        if (launch(L, s)) return 1;
      return 0;
    }
    hipStream_t a = side.get();
    hipEvent_t fork = forks.next(), join = forks.next();
    FE_CHECK(hipEventRecord(fork, s));
    for (const Launch& L : launches)
      if (stage(L.kind) == 0 && launch(L, s)) return 1;
    FE_CHECK(hipStreamWaitEvent(a, fork, 0));
    for (const Launch& L : launches)
      if (stage(L.kind) == 1 && launch(L, a)) return 1;
    FE_CHECK(hipEventRecord(join, a));
    for (const Launch& L : launches)
      if (stage(L.kind) == 2 && launch(L, s)) return 1;
    FE_CHECK(hipStreamWaitEvent(s, join, 0));
    for (const Launch& L : launches)
      if (stage(L.kind) == 3 && launch(L, s)) return 1;
    return 0;
  }

  int launch(const Launch& L, hipStream_t s) {
    {
      const int* sl = d_tab + L.off;
      const int* fi = sl + L.m;
      switch (L.kind) {
        case 0:
          batch_splat_count<<<L.blocks, kSplatThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
        case 11:
          batch_splat_colscan<<<L.blocks, kSplatThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
        case 6:
          batch_splat_scatter<<<L.blocks, kSplatThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
        case 7:
          batch_splat_tiles<<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
        case 9:
          batch_finalize<<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
#define FE_B_CONV(FF)                                                                          \
  case FF:                                                                                     \
    if (L.kind == 1)                                                                           \
      batch_rows_fwd<FF><<<L.blocks, kFftThreads, 0, s>>>(d_desc, sl, fi, L.m);                   \
    else if (L.kind == 2)                                                                      \
      batch_cols<FF><<<L.blocks, kFftThreads, conv_cols_lds<FF>(), s>>>(d_desc, sl, fi, L.m); \
    else                                                                                       \
      batch_rows_inv<FF><<<L.blocks, kFftThreads, 0, s>>>(d_desc, sl, fi, L.m);                   \
    break;
        case 1:
        case 2:
        case 3:
          switch (L.param) {
            FE_B_CONV(32)
            FE_B_CONV(64)
            FE_B_CONV(128)
            FE_B_CONV(256)
            FE_B_CONV(512)
            FE_B_CONV(1024)
            FE_B_CONV(2048)
            default:
              return fail("unsupported FFT size");
          }
          break;
#undef FE_B_CONV
        case 4:
          switch (L.param) {
            case kStagedLaunch:
              batch_attract_blocked<true><<<L.blocks, kStagedThreads, kStagedLds, s>>>(
                  d_desc, sl, fi, L.m);
              break;
            case kLanesLaunch:
              batch_attract_blocked<false><<<L.blocks, kStagedThreads, 0, s>>>(d_desc, sl, fi, L.m);
              break;
            case kPairsLaunch:
              batch_attract_pairs<<<L.blocks, kPairThreads, kPairLds, s>>>(d_desc, sl, fi, L.m);
              break;
            case 4: batch_attract<4><<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m); break;
            case 8: batch_attract<8><<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m); break;
            case 16: batch_attract<16><<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m); break;
            case 32: batch_attract<32><<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m); break;
            default: batch_attract<64><<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m); break;
          }
          break;
        case 5:
          batch_update<<<L.blocks, kThreads, 0, s>>>(d_desc, sl, fi, L.m);
          break;
        default:
          return fail("unknown batch stage");
      }
      FE_LAUNCH_CHECK("batch kernel");
    }
    return 0;
  }

  int step(int iters) {
    if (dirty) {
      graph.reset();
      if (rebuild()) return 1;
    }
    if (launches.empty()) return 0;
    static bool attr_done = false;
    if (!attr_done) {
      FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_cols<2048>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(conv_cols_lds<2048>())));
      FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_attract_blocked<true>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(kStagedLds)));
      FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_attract_pairs),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(kPairLds)));
      attr_done = true;
    }
    // Captured chunks of kChunk iterations replay the current launch table;
    // a job that finishes inside a chunk goes inactive on the device and its
    // blocks exit, so the graph stays valid until a job joins or leaves.
    constexpr int kChunk = 25;
    static_assert(2 * kChunk <= kForkEvents, "a captured chunk reuses no fork event");
    hipStream_t s = stream.get();
    int left = iters;
    if (graph_capture_enabled() && left >= kChunk) {
      if (!graph) {
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        for (int it = 0; ok && it < kChunk; ++it) ok = enqueue_iteration() == 0;
        const hipError_t end = hipStreamEndCapture(s, &g);
        if (ok && end == hipSuccess && g) {
          auto exec = std::make_unique<GraphExec>();
          if (exec->instantiate(g) == 0) graph = std::move(exec);
        } else if (g) {
          (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
        clear_error();
      }
      if (graph) {
        while (left >= kChunk) {
          FE_CHECK(hipGraphLaunch(graph->get(), s));
          left -= kChunk;
        }
      }
    }
    for (; left > 0; --left)
      if (enqueue_iteration()) return 1;
    return 0;
  }

  int status(int* iters, int* active) {
    std::vector<JobDesc> h(max_jobs);
    FE_CHECK(hipMemcpyAsync(h.data(), d_desc, max_jobs * sizeof(JobDesc), hipMemcpyDeviceToHost,
                            stream.get()));
    FE_CHECK(hipStreamSynchronize(stream.get()));
    for (int s = 0; s < max_jobs; ++s) {
      if (iters) iters[s] = slots[s] ? h[s].iter : -1;
      if (active) active[s] = slots[s] ? h[s].active : 0;
      if (slots[s]) slots[s]->cur = h[s].cur;
    }
    return 0;
  }

  int layout(int slot, float* out) {
    if (slot < 0 || slot >= max_jobs || !slots[slot]) return fail("invalid batch slot");
    if (status(nullptr, nullptr)) return 1;
    return slots[slot]->layout(out);
  }
};

}  // namespace fe

using fe::fail;

struct fastembedr_hip_tsne_plan {
  fe::TsnePlan impl;
};

extern "C" {

int fastembedr_hip_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

const char* fastembedr_hip_last_error(void) { return fe::last_error().c_str(); }

int fastembedr_hip_opentsne_fft_grid_size(int n) { return fe::resolve_grid_size(n, 0); }

int fastembedr_hip_opentsne_fft_from_knn(const int* indices, const double* distances,
                                         const float* init, int has_init, int n, int k,
                                         int n_components, float perplexity,
                                         int early_exaggeration_iter, int n_iter,
                                         float early_exaggeration, float exaggeration,
                                         float learning_rate, int learning_rate_auto,
                                         float initial_momentum, float final_momentum,
                                         float min_gain, float max_step_norm, unsigned int seed,
                                         int index_offset, int grid_size, float* out) {
  return fe::tsne_one_shot<double>(indices, distances, false, init, has_init, n, k, n_components,
                                   perplexity, early_exaggeration_iter, n_iter,
                                   early_exaggeration, exaggeration, learning_rate,
                                   learning_rate_auto, initial_momentum, final_momentum, min_gain,
                                   max_step_norm, seed, index_offset, grid_size, out);
}

int fastembedr_hip_opentsne_fft_from_knn_float(const int* indices, const float* distances,
                                               const float* init, int has_init, int n, int k,
                                               int n_components, float perplexity,
                                               int early_exaggeration_iter, int n_iter,
                                               float early_exaggeration, float exaggeration,
                                               float learning_rate, int learning_rate_auto,
                                               float initial_momentum, float final_momentum,
                                               float min_gain, float max_step_norm,
                                               unsigned int seed, int index_offset,
                                               int grid_size, float* out) {
  return fe::tsne_one_shot<float>(indices, distances, false, init, has_init, n, k, n_components,
                                  perplexity, early_exaggeration_iter, n_iter, early_exaggeration,
                                  exaggeration, learning_rate, learning_rate_auto,
                                  initial_momentum, final_momentum, min_gain, max_step_norm, seed,
                                  index_offset, grid_size, out);
}

int fastembedr_hip_opentsne_fft_from_device_knn_float(
    const int* indices, const float* distances, const float* init, int has_init, int n, int k,
    int n_components, float perplexity, int early_exaggeration_iter, int n_iter,
    float early_exaggeration, float exaggeration, float learning_rate, int learning_rate_auto,
    float initial_momentum, float final_momentum, float min_gain, float max_step_norm,
    unsigned int seed, int index_offset, int grid_size, float* out) {
  return fe::tsne_one_shot<float>(indices, distances, true, init, has_init, n, k, n_components,
                                  perplexity, early_exaggeration_iter, n_iter, early_exaggeration,
                                  exaggeration, learning_rate, learning_rate_auto,
                                  initial_momentum, final_momentum, min_gain, max_step_norm, seed,
                                  index_offset, grid_size, out);
}

int fastembedr_hip_tsne_gradient_diagnostic(const int* indices, const float* distances, int n,
                                            int k, int index_offset, float perplexity,
                                            float exaggeration, int grid_size,
                                            const float* layout, float* attractive,
                                            float* repulsive, float* gradient, double* sum_q,
                                            int* row_ptr, int* col, float* val,
                                            long long* nnz) {
  fe::clear_error();
  if (fe::validate_tsne(indices, distances, n, k, 2, perplexity, 0, 1, 1.0f, 1.0f, 1.0f, 1, 0.0f,
                        0.0f, 1.0f, index_offset))
    return 1;
  if (!layout) return fail("`layout` must have one row per KNN row and exactly two columns.");
  if (!(exaggeration > 0.0f) || !std::isfinite(exaggeration))
    return fail("`exaggeration` must be positive and finite.");
  if (grid_size < 16 || grid_size > 1024 || (grid_size & (grid_size - 1)) != 0)
    return fail("`grid_size` must be a power of two between 16 and 1024.");
  fe::TsnePlan p;
  if (p.create<float>(indices, distances, false, n, k, index_offset, perplexity, grid_size, layout,
                      0u, false))
    return 1;
  if (p.enqueue_diag(fe::StepParams{exaggeration, 0.f, 0.f, 0.01f, 0.f})) return 1;
  hipStream_t s = p.stream.get();
  std::vector<float> a(2 * static_cast<size_t>(n)), r(2 * static_cast<size_t>(n));
  FE_CHECK(hipMemcpyAsync(a.data(), p.diag, a.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(r.data(), p.diag + n, r.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  fe::TsneStats hst;
  FE_CHECK(hipMemcpyAsync(&hst, p.st, sizeof(hst), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  if (attractive) std::memcpy(attractive, a.data(), a.size() * sizeof(float));
  if (repulsive) std::memcpy(repulsive, r.data(), r.size() * sizeof(float));
  if (gradient)
    for (size_t t = 0; t < a.size(); ++t) gradient[t] = a[t] + r[t];
  if (sum_q) *sum_q = std::fmax(hst.sum_q_total - static_cast<double>(n), static_cast<double>(FLT_MIN));
  if (nnz) *nnz = p.csr.nnz;
  if (row_ptr)
    FE_CHECK(hipMemcpy(row_ptr, p.csr.row_ptr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
  if (col) FE_CHECK(hipMemcpy(col, p.csr.col, p.csr.nnz * sizeof(int), hipMemcpyDeviceToHost));
  if (val) FE_CHECK(hipMemcpy(val, p.csr.val, p.csr.nnz * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fastembedr_hip_opentsne_force_diagnostic(
    const int* indices, const float* distances, int n, int k, int index_offset, float perplexity,
    float exaggeration, int grid_size, const float* layout, float* attractive,
    float* repulsive_exact, float* gradient_exact, float* repulsive_fft, float* gradient_fft,
    double* sum_q, double* kl, int* row_ptr, int* col, float* val, long long* nnz) {
  fe::clear_error();
  if (fe::validate_tsne(indices, distances, n, k, 2, perplexity, 0, 1, 1.0f, 1.0f, 1.0f, 1, 0.0f,
                        0.0f, 1.0f, index_offset))
    return 1;
  if (!layout) return fail("`layout` must have one row per KNN row and exactly two columns.");
  if (!(perplexity <= static_cast<float>(k)))
    return fail("`perplexity` must be positive and no larger than the KNN width.");
  if (!(exaggeration > 0.0f) || !std::isfinite(exaggeration))
    return fail("`exaggeration` must be positive and finite.");
  if (grid_size < 16 || grid_size > 1024 || (grid_size & (grid_size - 1)) != 0)
    return fail("`grid_size` must be a power of two between 16 and 1024.");
  for (size_t t = 0; t < 2 * static_cast<size_t>(n); ++t)
    if (!std::isfinite(layout[t])) return fail("`layout` must contain only finite coordinates.");
  fe::TsnePlan p;
  if (p.create<float>(indices, distances, false, n, k, index_offset, perplexity, grid_size, layout,
                      0u, false))
    return 1;
  hipStream_t s = p.stream.get();
  const size_t m = 2 * static_cast<size_t>(n);
  std::vector<float> a(m), rf(m), re(m);
  // FFT route: attraction into diag[0, n), grid repulsion into diag[n, 2n)
  if (p.enqueue_diag(fe::StepParams{exaggeration, 0.f, 0.f, 0.01f, 0.f})) return 1;
  FE_CHECK(hipMemcpyAsync(a.data(), p.diag, m * sizeof(float), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(rf.data(), p.diag + n, m * sizeof(float), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  // exact route: pair repulsion into diag[n, 2n), then Z and KL
  if (p.ensure_exact() || p.launch_exact(true)) return 1;
  fe::update_exact_kernel<1><<<p.stat_blocks, fe::kThreads, 0, s>>>(
      p.y[p.cur], nullptr, nullptr, nullptr, p.diag, p.ex_part, p.ex_splits, p.ex_zpart,
      p.ex_splits * p.ex_bx, p.st, p.lpart, n, fe::StepParams{}, p.diag + n);
  FE_LAUNCH_CHECK("update_exact_kernel");
  FE_CHECK(hipMemcpyAsync(re.data(), p.diag + n, m * sizeof(float), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  double h_kl = 0.0, h_sq = 0.0;
  if (p.kl(&h_kl, &h_sq)) return 1;
  if (attractive) std::memcpy(attractive, a.data(), m * sizeof(float));
  if (repulsive_fft) std::memcpy(repulsive_fft, rf.data(), m * sizeof(float));
  if (repulsive_exact) std::memcpy(repulsive_exact, re.data(), m * sizeof(float));
  if (gradient_fft)
    for (size_t t = 0; t < m; ++t) gradient_fft[t] = a[t] + rf[t];
  if (gradient_exact)
    for (size_t t = 0; t < m; ++t) gradient_exact[t] = a[t] + re[t];
  if (sum_q) *sum_q = h_sq;
  if (kl) *kl = h_kl;
  if (nnz) *nnz = p.csr.nnz;
  if (row_ptr)
    FE_CHECK(hipMemcpy(row_ptr, p.csr.row_ptr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
  if (col) FE_CHECK(hipMemcpy(col, p.csr.col, p.csr.nnz * sizeof(int), hipMemcpyDeviceToHost));
  if (val) FE_CHECK(hipMemcpy(val, p.csr.val, p.csr.nnz * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fastembedr_hip_tsne_plan_kl(fastembedr_hip_tsne_plan* plan, double* kl, double* sum_q) {
  fe::clear_error();
  if (!plan) return fail("null plan");
  FE_CHECK(hipStreamSynchronize(plan->impl.stream.get()));
  return plan->impl.kl(kl, sum_q);
}

int fastembedr_hip_tsne_plan_create(const int* indices, const float* distances,
                                    int device_inputs, int n, int k, int index_offset,
                                    float perplexity, int grid_size, const float* init,
                                    unsigned int seed, fastembedr_hip_tsne_plan** plan) {
  return fastembedr_hip_tsne_plan_create_shaped(indices, distances, device_inputs, n, k,
                                                index_offset, perplexity, grid_size, 0, 0, init,
                                                seed, plan);
}

int fastembedr_hip_tsne_plan_create_shaped(const int* indices, const float* distances,
                                           int device_inputs, int n, int k, int index_offset,
                                           float perplexity, int grid_size, int attraction,
                                           int pair_workgroups, const float* init,
                                           unsigned int seed, fastembedr_hip_tsne_plan** plan) {
  fe::clear_error();
  if (!plan) return fail("null plan pointer");
  *plan = nullptr;
  if (fe::validate_tsne(indices, distances, n, k, 2, perplexity, 0, 1, 1.0f, 1.0f, 1.0f, 1, 0.0f,
                        0.0f, 1.0f, index_offset))
    return 1;
  if (pair_workgroups < 0) return fail("pair_workgroups must be non-negative");
  auto* p = new (std::nothrow) fastembedr_hip_tsne_plan;
  if (!p) return fail("out of host memory");
  p->impl.pair_wg_request = pair_workgroups;
  if (fe::attraction_request(attraction, &p->impl.attr_request) ||
      p->impl.create<float>(indices, distances, device_inputs != 0, n, k, index_offset, perplexity,
                            grid_size, init, seed, true)) {
    delete p;
    return 1;
  }
  *plan = p;
  return 0;
}

int fastembedr_hip_tsne_plan_run(fastembedr_hip_tsne_plan* plan, int n_iter, float exaggeration,
                                 float momentum, float learning_rate, float min_gain,
                                 float max_step_norm) {
  fe::clear_error();
  if (!plan) return fail("null plan");
  if (n_iter < 0) return fail("iteration counts must be non-negative.");
  if (!(exaggeration > 0.0f) || !(learning_rate > 0.0f) || momentum < 0.0f || !(min_gain > 0.0f))
    return fail("invalid t-SNE phase parameters");
  return plan->impl.run(n_iter,
                        fe::StepParams{exaggeration, learning_rate, momentum, min_gain, max_step_norm});
}

int fastembedr_hip_tsne_plan_prepare(fastembedr_hip_tsne_plan* plan, int n_iter,
                                     float exaggeration, float momentum, float learning_rate,
                                     float min_gain, float max_step_norm, long long after) {
  fe::clear_error();
  if (!plan) return fail("null plan");
  if (n_iter < 0 || after < 0) return fail("iteration counts must be non-negative.");
  if (!(exaggeration > 0.0f) || !(learning_rate > 0.0f) || momentum < 0.0f || !(min_gain > 0.0f))
    return fail("invalid t-SNE phase parameters");
  return plan->impl.prepare(
      n_iter, fe::StepParams{exaggeration, learning_rate, momentum, min_gain, max_step_norm}, after);
}

int fastembedr_hip_tsne_plan_sync(fastembedr_hip_tsne_plan* plan) {
  if (!plan) return fail("null plan");
  return fe::check(hipStreamSynchronize(plan->impl.stream.get()), "hipStreamSynchronize");
}

int fastembedr_hip_tsne_plan_layout(fastembedr_hip_tsne_plan* plan, float* out) {
  fe::clear_error();
  if (!plan || !out) return fail("null plan or output");
  return plan->impl.layout(out);
}

int fastembedr_hip_tsne_plan_info(fastembedr_hip_tsne_plan* plan, long long* info) {
  if (!plan || !info) return fail("null plan or output");
  const fe::TsnePlan& p = plan->impl;
  info[0] = p.n;
  info[1] = p.k;
  info[2] = p.exact ? -1 : p.g;
  info[3] = p.csr.nnz;
  info[4] = p.attr_kind == fe::kAttrRows ? p.tpr : 0;
  info[5] = p.device_bytes();
  info[6] = p.iters_run;
  info[7] = static_cast<long long>(p.setup_us);
  return 0;
}

int fastembedr_hip_tsne_plan_attraction_info(fastembedr_hip_tsne_plan* plan, long long* info) {
  if (!plan || !info) return fail("null plan or output");
  const fe::TsnePlan& p = plan->impl;
  info[0] = p.attr_kind;
  info[1] = p.attr_kind == fe::kAttrPairs ? p.pair_entries : p.seg_entries;
  info[2] = p.attr_kind == fe::kAttrPairs ? p.rest_entries : 0;
  info[3] = p.attr_kind == fe::kAttrPairs ? p.pair_ranges : p.groups;
  info[4] = p.attr_kind == fe::kAttrPairs ? p.pair_nchunks : 0;
  info[5] = p.attr_kind == fe::kAttrPairs ? p.pair_nslots : 0;
  return 0;
}

int fastembedr_hip_tsne_plan_affinities(fastembedr_hip_tsne_plan* plan, int* row_ptr, int* col,
                                        float* val) {
  fe::clear_error();
  if (!plan) return fail("null plan");
  const fe::TsnePlan& p = plan->impl;
  FE_CHECK(hipStreamSynchronize(p.stream.get()));
  if (row_ptr)
    FE_CHECK(hipMemcpy(row_ptr, p.csr.row_ptr, (p.n + 1) * sizeof(int), hipMemcpyDeviceToHost));
  if (col) FE_CHECK(hipMemcpy(col, p.csr.col, p.csr.nnz * sizeof(int), hipMemcpyDeviceToHost));
  if (val) FE_CHECK(hipMemcpy(val, p.csr.val, p.csr.nnz * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fastembedr_hip_tsne_plan_profile(fastembedr_hip_tsne_plan* plan, int n_iter,
                                     float exaggeration, float momentum, float learning_rate,
                                     float min_gain, float max_step_norm, double* kernel_ms) {
  fe::clear_error();
  if (!plan || !kernel_ms || n_iter < 1) return fail("invalid profile arguments");
  return plan->impl.profile(
      n_iter, fe::StepParams{exaggeration, learning_rate, momentum, min_gain, max_step_norm},
      kernel_ms);
}

void fastembedr_hip_tsne_plan_destroy(fastembedr_hip_tsne_plan* plan) {
  if (plan) {
    (void)hipStreamSynchronize(plan->impl.stream.get());
    delete plan;
  }
}

}  // extern "C"

struct fastembedr_hip_tsne_batch {
  fe::TsneBatch impl;
};

hipStream_t fe::tsne_batch_stream(const fastembedr_hip_tsne_batch* batch) {
  return batch ? batch->impl.stream.get() : nullptr;
}

void fe::tsne_batch_set_budget(fastembedr_hip_tsne_batch* batch, long long budget) {
  if (batch) batch->impl.budget = budget;
}

long long fe::tsne_batch_used(const fastembedr_hip_tsne_batch* batch) {
  return batch ? batch->impl.used : 0;
}

extern "C" {

int fastembedr_hip_tsne_batch_create(int max_jobs, long long hbm_budget_bytes,
                                     fastembedr_hip_tsne_batch** batch) {
  fe::clear_error();
  if (!batch) return fail("null batch pointer");
  *batch = nullptr;
  if (max_jobs < 1 || max_jobs > 4096) return fail("max_jobs must lie in [1, 4096]");
  auto* b = new (std::nothrow) fastembedr_hip_tsne_batch;
  if (!b) return fail("out of host memory");
  if (b->impl.init(max_jobs, hbm_budget_bytes)) {
    delete b;
    return 1;
  }
  *batch = b;
  return 0;
}

long long fastembedr_hip_tsne_job_bytes(int n, int k, int grid_size, int setup_peak) {
  const int g = fe::resolve_grid_size(n, grid_size);
  return fe::job_resident_bytes(n, k, g) + (setup_peak ? fe::job_setup_bytes(n, k) : 0);
}

int fastembedr_hip_tsne_batch_add(fastembedr_hip_tsne_batch* batch,
                                  const fastembedr_hip_tsne_job* job, int device_inputs,
                                  int* slot) {
  fe::clear_error();
  if (!batch || !job || !slot) return fail("null batch argument");
  return batch->impl.add(*job, device_inputs != 0, slot);
}

int fastembedr_hip_tsne_batch_step(fastembedr_hip_tsne_batch* batch, int n_iter) {
  fe::clear_error();
  if (!batch || n_iter < 0) return fail("invalid batch step");
  return batch->impl.step(n_iter);
}

int fastembedr_hip_tsne_batch_status(fastembedr_hip_tsne_batch* batch, int* iterations,
                                     int* active) {
  fe::clear_error();
  if (!batch) return fail("null batch");
  return batch->impl.status(iterations, active);
}

int fastembedr_hip_tsne_batch_layout(fastembedr_hip_tsne_batch* batch, int slot, float* out) {
  fe::clear_error();
  if (!batch || !out) return fail("null batch argument");
  return batch->impl.layout(slot, out);
}

int fastembedr_hip_tsne_batch_remove(fastembedr_hip_tsne_batch* batch, int slot) {
  fe::clear_error();
  if (!batch) return fail("null batch");
  return batch->impl.remove(slot);
}

void fastembedr_hip_tsne_batch_destroy(fastembedr_hip_tsne_batch* batch) {
  if (batch) {
    (void)hipStreamSynchronize(batch->impl.stream.get());
    delete batch;
  }
}

}  // extern "C"
