// Attraction with every symmetric pair of P streamed once (gfx950).
//
// The attractive force of the reference (add_sparse_attractive_gradient_f,
// tsne_neighbors.cpp:1061-1088; CUDA twin embedding_cuda_kernels.cpp:
// 2788-2817) is attr_i = exaggeration * sum_j P_ij q_ij (y_i - y_j).  P is
// symmetric, so the pair {i, j} contributes t = P_ij q_ij (y_i - y_j) to i
// and -t to j: it is enough to read each pair once (the symmetric CSR stores
// it twice).
//
// Points are cut into blocks of kPairBlock consecutive indices.  The pairs
// between blocks I <= J form tile (I, J) (on the diagonal, the pairs with
// i < j).  A tile with many pairs is "staged": its pairs are stored once, in
// runs of one row padded to whole lanes of kLaneEntries, and a workgroup
// stages both blocks of the layout in LDS, so every gather reads LDS.  Pairs
// of the other tiles go to the "rest" stream, stored in both directions with
// global columns, and add to their row only (gathers from the layout).
//
// Sums are integers: every pair's contribution is scaled by a power of two,
// truncated toward zero and packed as the 64-bit integer y * 2^32 + x; its
// column adds it with one LDS 64-bit atomic into one accumulator per staged
// point, and its row adds a lane's 8 of them at once; a work item writes its
// accumulators to its partial slots, and the update adds a point's slots.
// A sum of packed values is the packed sum Y * 2^32 + X exactly, and while
// |X| < 2^31 and |Y| < 2^30 the low 32 bits read as a signed integer are X
// and the rest is Y.  The result is therefore
// the same bits whatever the order of the work, the waves or the runs in the
// streams.  Every partial sum is at most the largest row sum of P times
// max |q (y_i - y_j)| <= min(1/2, diagonal of the layout's bounding box); the
// scale, a power of two set per launch from the current bounds, puts that
// bound below 2^30, so the sums stay inside 32 bits and a small early layout
// keeps as many digits as a spread one.
#pragma once

#include "tsne_device.hpp"

namespace fe {
namespace {

constexpr int kPairBlock = 4096;  // points per block
// 16 waves (four per SIMD at 80 VGPRs): the stream is bound by the bytes a
// CU keeps in flight, and at 1M points 16 waves take 145 us alone on every
// CU against 160 with 8 (4015 against 3992 it/s in the overlapped iteration)
constexpr int kPairWaves = 16;
static_assert(kPairBlock % 64 == 0 && kPairBlock <= 65536,
              "blocks of whole waves' worth of points, with 16-bit offsets");
constexpr int kPairThreads = 64 * kPairWaves;
// staged rows, staged columns, their packed accumulators: 128 KB
constexpr size_t kPairLds = 2 * kPairBlock * sizeof(float2) + 2 * kPairBlock * sizeof(long long);
// Regions are cut into chunks of at most 524288 entries, and again where a
// workgroup's range ends: every chunk starts its stream reads cold, so few
// and long chunks pay (1M points: 64K-entry chunks ran the overlapped
// iteration at 3890 it/s, 128K at 3965, whole regions cut at the range
// boundaries at 3996)
constexpr int kPairChunkEntries = 1024 * kLaneStep;
constexpr int kPairDenseMin = 8192;              // pairs that pay for staging an off-diagonal tile
constexpr int kPairMaxBlocks = 4096;             // tile table nb^2 ints (n <= 16.7M)
// the streams are read two lane-steps ahead by every wave:
// padding past their end for those reads
constexpr int kPairPad = 5 * kLaneStep;

// A chunk is a piece of one tile's region of the staged stream, or of one row
// block's region of the rest stream.  Workgroups take contiguous ranges of
// the chunk list (tiles in (I, J) order) and keep their accumulators across
// chunks of one row block (row side) and one tile (column side): `flush`
// marks the chunks after which a side is written to its slot and cleared.
struct PairChunk {
  int I, J;          // row block; column block (staged chunks)
  int e0, e1;        // entries [e0, e1) of its stream (multiples of kLaneEntries)
  int slot_i, slot_j;
  int rest;          // 1: the rest stream (global columns, row side only)
  int flush;         // bit 0: write and clear the row side after it; bit 1: the column side
};

struct PairTables {
  const unsigned short* col16;  // staged stream: column - J * kPairBlock
  const float* val;             // staged stream: P_ij (0 in padding)
  const unsigned short* row16;  // staged stream: row - I * kPairBlock, per kLaneEntries entries
  const int* rcol;              // rest stream: columns
  const float* rval;
  const unsigned short* rrow16;
  const PairChunk* chunk;
  const int* range_first;       // workgroup w: chunks range_first[w] .. range_first[w + 1]
  unsigned long long* part;     // partial slot s: part[s * kPairBlock + local point], packed
  const int* slot_ptr;          // block b: slots slot_list[slot_ptr[b] .. slot_ptr[b + 1])
  const int* slot_list;
  double row_bound;             // the largest row sum of P, with 1% headroom
};

// Exponent of the fixed-point scale for the current layout: 2^e times
// row_bound * min(1/2, diagonal of the bounds) stays below 2^30.
__device__ __forceinline__ int pair_scale_exp(const PairTables& t, const TsneStats* st) {
  const double sx = static_cast<double>(st->max_x) - st->min_x;
  const double sy = static_cast<double>(st->max_y) - st->min_y;
  double span = sqrt(sx * sx + sy * sy);
  if (!(span < 0.5)) span = 0.5;  // NaN too
  const double b = fmax(t.row_bound * span, 1e-300);
  return min(120, max(-120, 29 - ilogb(b)));
}

// (x, y), already scaled, truncated toward zero and packed as y * 2^32 + x:
// one conversion instruction a component, and the pack of -(x, y) is the
// negated pack exactly
__device__ __forceinline__ unsigned long long pack_fixed_rz(float x, float y) {
  const long long ix = static_cast<int>(x), iy = static_cast<int>(y);
  return static_cast<unsigned long long>(iy * 4294967296ll + ix);
}

// The LDS holds, per staged block, the points' positions followed by their
// accumulators, kPairBlock float2 (32 KB) apart: the accumulator of a staged
// point sits at a fixed offset from its position, so one address serves the
// gather and the atomic.
__device__ __forceinline__ unsigned long long* pair_acc(const float2* p) {
  return reinterpret_cast<unsigned long long*>(const_cast<float2*>(p + kPairBlock));
}

__device__ __forceinline__ void unpack_fixed(unsigned long long v, long long* x, long long* y) {
  const long long lo = static_cast<int>(static_cast<unsigned int>(v));
  *x = lo;
  *y = (static_cast<long long>(v) - lo) >> 32;
}

__device__ __forceinline__ int pad_lane(int len) {
  return (len + kLaneEntries - 1) / kLaneEntries * kLaneEntries;
}

// a block of the layout into LDS, centred (fl(y - mean) as every reader);
// every load of a thread is issued before its first LDS write
__device__ __forceinline__ void pair_stage(const float2* __restrict__ y, int lo, int m, float mx,
                                           float my, float2* s) {
  constexpr int L = (kPairBlock + kPairThreads - 1) / kPairThreads;
  float2 r[L];
#pragma unroll
  for (int u = 0; u < L; ++u) {
    const int i = static_cast<int>(threadIdx.x) + u * kPairThreads;
    if (i < m) r[u] = y[lo + i];
  }
#pragma unroll
  for (int u = 0; u < L; ++u) {
    const int i = static_cast<int>(threadIdx.x) + u * kPairThreads;
    if (i < m) s[i] = make_float2(r[u].x - mx, r[u].y - my);
  }
}

// Within each lane-step of kLaneStep entries, the 32-bit arrays (values, the
// rest stream's columns) hold entries 0..3 of every lane first (lane-major,
// 16 bytes per lane) and entries 4..7 after them, so each of a wave's two
// 16-byte loads reads 1 KB of whole cache lines.  Stream position p (lane
// l = p % 512 / 8, entry u = p % 8 of step p / 512) is stored at:
__host__ __device__ __forceinline__ int split_index(int p) {
  return (p & ~(kLaneStep - 1)) + ((p & 7) >> 2) * (kLaneStep / 2) + ((p & (kLaneStep - 1)) >> 3) * 4 +
         (p & 3);
}

// One lane-step's loads: 8 columns, 8 values and the row of the lane.
struct PairLane {
  int j[kLaneEntries];
  float v[kLaneEntries];
  int row;
};

template <bool REST>
__device__ __forceinline__ PairLane pair_lane_load(const PairTables& t, int eb) {
  PairLane L;
  const int lo = split_index(eb), hi = lo + kLaneStep / 2;  // entries 0..3, 4..7
  if constexpr (REST) {
    const int4 a = stream_load(reinterpret_cast<const int4*>(t.rcol + lo));
    const int4 b = stream_load(reinterpret_cast<const int4*>(t.rcol + hi));
    L.j[0] = a.x; L.j[1] = a.y; L.j[2] = a.z; L.j[3] = a.w;
    L.j[4] = b.x; L.j[5] = b.y; L.j[6] = b.z; L.j[7] = b.w;
    const float4 va = stream_load(reinterpret_cast<const float4*>(t.rval + lo));
    const float4 vb = stream_load(reinterpret_cast<const float4*>(t.rval + hi));
    L.v[0] = va.x; L.v[1] = va.y; L.v[2] = va.z; L.v[3] = va.w;
    L.v[4] = vb.x; L.v[5] = vb.y; L.v[6] = vb.z; L.v[7] = vb.w;
    L.row = t.rrow16[eb / kLaneEntries];
  } else {
    const int4 a = stream_load(reinterpret_cast<const int4*>(t.col16 + eb));
    const unsigned int w[4] = {static_cast<unsigned int>(a.x), static_cast<unsigned int>(a.y),
                               static_cast<unsigned int>(a.z), static_cast<unsigned int>(a.w)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      L.j[2 * q] = static_cast<int>(w[q] & 0xffffu);
      L.j[2 * q + 1] = static_cast<int>(w[q] >> 16);
    }
    const float4 va = stream_load(reinterpret_cast<const float4*>(t.val + lo));
    const float4 vb = stream_load(reinterpret_cast<const float4*>(t.val + hi));
    L.v[0] = va.x; L.v[1] = va.y; L.v[2] = va.z; L.v[3] = va.w;
    L.v[4] = vb.x; L.v[5] = vb.y; L.v[6] = vb.z; L.v[7] = vb.w;
    L.row = t.row16[eb / kLaneEntries];
  }
  return L;
}

// Entries [e0, e1) of one wave (multiples of kLaneEntries): lane l of a step
// takes entries base + 8l .. +7, one row's.  Every column and row index read
// here lies inside its block or the layout (the streams are filled whole and
// padded by kPairPad entries past their end), so lanes outside [e0, e1) load
// safely and only their sums are dropped.
// One lane-step: the gathers of the lane's 8 columns and its row from the
// staged blocks (or the layout for the rest stream), the terms, the column
// side into its accumulators and the lane's row sum into the row's.
template <bool REST>
__device__ __forceinline__ void pair_step(const float2* __restrict__ y, const PairLane& L,
                                          bool valid, const float2* yI, const float2* yJ,
                                          float mx, float my, float nscale) {
  float2 pj[kLaneEntries];
#pragma unroll
  for (int u = 0; u < kLaneEntries; ++u) pj[u] = REST ? y[L.j[u]] : yJ[L.j[u]];
  const float2 pi = yI[L.row];
  if (!valid) return;
  // nt = the pack of -t (t the entry's term for its row, scaled): the column
  // adds it, the row subtracts the lane's sum of them once
  unsigned long long nrow = 0ull;
#pragma unroll
  for (int u = 0; u < kLaneEntries; ++u) {
    const float2 q2 = REST ? make_float2(pj[u].x - mx, pj[u].y - my) : pj[u];
    const float dx = pi.x - q2.x, dy = pi.y - q2.y;
    const float q = __builtin_amdgcn_rcpf(1.0f + dx * dx + dy * dy);
    const float c = (L.v[u] * nscale) * q;
    const unsigned long long nt = pack_fixed_rz(c * dx, c * dy);
    nrow += nt;
    if constexpr (!REST) {
      // padding entries (value 0) add 0: cheaper than a branch per entry
      atomicAdd(pair_acc(yJ + L.j[u]), nt);
    }
  }
  atomicAdd(pair_acc(yI + L.row), 0ull - nrow);
}

// Entries [e0, e1) of one wave (multiples of kLaneEntries): lane l of a step
// takes entries base + 8l .. +7, one row's.  Two buffers alternate (the loop
// is unrolled by two): a step's buffer is reloaded with the step two ahead as
// soon as the step is done, with no register copies that would wait on loads
// in flight.  Every column and row index read here lies inside its block or
// the layout (the streams are filled whole and padded by kPairPad entries
// past their end), so lanes outside [e0, e1) load safely and only their sums
// are dropped.
template <bool REST>
__device__ __forceinline__ void pair_range(const float2* __restrict__ y, const PairTables& t,
                                           const float2* yI, const float2* yJ, int e0, int e1,
                                           float mx, float my, float nscale) {
  const int lane = threadIdx.x & 63;
  int base = e0 & ~(kLaneStep - 1);
  PairLane a = pair_lane_load<REST>(t, base + kLaneEntries * lane);
  PairLane b = pair_lane_load<REST>(t, base + kLaneStep + kLaneEntries * lane);
  for (; base < e1; base += 2 * kLaneStep) {
    const int ea = base + kLaneEntries * lane, eb = ea + kLaneStep;
    pair_step<REST>(y, a, ea >= e0 && ea < e1, yI, yJ, mx, my, nscale);
    a = pair_lane_load<REST>(t, ea + 2 * kLaneStep);
    if (base + kLaneStep >= e1) break;
    pair_step<REST>(y, b, eb >= e0 && eb < e1, yI, yJ, mx, my, nscale);
    b = pair_lane_load<REST>(t, eb + 2 * kLaneStep);
  }
}

// One workgroup's range of chunks: stage a row block when it changes and a
// column block when the tile changes, stream each chunk, and write a side's
// accumulators to their slot (then clear them) where the chunk says so.
__device__ __forceinline__ void attract_pairs_range(const float2* __restrict__ y,
                                                    const TsneStats* __restrict__ st,
                                                    const PairTables& t, int n, int w,
                                                    float2* lds) {
  const float mx = st->mean_xf, my = st->mean_yf;
  const float scale = ldexpf(1.0f, pair_scale_exp(t, st));
  // [row block positions | their accumulators | column block positions | theirs]
  float2* yI = lds;
  float2* yJ = lds + 2 * kPairBlock;
  unsigned long long* accI = pair_acc(yI);
  unsigned long long* accJ = pair_acc(yJ);
  for (int q = static_cast<int>(threadIdx.x); q < kPairBlock; q += kPairThreads) {
    accI[q] = 0ull;
    accJ[q] = 0ull;
  }
  const int c0 = t.range_first[w], c1 = t.range_first[w + 1];
  int curI = -1, curJ = -1;
  const int wave = static_cast<int>(threadIdx.x >> 6);
  for (int c = c0; c < c1; ++c) {
    const PairChunk ch = t.chunk[c];
    const bool rest = ch.rest != 0;
    const bool diag = !rest && ch.J == ch.I;
    const int ni = min(kPairBlock, n - ch.I * kPairBlock);
    const int nj = rest ? 0 : min(kPairBlock, n - ch.J * kPairBlock);
    if (ch.I != curI) pair_stage(y, ch.I * kPairBlock, ni, mx, my, yI);
    if (!rest && !diag && (ch.I != curI || ch.J != curJ))
      pair_stage(y, ch.J * kPairBlock, nj, mx, my, yJ);
    curI = ch.I;
    curJ = rest ? -1 : ch.J;
    __syncthreads();
    const int s0 = ch.e0 / kLaneStep, steps = (ch.e1 - 1) / kLaneStep - s0 + 1;
    const int e0 = max(ch.e0, (s0 + wave * steps / kPairWaves) * kLaneStep);
    const int e1 = min(ch.e1, (s0 + (wave + 1) * steps / kPairWaves) * kLaneStep);
    if (e0 < e1) {
      if (rest)
        pair_range<true>(y, t, yI, nullptr, e0, e1, mx, my, -scale);
      else
        pair_range<false>(y, t, yI, diag ? yI : yJ, e0, e1, mx, my, -scale);
    }
    __syncthreads();
    if (ch.flush & 1) {
      unsigned long long* pi = t.part + static_cast<size_t>(ch.slot_i) * kPairBlock;
      for (int q = static_cast<int>(threadIdx.x); q < ni; q += kPairThreads) {
        pi[q] = accI[q];
        accI[q] = 0ull;
      }
    }
    if (ch.flush & 2) {
      unsigned long long* pj = t.part + static_cast<size_t>(ch.slot_j) * kPairBlock;
      for (int q = static_cast<int>(threadIdx.x); q < nj; q += kPairThreads) {
        pj[q] = accJ[q];
        accJ[q] = 0ull;
      }
    }
    // the next chunk's staging writes the blocks only after every wave is done
    // (the barrier above) and its streaming reads them after the next barrier
  }
}

__global__ __launch_bounds__(kPairThreads) void attract_pairs(const float2* __restrict__ y_in,
                                                              PairTables t,
                                                              const TsneStats* __restrict__ st,
                                                              int n) {
  extern __shared__ float2 lds[];
  attract_pairs_range(y_in, st, t, n,
                      xcd_remap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)), lds);
}

// attr_i = exaggeration * (sum of the point's slots) / scale; the slots are
// added as 64-bit integers, so their order does not matter.
__device__ __forceinline__ void pairs_reduce_body(const PairTables& t,
                                                  const TsneStats* __restrict__ st,
                                                  float exaggeration, float2* __restrict__ attr,
                                                  int n, int blk, int nblk) {
  const double inv_scale = ldexp(1.0, -pair_scale_exp(t, st));
  int beg, end;
  point_range(n, xcd_remap(blk, nblk), nblk, &beg, &end);
  for (int p = beg + static_cast<int>(threadIdx.x); p < end; p += kThreads) {
    const int b = p / kPairBlock, l = p - b * kPairBlock;
    unsigned long long acc = 0ull;
    for (int s = t.slot_ptr[b]; s < t.slot_ptr[b + 1]; ++s)
      acc += t.part[static_cast<size_t>(t.slot_list[s]) * kPairBlock + l];
    long long sx, sy;
    unpack_fixed(acc, &sx, &sy);
    attr[p] = make_float2(exaggeration * static_cast<float>(static_cast<double>(sx) * inv_scale),
                          exaggeration * static_cast<float>(static_cast<double>(sy) * inv_scale));
  }
}

// The update reading the attraction straight from the partial slots (the
// same sum and rounding as pairs_reduce_body, so the same bits), which saves
// the reduce launch and the force array's write and read.
struct AttrPairs {
  PairTables t;
  double inv_scale;
  float exaggeration;
  __device__ __forceinline__ float2 operator()(int p) const {
    const int b = p / kPairBlock, l = p - b * kPairBlock;
    const int s0 = t.slot_ptr[b], s1 = t.slot_ptr[b + 1];
    // four slots' loads in flight at a time (integer sums: any order)
    unsigned long long acc = 0ull;
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
      const int a0 = t.slot_list[s], a1 = t.slot_list[s + 1];
      const int a2 = t.slot_list[s + 2], a3 = t.slot_list[s + 3];
      const unsigned long long v0 = t.part[static_cast<size_t>(a0) * kPairBlock + l];
      const unsigned long long v1 = t.part[static_cast<size_t>(a1) * kPairBlock + l];
      const unsigned long long v2 = t.part[static_cast<size_t>(a2) * kPairBlock + l];
      const unsigned long long v3 = t.part[static_cast<size_t>(a3) * kPairBlock + l];
      acc += (v0 + v1) + (v2 + v3);
    }
    for (; s < s1; ++s) acc += t.part[static_cast<size_t>(t.slot_list[s]) * kPairBlock + l];
    long long sx, sy;
    unpack_fixed(acc, &sx, &sy);
    return make_float2(exaggeration * static_cast<float>(static_cast<double>(sx) * inv_scale),
                       exaggeration * static_cast<float>(static_cast<double>(sy) * inv_scale));
  }
};

__device__ __forceinline__ AttrPairs attr_pairs(const PairTables& t, const TsneStats* st,
                                                float exaggeration) {
  return AttrPairs{t, ldexp(1.0, -pair_scale_exp(t, st)), exaggeration};
}

__global__ __launch_bounds__(kThreads) void update_pairs_kernel(
    const float2* __restrict__ y_in, float2* __restrict__ y_out, float2* __restrict__ gains,
    float2* __restrict__ upd, PairTables t, const float4* __restrict__ field, TsneStats* st,
    LayoutPartial* partials, int n, int g, StepParams sp) {
  update_body<0>(y_in, y_out, gains, upd, attr_pairs(t, st, sp.exaggeration), field, st,
                 partials, n, g, sp, nullptr, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kThreads) void pairs_reduce(PairTables t,
                                                         const TsneStats* __restrict__ st,
                                                         float exaggeration,
                                                         float2* __restrict__ attr, int n) {
  pairs_reduce_body(t, st, exaggeration, attr, n, blockIdx.x, gridDim.x);
}

// ---- set-up of the streams ----

// first entry of row i's sorted columns with column > i
__device__ __forceinline__ int first_above(const int* __restrict__ col, int e0, int e1, int i) {
  int lo = e0, hi = e1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] <= i) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Per row: its pairs with j > i counted per tile (and padded to whole lanes),
// and the largest row sum of P (positive floats order as their bits).
__global__ __launch_bounds__(kThreads) void pair_count_kernel(
    const int* __restrict__ row_ptr, const int* __restrict__ col, const float* __restrict__ val,
    int n, int nb, unsigned* __restrict__ tile_cnt, unsigned* __restrict__ tile_pad,
    unsigned* __restrict__ max_row_bits) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int I = i / kPairBlock, e0 = row_ptr[i], e1 = row_ptr[i + 1];
  float rs = 0.0f;
  for (int e = e0; e < e1; ++e) rs += val[e];
  atomicMax(max_row_bits, __float_as_uint(rs));
  for (int e = first_above(col, e0, e1, i); e < e1;) {
    const int J = col[e] / kPairBlock;
    int f = e + 1;
    while (f < e1 && col[f] / kPairBlock == J) ++f;
    const size_t tile = static_cast<size_t>(I) * nb + J;
    atomicAdd(&tile_cnt[tile], static_cast<unsigned>(f - e));
    atomicAdd(&tile_pad[tile], static_cast<unsigned>(pad_lane(f - e)));
    e = f;
  }
}

// Per row: its entries (both directions) in tiles that are not staged.
__global__ __launch_bounds__(kThreads) void pair_rest_count_kernel(
    const int* __restrict__ row_ptr, const int* __restrict__ col, int n, int nb,
    const int* __restrict__ tile_id, int* __restrict__ rest_len, unsigned* __restrict__ rest_pad) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int I = i / kPairBlock;
  int cnt = 0;
  for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
    const int J = col[e] / kPairBlock;
    const int a = min(I, J), b = max(I, J);
    if (tile_id[static_cast<size_t>(a) * nb + b] < 0) ++cnt;
  }
  rest_len[i] = cnt;
  if (cnt) atomicAdd(&rest_pad[I], static_cast<unsigned>(pad_lane(cnt)));
}

// Per row: its runs of staged pairs (j > i, one run per tile, placed by the
// tile's cursor) and its rest entries (placed by the row block's cursor).
// Padding entries carry value 0; a staged run's padding repeats the run's
// last column, so its column atomics (which add 0) do not pile the padding
// lanes of a wave onto one accumulator.  Run placement depends
// on timing; no result does (integer sums).
__global__ __launch_bounds__(kThreads) void pair_fill_kernel(
    const int* __restrict__ row_ptr, const int* __restrict__ col, const float* __restrict__ val,
    int n, int nb, const int* __restrict__ tile_id, const int* __restrict__ rest_len,
    unsigned* __restrict__ tile_cursor, unsigned* __restrict__ rest_cursor,
    unsigned short* __restrict__ col16, float* __restrict__ pval, unsigned short* __restrict__ row16,
    int* __restrict__ rcol, float* __restrict__ rval, unsigned short* __restrict__ rrow16) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int I = i / kPairBlock, il = i - I * kPairBlock;
  const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
  for (int e = first_above(col, e0, e1, i); e < e1;) {
    const int J = col[e] / kPairBlock;
    int f = e + 1;
    while (f < e1 && col[f] / kPairBlock == J) ++f;
    const int t = tile_id[static_cast<size_t>(I) * nb + J];
    if (t >= 0) {
      const int len = f - e, pl = pad_lane(len);
      const unsigned pos = atomicAdd(&tile_cursor[t], static_cast<unsigned>(pl));
      for (int q = 0; q < pl; ++q) {
        col16[pos + q] = static_cast<unsigned short>(col[e + min(q, len - 1)] - J * kPairBlock);
        pval[split_index(static_cast<int>(pos) + q)] = q < len ? val[e + q] : 0.0f;
      }
      for (int r = 0; r < pl / kLaneEntries; ++r)
        row16[pos / kLaneEntries + r] = static_cast<unsigned short>(il);
    }
    e = f;
  }
  const int cnt = rest_len[i];
  if (cnt) {
    const int pl = pad_lane(cnt);
    const unsigned pos = atomicAdd(&rest_cursor[I], static_cast<unsigned>(pl));
    int w = 0;
    for (int e = e0; e < e1; ++e) {
      const int j = col[e], J = j / kPairBlock;
      const int a = min(I, J), b = max(I, J);
      if (tile_id[static_cast<size_t>(a) * nb + b] < 0) {
        rcol[split_index(static_cast<int>(pos) + w)] = j;
        rval[split_index(static_cast<int>(pos) + w)] = val[e];
        ++w;
      }
    }
    for (; w < pl; ++w) {
      rcol[split_index(static_cast<int>(pos) + w)] = i;
      rval[split_index(static_cast<int>(pos) + w)] = 0.0f;
    }
    for (int r = 0; r < pl / kLaneEntries; ++r)
      rrow16[pos / kLaneEntries + r] = static_cast<unsigned short>(il);
  }
}

}  // namespace
}  // namespace fe
