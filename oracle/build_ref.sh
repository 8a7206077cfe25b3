#!/bin/sh
# Builds oracle/_ref/libref.so: the reference's own C++ for the functions on
# the hot path that need nothing but the standard library, compiled from line
# ranges of the reference sources where they lie (piped to the compiler; no
# reference source is written anywhere) together with the harnesses
# ref_tsne_harness.cpp / ref_umap_harness.cpp, which expose them through
# extern "C" for the tests and the CPU baseline.
#
# TEST INFRASTRUCTURE ONLY.  The ranges (file:first-last):
#   src/tsne_neighbors.cpp
#     43-88      SparseProbabilitiesF, pair keys, mix_uint32 / xorshift32 /
#                uniform_index (the transform RNG)
#     113-284    ParallelExecutor, ParallelExecutorScope, parallel_for
#     293-302    env_positive_int
#     403-413    tsne_fft_grid_size
#     532-614    compute_row_probabilities_float
#     725-1365   FftPlanT ... evaluate_kl_f (FFT, grid gradient, sum_q,
#                attraction, exact pair gradient, zero mean, update, KL)
#   src/fast_knn_umap.cpp
#     59-89      ReusableBarrier, effective_cpu_threads
#     142-148    clip_value, clip4f
#     156-215    umap_pow, umap_powf_fast
#     437-497    find_ab_params
#     513-556    mix_seed, deterministic_unit, deterministic_vertex,
#                deterministic_vertex_cuda_style
#     604-701    the positive / negative epoch schedules, atomic_add_float,
#                TauPrng, make_tau_prng
#     2663-2679  add_delta_2d
#     3273-3325  the hogwild worker's per-epoch edge loop of
#                optimize_layout_csr (compiled inside the harness's worker)
# Flags: -O2, no -ffast-math / -march=native (docs/implementation.md:47-53).
# Without the reference sources (e.g. on a machine that only runs the
# tests) nothing is built and the tests that need _ref skip.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${FASTEMBEDR_REFERENCE:-/root/reference}
SRC=$REF/src
OUT=$HERE/_ref
if [ ! -f "$SRC/tsne_neighbors.cpp" ] || [ ! -f "$SRC/fast_knn_umap.cpp" ]; then
  echo "build_ref.sh: reference sources not found under $REF; oracle/_ref not built" >&2
  exit 0
fi
CXX=${CXX:-g++}
FLAGS="-O2 -fPIC -pthread -std=gnu++17 -w"
mkdir -p "$OUT"

# prints the lines first..last of $1, preceded by a #line directive so that
# diagnostics name the reference file and line
ranges() {
  file=$1
  shift
  for r in "$@"; do
    first=${r%-*}
    last=${r#*-}
    printf '#line %s "%s"\n' "$first" "$file"
    sed -n "${first},${last}p" "$file"
  done
}

{
  printf '#include "%s"\nnamespace ref_tsne {\nnamespace {\n' "$HERE/ref_prelude.hpp"
  ranges "$SRC/tsne_neighbors.cpp" 43-88 113-284 293-302 403-413 532-614 725-1365
  printf '}  // namespace\n}  // namespace ref_tsne\n#line 1 "%s"\n' "$HERE/ref_tsne_harness.cpp"
  cat "$HERE/ref_tsne_harness.cpp"
} | $CXX $FLAGS -x c++ -c - -o "$OUT/ref_tsne.o"

{
  printf '#include "%s"\nnamespace ref_umap {\nnamespace {\n' "$HERE/ref_prelude.hpp"
  ranges "$SRC/fast_knn_umap.cpp" 59-89 142-148 156-215 437-497 513-556 604-701 2663-2679
  printf '}  // namespace\n}  // namespace ref_umap\n'
  # the harness holds the worker loop; the reference's edge loop goes where
  # it says REF_HOGWILD_EDGE_LOOP
  sed -n '1,/REF_HOGWILD_EDGE_LOOP/p' "$HERE/ref_umap_harness.cpp" | sed '$d' |
    { printf '#line 1 "%s"\n' "$HERE/ref_umap_harness.cpp"; cat; }
  ranges "$SRC/fast_knn_umap.cpp" 3273-3325
  line=$(grep -n REF_HOGWILD_EDGE_LOOP "$HERE/ref_umap_harness.cpp" | head -1 | cut -d: -f1)
  printf '#line %s "%s"\n' "$((line + 1))" "$HERE/ref_umap_harness.cpp"
  sed -n "$((line + 1)),\$p" "$HERE/ref_umap_harness.cpp"
} | $CXX $FLAGS -x c++ -c - -o "$OUT/ref_umap.o"

$CXX -shared -pthread -o "$OUT/libref.so.tmp" "$OUT/ref_tsne.o" "$OUT/ref_umap.o"
mv "$OUT/libref.so.tmp" "$OUT/libref.so"
rm -f "$OUT/ref_tsne.o" "$OUT/ref_umap.o"
