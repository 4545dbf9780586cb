#!/usr/bin/env bash
# Builds the reference CPU join (CPU_version1 of the reference repository) from its
# own sources where they lie, plus the one-config driver oracle/ref_driver.cpp.
# Output goes only to oracle/_ref/ (git-ignored).  The reference's AAAA.cpp main is
# left out because it ignores argv[3..4] and sweeps a fixed (m, r) grid
# (CPU_version1/AAAA.cpp:30-38); the driver runs exactly one (m, w).
# Four reference files use memcpy/strlen without including <cstring>; they are
# compiled with "-include cstring" (a compiler flag, no source is changed).
set -euo pipefail
REF=${REF:-/root/reference/CPU_version1}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "reference sources not present at $REF; skipping oracle/_ref build" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
# Source list = the reference's add_executable list (CPU_version1/CMakeLists.txt:15-153)
# minus headers and minus AAAA.cpp.
SRCS=(
dist_mpx_v201_fast6.cpp allmatrixop/matrix_transpose.cpp allmatrixop/matrix_false.cpp
allmatrixop/matrix_convertTo2DRowOrder.cpp allmatrixop/matrix_convertTo2DColOrder.cpp
allmatrixop/movmean.cpp allmatrixop/movstd.cpp allmatrixop/doubledividedbymatrix.cpp
allmatrixop/extractVector.cpp allmatrixop/movmax.cpp allmatrixop/movmin.cpp dr.cpp dr_f.cpp
allmatrixop/plus.cpp allmatrixop/plusvector.cpp allmatrixop/movsum.cpp allmatrixop/vectormultiply.cpp
allmatrixop/norm_vector.cpp allmatrixop/sum_vector.cpp allmatrixop/extractvectorfromvector.cpp
allmatrixop/or.cpp allmatrixop/sum_bool_matrix.cpp alltest/test2fast6.cpp allmatrixop/vectorcopy.cpp
allmatrixop/addelementtofront.cpp "dist mpx_v201 selectedV4_.cpp" allmatrixop/substractvector.cpp
allmatrixop/std.cpp alltest/test2V4.cpp dist_mpx_DIAGV3_fastF.cpp compute_shared_dataV5.cpp
allmatrixop/maxinmatrix.cpp allmatrixop/mininmatrix.cpp alltest/test2dataV5.cpp
allmatrixop/vectordivision.cpp allmatrixop/meanwithoutNAN.cpp alltest/test2fasfF.cpp MY_dtw_mpGUIV2.cpp
lb_upd.cpp dtw_upd.cpp NEXT.cpp allmatrixop/isempty.cpp alltest/test2NEXT.cpp
dist_mpx_v2O1_selectedV6_test_pro_max.cpp linspace.cpp allmatrixop/mergevector.cpp
allmatrixop/sum_vector_bool.cpp compute_shared_data.cpp alltest/test2data.cpp allmatrixop/sum_matrix.cpp
dist_LB_in_one_Line_O1_V2.cpp allmatrixop/stdwithomitnan.cpp allmatrixop/dotproduct.cpp
allmatrixop/elementwisemultiply_nv.cpp allmatrixop/Mean.cpp alltest/test2LB.cpp alltest/test2promax.cpp
dist_mpx_DIAGV3_fast.cpp compute_shared_dataV3.cpp alltest/test2fast.cpp alltest/test2data3.cpp
dist_mpx_DIAGV3.cpp alltest/test2DIAGV3.cpp dist_mpx_v2O1_selectedV6_.cpp mpx_v2.cpp
allmatrixop/isinfinite.cpp findNonZero.cpp allmatrixop/isNAN.cpp allmatrixop/moving_mean.cpp
findDiscords.cpp allmatrixop/sortvec.cpp findMotifs.cpp allmatrixop/convolve.cpp allmatrixop/nextpow2.cpp
fft.cpp allmatrixop/conj.cpp allmatrixop/elementmultiply_complex.cpp ifft.cpp allmatrixop/crosscov.cpp
allmatrixop/min_nv_includenan.cpp allmatrixop/max_v_idx.cpp dtw_motifGUI.cpp allmatrixop/min_v.cpp
DTWMotifDiscoveryGUI.cpp loadfile.cpp alltest/test2selected6.cpp alltest/test2mpx_v2.cpp
alltest/test2findIdx.cpp alltest/test2findmotifs.cpp alltest/test2moving_mean.cpp
alltest/test2dtw_motifGUI.cpp alltest/test2MY-dtw_mpguiv2.cpp alltest/test2DTW.cpp alltest/test2loadfile.cpp
DTW_Motif_of_one_layer.cpp allmatrixop/vectordividedbydouble.cpp allmatrixop/lower_upper_lemire.cpp
sampling.cpp allmatrixop/nomalise.cpp diag_fast.cpp diag_mask_local_up.cpp new_dtw_motifGUI.cpp
lb_kim.cpp lb_keogh.cpp LB_PETITJEAN.cpp dtw.cpp diag_mask_global.cpp alpha_generater.cpp
diag_mask_local_down.cpp diag_mask_local.cpp SS-PrunedDTW.cpp alltest/test_for_SS_dtw.cpp MON_DTW.cpp
compute_shared_data_global.cpp compute_MASK_global.cpp
)
CXXFLAGS="-std=c++17 -O2 -w -include cstring -I$REF"
JOBS=${JOBS:-8}
OBJS=()
i=0
pids=()
for s in "${SRCS[@]}"; do
  o="$OUT/obj/$(echo "$s" | tr '/ ' '__').o"
  OBJS+=("$o")
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    g++ $CXXFLAGS -c "$REF/$s" -o "$o" &
    pids+=($!)
    i=$((i+1))
    if [ $((i % JOBS)) -eq 0 ]; then for p in "${pids[@]}"; do wait "$p"; done; pids=(); fi
  fi
done
for p in "${pids[@]}"; do wait "$p"; done
g++ $CXXFLAGS -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver.o"
g++ -o "$OUT/ref_dtw" "${OBJS[@]}" "$OUT/obj/ref_driver.o"
echo "built $OUT/ref_dtw"
