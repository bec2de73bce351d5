#!/usr/bin/env bash
# ORACLE -- test infrastructure only.
#
# Builds oracle/_ref/libref_acoustic.so: the reference's own acoustic 2-D/3-D
# and elastic 2-D translation units, compiled from where they lie under /root/reference with
# g++ as C++ (the way the reference builds its CPU "host core",
# backend/c/hostcore.py:64-82: its CUDA stand-in headers under
# csrc/host/include first on the include path, cuda_runtime.h forced in),
# linked with oracle/ref_harness.cpp.  Nothing is copied out of the reference;
# objects and the library go to oracle/_ref/ only (git-ignored).  Exits 0
# without building when the reference is absent.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
REF="${SWEEP_REFERENCE:-/root/reference}/src/sweep/csrc"
[ -d "$REF" ] || { echo "build_ref: no reference tree at $REF, skipped"; exit 0; }
out="$here/_ref"
obj="$out/obj"
mkdir -p "$obj"
CXX="${CXX:-g++}"
FLAGS=(-x c++ -std=c++17 -O3 -fPIC -fopenmp -fvisibility=hidden -ffp-contract=off
       -ftls-model=initial-exec -DSWEEP_HOST_CORE -Wno-unknown-pragmas -Wno-attributes -w
       -isystem "$REF/host/include" -include "$REF/host/include/cuda_runtime.h"
       -I"$REF" -I"$REF/bindings" -I"$REF/shared" -I"$REF/cuda" -I"$REF/cuda/common" -I"$REF/cuda/equations")
SRCS=(
  cuda/equations/acoustic2d/forward.cu cuda/equations/acoustic2d/forward_split.cu
  cuda/equations/acoustic2d/backward.cu cuda/equations/acoustic2d/adjoint_split.cu
  cuda/equations/acoustic2d/kernels.cu
  cuda/equations/acoustic3d/forward.cu cuda/equations/acoustic3d/forward_split.cu
  cuda/equations/acoustic3d/backward.cu cuda/equations/acoustic3d/adjoint_split.cu
  cuda/equations/acoustic3d/kernels.cu
  cuda/equations/elastic2d/forward.cu cuda/equations/elastic2d/backward.cu cuda/equations/elastic2d/kernels.cu
  cuda/common/common.cu cuda/common/boundarysaver.cu cuda/common/derived_models.cu
  cuda/common/layout.cu cuda/common/boundary/session.cu
)
objs=()
pids=()
for s in "${SRCS[@]}"; do
  o="$obj/$(echo "$s" | tr '/' '_').o"
  objs+=("$o")
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    "$CXX" "${FLAGS[@]}" -c "$REF/$s" -o "$o" &
    pids+=($!)
  fi
done
for p in "${pids[@]}"; do wait "$p"; done
"$CXX" "${FLAGS[@]}" -c "$here/ref_harness.cpp" -o "$obj/ref_harness.o"
"$CXX" -shared -fopenmp -Wl,-z,defs -Wl,-Bsymbolic-functions -o "$out/libref_acoustic.so" \
  "$obj/ref_harness.o" "${objs[@]}"
echo "build_ref: $out/libref_acoustic.so"
