#!/usr/bin/env bash
# Builds the host library (scene, BVH, renderer, batch, C-ABI, image writers)
# against hipstub.cpp, a CPU stand-in for the HIP runtime, with
# AddressSanitizer, into $1/libomv_hostsim.so.  Host code only (g++); no
# device code is built.
set -euo pipefail
OUT="${1:?output directory}"
HERE="$(cd "$(dirname "$0")" && pwd)"
C="$HERE/../../openmaterial-validation_amd/csrc"
FL=(-std=c++17 -O1 -g -fPIC -fsanitize=address -fno-omit-frame-pointer -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I"$C" -ffp-contract=off)
mkdir -p "$OUT"
pids=()
for f in scene bvh renderer batch capi image_io; do g++ "${FL[@]}" -c "$C/$f.cpp" -o "$OUT/$f.o" & pids+=($!); done
g++ "${FL[@]}" -c "$HERE/hipstub.cpp" -o "$OUT/hipstub.o" & pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done
g++ -shared -fsanitize=address -Wl,-z,defs -o "$OUT/libomv_hostsim.so" "$OUT"/scene.o "$OUT"/bvh.o "$OUT"/renderer.o "$OUT"/batch.o \
    "$OUT"/capi.o "$OUT"/image_io.o "$OUT"/hipstub.o -lpthread
