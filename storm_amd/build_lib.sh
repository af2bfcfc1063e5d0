#!/bin/sh
# Builds the gfx950 engine library storm_amd/_lib/libstorm_amd.so.
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_lib"
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"$here/../include" \
  -o "$here/_lib/libstorm_amd.so.tmp" "$here/csrc/engine.hip"
mv "$here/_lib/libstorm_amd.so.tmp" "$here/_lib/libstorm_amd.so"
