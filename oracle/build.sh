#!/bin/sh
# Builds the CPU checkers (test infrastructure) into oracle/_build/:
#   liboracle.so  FP64 restatement of the reference's serial IMC path
#   libreplay.so  CPU replay of the engine's shared transport/sourcing headers
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_build"
# per-process temporaries and a rename, so that concurrent builds (parallel
# test workers) never see a half-written library
g++ -std=c++17 -O2 -fPIC -shared -o "$here/_build/liboracle.so.$$" "$here/oracle_capi.cpp"
mv -f "$here/_build/liboracle.so.$$" "$here/_build/liboracle.so"
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -o "$here/_build/libreplay.so.$$" "$here/replay.cpp"
mv -f "$here/_build/libreplay.so.$$" "$here/_build/libreplay.so"
