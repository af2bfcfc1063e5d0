#!/bin/sh
# Builds the CPU checkers (test infrastructure) into oracle/_build/:
#   liboracle.so  FP64 restatement of the reference's serial IMC path
#   libreplay.so  CPU replay of the engine's shared transport/sourcing headers
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_build"
g++ -std=c++17 -O2 -fPIC -shared -o "$here/_build/liboracle.so.tmp" "$here/oracle_capi.cpp"
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -o "$here/_build/libreplay.so.tmp" "$here/replay.cpp"
mv "$here/_build/libreplay.so.tmp" "$here/_build/libreplay.so"
