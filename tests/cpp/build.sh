#!/bin/sh
# Builds the C++ programs that drive the engine through the drop-in manager
# header (include/storm_amd/MonteCarloManagerGPU.hpp) into tests/cpp/_build/.
# Needs storm_amd/_lib/libstorm_amd.so (storm_amd/build_lib.sh) first.
set -e
here=$(cd "$(dirname "$0")" && pwd)
root="$here/../.."
mkdir -p "$here/_build"
g++ -std=c++17 -O2 -Wall -Wextra -I"$root/include" -o "$here/_build/manager_slab.tmp" "$here/manager_slab.cpp" \
  -L"$root/storm_amd/_lib" -lstorm_amd -Wl,-rpath,'$ORIGIN/../../../storm_amd/_lib'
mv "$here/_build/manager_slab.tmp" "$here/_build/manager_slab"
g++ -std=c++17 -O2 -Wall -Wextra -Wno-overloaded-virtual -I"$root/include" -o "$here/_build/plugin_checks.tmp" \
  "$here/plugin_checks.cpp" -L"$root/storm_amd/_lib" -lstorm_amd -Wl,-rpath,'$ORIGIN/../../../storm_amd/_lib'
mv "$here/_build/plugin_checks.tmp" "$here/_build/plugin_checks"
g++ -std=c++17 -O2 -Wall -Wextra -I"$root/include" -o "$here/_build/manager_groups.tmp" "$here/manager_groups.cpp" \
  -L"$root/storm_amd/_lib" -lstorm_amd -Wl,-rpath,'$ORIGIN/../../../storm_amd/_lib'
mv "$here/_build/manager_groups.tmp" "$here/_build/manager_groups"
# the launch policy alone (host code, no library)
g++ -std=c++17 -O2 -Wall -Wextra -I"$root/include" -I"$root/storm_amd/csrc" -o "$here/_build/launch_policy.tmp" \
  "$here/launch_policy.cpp"
mv "$here/_build/launch_policy.tmp" "$here/_build/launch_policy"
