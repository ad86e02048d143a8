#!/bin/bash
# Builds the CPU oracle (test infrastructure) into oracle/_build/liborc.so.
# -ffp-contract=off and no -march flags: plain x86-64 SSE arithmetic, the same float bits as
# dray's OpenMP CPU build of the restated functions.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
g++ -std=c++17 -O2 -fopenmp -fPIC -shared -ffp-contract=off -fno-fast-math \
    -o "$here/_build/liborc.so.tmp" "$here/drx_oracle.cpp"
mv "$here/_build/liborc.so.tmp" "$here/_build/liborc.so"
