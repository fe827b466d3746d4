#!/usr/bin/env bash
# Builds librsa_mpi_amd.so for gfx950 in-tree (rsa_mpi_amd/lib/).
# -ffp-contract=off keeps every FP64 expression unfused, as the reference's
# CPU build evaluates it; packed FP32 instructions are disabled.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function \
  -o "$here/lib/librsa_mpi_amd.so.tmp" \
  "$here/csrc/engine.hip" "$here/csrc/capi.cpp" "$here/csrc/mt19937.cpp" 2>&1 | grep -v "packed-fp32-ops" || true
test -f "$here/lib/librsa_mpi_amd.so.tmp"
mv "$here/lib/librsa_mpi_amd.so.tmp" "$here/lib/librsa_mpi_amd.so"
