#!/usr/bin/env bash
# Builds libfa_hip.so for gfx950 in-tree (frame-aligner_amd/lib/).
# Packed FP32 instructions are switched off (-packed-fp32-ops).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
mkdir -p "$HERE/lib"
OUT="$HERE/lib/libfa_hip.so"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"$ROOT/include" -I"$HERE/csrc" \
  "$HERE/csrc/fa_kernels.hip" "$HERE/csrc/fa_api.hip" "$HERE/csrc/fa_frontend.hip" \
  -L/opt/rocm/lib -lrocfft -ldl -Wl,-rpath,/opt/rocm/lib \
  -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# C++ host mirror check program (plain g++ against the C ABI)
g++ -O2 -std=c++17 -I"$ROOT/include" -I"$HERE/host" "$HERE/host/fa_host_check.cpp" \
  -L"$HERE/lib" -lfa_hip -Wl,-rpath,'$ORIGIN' -o "$HERE/lib/fa_host_check"
echo "built $HERE/lib/fa_host_check"
