#!/bin/sh
# Build libdcs_amd.so for gfx950 in place (next to the Python package).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/../libdcs_amd.so"
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -Wno-c23-extensions \
  -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"$HERE/../../../include" \
  -o "$OUT.tmp" "$HERE/dcs_amd.hip" -L/opt/rocm/lib -lhiprtc
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
