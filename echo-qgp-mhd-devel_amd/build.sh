#!/bin/bash
# Builds libechoamd.so (HIP kernels for gfx950 + the C-ABI of include/echo_amd.h).
# FP64 multiply-adds are contracted to FMA (-ffp-contract=fast); results then
# agree with the reference to rounding rather than bit for bit.  Packed-FP32
# instructions are disabled (the kernels are FP64 and do not need them).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
ARCH=${ECHO_ARCH:-gfx950}
"$HIPCC" --offload-arch="$ARCH" -O3 -std=c++17 -fPIC -shared \
  -Wno-unused-value -Wno-unused-result -ffp-contract=fast -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"$HERE/../include" \
  -o "$HERE/libechoamd.so" "$HERE/csrc/echo_capi.hip"
echo "build: $HERE/libechoamd.so"
"$HIPCC" --offload-arch="$ARCH" -O3 -o "$HERE/copy_calib" "$HERE/csrc/copy_calib.hip"
