#!/bin/bash
# Builds libechoamd.so (HIP kernels for gfx950 + the C-ABI of include/echo_amd.h)
# from three HIP translation units compiled in parallel: the grid update
# (echo_capi.hip), the batched / learned con2prim (c2p_mlp.hip) and the
# freeze-out hypersurface finder (freezeout.hip), and one host-only C++ unit
# (glauber_mc.cpp, the Glauber-MC event generator).
# FP64 multiply-adds are contracted to FMA (-ffp-contract=fast); results then
# agree with the reference to rounding rather than bit for bit.  Packed-FP32
# instructions are disabled.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
ARCH=${ECHO_ARCH:-gfx950}
OBJ="$HERE/_obj"
mkdir -p "$OBJ"
FLAGS=(--offload-arch="$ARCH" -O3 -std=c++17 -fPIC
       -Wno-unused-value -Wno-unused-result -ffp-contract=fast
       -Xclang -target-feature -Xclang -packed-fp32-ops -I"$HERE/../include")
"$HIPCC" "${FLAGS[@]}" -c -o "$OBJ/echo_capi.o" "$HERE/csrc/echo_capi.hip" &
P1=$!
# the network kernels take their MFMA accumulators in VGPRs rather than AGPRs:
# the activations read every accumulator, and AGPR results cost one
# v_accvgpr_read per value (208 per 64-state tile)
"$HIPCC" "${FLAGS[@]}" -mllvm -amdgpu-mfma-vgpr-form -c -o "$OBJ/c2p_mlp.o" "$HERE/csrc/c2p_mlp.hip" &
P2=$!
"$HIPCC" "${FLAGS[@]}" -c -o "$OBJ/freezeout.o" "$HERE/csrc/freezeout.hip" &
P3=$!
# host-only code (the Glauber-MC event generator)
g++ -O2 -std=c++17 -fPIC -Wall -I"$HERE/../include" -c -o "$OBJ/glauber_mc.o" "$HERE/csrc/glauber_mc.cpp"
wait $P1
wait $P2
wait $P3
"$HIPCC" --offload-arch="$ARCH" -shared -fPIC -o "$HERE/libechoamd.so" "$OBJ/echo_capi.o" "$OBJ/c2p_mlp.o" \
  "$OBJ/freezeout.o" "$OBJ/glauber_mc.o"
echo "build: $HERE/libechoamd.so"
"$HIPCC" "${FLAGS[@]}" -o "$HERE/copy_calib" "$HERE/csrc/copy_calib.hip"
