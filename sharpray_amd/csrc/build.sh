#!/usr/bin/env bash
# Builds libsharpray_hip.so (gfx950 kernels + C ABI + host scene preparation) in-tree.
#  -ffp-contract=off            no FMA contraction: every float op rounds like the reference's
#  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
#  -fno-gpu-flush-denormals-to-zero           keep fp32 denormals (x64 .NET does)
#  -packed-fp32-ops disabled    no v_pk_*_f32 from the vectoriser
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
out="${SR_OUT:-${here}/../libsharpray_hip.so}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="${SR_OFFLOAD_ARCH:-gfx950}"
COMMON=(-std=c++17 -O3 -fPIC -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function ${SR_EXTRA_FLAGS:-})
DEVICE=(--offload-arch="${ARCH}" -fhip-fp32-correctly-rounded-divide-sqrt -fno-gpu-flush-denormals-to-zero
        -Xclang -target-feature -Xclang -packed-fp32-ops)
tmp="$(mktemp -d)"
trap 'rm -rf "${tmp}"' EXIT
"${HIPCC}" "${COMMON[@]}" -x hip --offload-host-only -c "${here}/sr_host.cpp" -o "${tmp}/sr_host.o" -fno-builtin-powf
"${HIPCC}" "${COMMON[@]}" -x hip --offload-host-only -c "${here}/sr_ingest.cpp" -o "${tmp}/sr_ingest.o" -fno-builtin-powf
"${HIPCC}" "${COMMON[@]}" "${DEVICE[@]}" -c "${here}/sr_render.hip" -o "${tmp}/sr_render.o"
"${HIPCC}" "${COMMON[@]}" "${DEVICE[@]}" -c "${here}/sr_cnn.hip" -o "${tmp}/sr_cnn.o"
"${HIPCC}" "${COMMON[@]}" "${DEVICE[@]}" -c "${here}/sr_bvh.hip" -o "${tmp}/sr_bvh.o"
"${HIPCC}" -shared --offload-arch="${ARCH}" "${tmp}/sr_host.o" "${tmp}/sr_ingest.o" "${tmp}/sr_render.o" "${tmp}/sr_cnn.o" "${tmp}/sr_bvh.o" -o "${out}.tmp"
mv "${out}.tmp" "${out}"
echo "built ${out}"
