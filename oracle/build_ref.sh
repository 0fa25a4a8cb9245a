#!/usr/bin/env bash
# Compiles the reference's CPU SynC (src/algorithms/SynC.cpp, src/utils/CPU_math.cpp)
# straight from /root/reference with g++, together with oracle/ref_harness.cpp,
# into oracle/_ref/sync_ref.  Nothing from the reference is copied into the
# repository; the binary stays out of git (oracle/_ref/ is ignored).
#
# Flags follow the reference's own build (python/algorithms.py:15-16:
# "-w --std=c++17 -fopenmp"); -O2 is added, which does not change IEEE results
# on x86-64 (SSE arithmetic, no FMA contraction without -mfma).
# CPU_math.cpp uses printf without including <cstdio> (CPU_math.cpp:165), so
# <cstdio> is force-included.  ATen headers and libtorch_cpu/libc10 come from
# the installed torch; the reference uses ATen only for result tensors.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/algorithms/SynC.cpp" ]; then
    echo "reference sources not found under $REF; skipping" >&2
    exit 0
fi
TORCH=$(python3 -c 'import os, torch; print(os.path.dirname(torch.__file__))')
mkdir -p "$OUT"
g++ -O2 -w --std=c++17 -fopenmp -include cstdio \
    -I"$REF/src" -I"$TORCH/include" -I"$TORCH/include/torch/csrc/api/include" \
    "$HERE/ref_harness.cpp" "$REF/src/algorithms/SynC.cpp" "$REF/src/utils/CPU_math.cpp" \
    -L"$TORCH/lib" -Wl,-rpath,"$TORCH/lib" -ltorch_cpu -lc10 \
    -o "$OUT/sync_ref"
echo "built $OUT/sync_ref"
