#!/usr/bin/env bash
# Builds the reference's own host-only scene exporter (detector/bench/
# scene_gen.cpp with scenes.hpp, test_support.hpp and tag_families.hpp) from
# its sources under /root/reference into oracle/_ref/.  It is a checker for
# the scene generator port; the reference's GPU pipeline needs nvcc and the
# CUDA runtime and is not buildable here.
set -euo pipefail
REF=/root/reference/detector
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
mkdir -p "$OUT"
g++ -std=c++20 -O2 -I"$REF/src" -I"$REF/tests" -I"$REF/bench" \
    -o "$OUT/bdcudatag_scenes" "$REF/bench/scene_gen.cpp"
echo "built $OUT/bdcudatag_scenes"
