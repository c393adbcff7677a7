#!/bin/sh
# Compiles the reference's own C++ kernel module (src/mom25d/core/mom25d_core.cpp, with the CDT headers it
# vendors beside it) into oracle/_ref/, straight from the sources where they lie. Test infrastructure only:
# the result is a checker for the oracle's C restatement and the CPU baseline, never part of the product.
# Usage: oracle/build_ref.sh [reference root, default /root/reference]
set -e
REF=${1:-/root/reference}
SRC="$REF/src/mom25d/core/mom25d_core.cpp"
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$SRC" ]; then
    echo "build_ref: $SRC not found (no reference here): skipped" >&2
    exit 0
fi
mkdir -p "$OUT"
EXT=$(python3 -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
INC=$(python3 -m pybind11 --includes)
g++ -O3 -w -std=c++17 -fopenmp -shared -fPIC $INC -I"$REF/src/mom25d/core" "$SRC" -o "$OUT/_core$EXT.tmp"
mv "$OUT/_core$EXT.tmp" "$OUT/_core$EXT"
echo "build_ref: $OUT/_core$EXT"
