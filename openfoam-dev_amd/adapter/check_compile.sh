#!/bin/sh
# Syntax-checks the OpenFOAM adapter against the OpenFOAM headers of a source
# tree (default /root/reference), the way wmake would see them: a flat
# lnInclude directory of symlinks to every header of src/OpenFOAM and
# src/OSspecific/POSIX, built under _build/.  Usage: check_compile.sh [FOAM_SRC]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
FOAM_SRC=${1:-/root/reference/src}
LN="$HERE/_build/lnInclude"
if [ ! -d "$LN" ]; then
    mkdir -p "$LN"
    for d in OpenFOAM OSspecific/POSIX; do
        find "$FOAM_SRC/$d" \( -name '*.H' -o -name '*.C' \) -exec ln -sf {} "$LN/" \;
    done
fi
for f in hipLduSolver.C hipLduPreconditioner.C; do
    g++ -std=c++14 -m64 -fsyntax-only -DWM_DP -DWM_LABEL_SIZE=32 -DNoRepository \
        -Wno-deprecated-declarations -I"$LN" -I"$ROOT/include" -I"$HERE" "$HERE/$f"
    echo "$f: OK"
done
