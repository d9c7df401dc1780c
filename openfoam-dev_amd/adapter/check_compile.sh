#!/bin/sh
# Syntax-checks the OpenFOAM adapter against the headers of an OpenFOAM source
# tree (default /root/reference/src) the way wmake would see them: one
# lnInclude directory of symlinks per library (OpenFOAM, OSspecific/POSIX and
# the libraries Make/options names: finiteVolume, meshTools, plus the ones
# their headers include), built under _build/, searched in that order.
# Usage: check_compile.sh [FOAM_SRC]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
FOAM_SRC=${1:-/root/reference/src}
LIBS="OpenFOAM OSspecific/POSIX finiteVolume meshTools triSurface surfMesh fileFormats"
INC=""
for d in $LIBS; do
    LN="$HERE/_build/$d/lnInclude"
    if [ ! -d "$LN" ]; then
        mkdir -p "$LN"
        find "$FOAM_SRC/$d" \( -name '*.H' -o -name '*.C' \) -exec ln -sf {} "$LN/" \;
    fi
    INC="$INC -I$LN"
done
for f in hipLduSolver.C hipLduPreconditioner.C; do
    g++ -std=c++14 -m64 -fsyntax-only -DWM_DP -DWM_LABEL_SIZE=32 -DNoRepository \
        -Wno-deprecated-declarations $INC -I"$ROOT/include" -I"$HERE" "$HERE/$f"
    echo "$f: OK"
done
