#!/bin/sh
# Compiles the reference's own DOP853 stepper (src/math_ode.cpp, which needs no third-party
# library) together with the test driver oracle/ref_driver.cpp into oracle/_ref/.
# The rest of the reference path (Multipole, splines, spherical harmonics) includes GSL
# headers that this image lacks and is therefore not built.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${AGAMA_REFERENCE:-/root/reference}
if [ ! -f "$REF/src/math_ode.cpp" ]; then
    echo "reference tree not found at $REF; skipping oracle/_ref" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
${CXX:-g++} -std=c++11 -O2 -march=x86-64-v3 -ffp-contract=off -fPIC -shared \
    -I"$REF/src" "$REF/src/math_ode.cpp" "$HERE/ref_driver.cpp" \
    -o "$HERE/_ref/libref_dop853.so"
echo "built $HERE/_ref/libref_dop853.so"
