#!/bin/bash
# Builds make_h5_fixtures.c against the HDF5 library of this image and writes
# the three fixture files next to this script.  Needs the HDF5 headers and
# library (>= 1.10) under HDF5_PREFIX.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
HDF5_PREFIX=${HDF5_PREFIX:-/opt/conda}
TMP=$(mktemp -d)
gcc -O1 -o "$TMP/make_h5_fixtures" "$HERE/make_h5_fixtures.c" -I"$HDF5_PREFIX/include" \
    -L"$HDF5_PREFIX/lib" -lhdf5 -Wl,-rpath,"$HDF5_PREFIX/lib"
"$TMP/make_h5_fixtures" "$HERE"
rm -rf "$TMP"
ls -l "$HERE"/pfield_*.h5
