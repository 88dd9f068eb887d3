#!/bin/sh
# Regenerates tests/golden/philox_rocrand.json from rocRAND's philox4x32_10 engine.
set -e
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
hipcc -O1 -o "$tmp/philox_rocrand" "$here/philox_rocrand.cpp"
"$tmp/philox_rocrand" > "$here/philox_rocrand.json"
rm -rf "$tmp"
