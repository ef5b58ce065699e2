#!/bin/sh
# Regenerates tests/golden/philox_kat.json from rocRAND's Philox4x32-10
# (host code only; no GPU needed).
set -e
here=$(cd "$(dirname "$0")" && pwd)
out=$(mktemp -d)
hipcc --offload-arch=gfx950 -O1 -o "$out/make_philox_vectors" "$here/make_philox_vectors.cpp"
"$out/make_philox_vectors" > "$here/philox_kat.json"
rm -rf "$out"
