#!/bin/bash
# Regenerates tests/golden/philox_rocrand.json from rocRAND's Philox header (host code).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
TMP="$(mktemp -d)"
/opt/rocm/bin/hipcc -x hip --offload-arch=gfx950 -O1 -o "$TMP/gen" "$HERE/gen_philox_rocrand.cpp"
"$TMP/gen" > "$HERE/philox_rocrand.json"
rm -rf "$TMP"
