#!/bin/bash
# Copies the reference repository's own input data used by tests/test_reference_assets.py
# into tests/golden/assets (data files, unchanged): the vortex-street configuration
# (assets/configs/vortexstreet2d.json: 801x201 grid, 1602 Wall + 201 InletLeft + 201
# OutletRight cells, a static sphere2d.mesh body) and three meshes (assets/mesh/
# sphere2d.mesh, fish2d.mesh, sphere_coarse.mesh).  Run in the container that holds
# the reference checkout; the copies travel with the tests.
set -e
SRC=${1:-/root/reference/assets}
DST=$(dirname "$0")/assets
mkdir -p "$DST"
cp "$SRC/configs/vortexstreet2d.json" "$DST/"
cp "$SRC/mesh/sphere2d.mesh" "$SRC/mesh/fish2d.mesh" "$SRC/mesh/sphere_coarse.mesh" "$DST/"
sha256sum "$DST"/* | sed "s#$DST/##" > "$DST/SHA256SUMS"
