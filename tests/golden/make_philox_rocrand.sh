#!/bin/sh
# Regenerates tests/golden/philox_rocrand.json from rocRAND's Philox engine.
set -e
d=$(dirname "$0")
/opt/rocm/bin/hipcc -O1 -std=c++17 -o /tmp/philox_rocrand "$d/philox_rocrand.cpp"
/tmp/philox_rocrand > "$d/philox_rocrand.json"
