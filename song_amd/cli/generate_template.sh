#!/usr/bin/env bash
# Build the song_amd libraries (the reference's template generation step has no
# equivalent: the HIP kernels take pq and dim at run time).
set -e
make -C "$(dirname "$0")/.." -j"$(nproc)"
