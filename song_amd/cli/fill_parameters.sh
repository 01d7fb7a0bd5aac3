#!/usr/bin/env bash
# Compile the `song` binary for a search budget (pq), dimension and metric and copy it to
# the current directory (same usage as the reference fill_parameters.sh).  The kernels
# take pq and dim at run time; the compiled values become the binary's defaults.
if [ "$#" -ne 3 ]; then
  echo "Usage: $0 <pq_size> <dim> <cos/l2/ip>" >&2
  echo "For example: $0 100 26 cos" >&2
  exit 1
fi
case "$3" in
  cos) metric=2 ;;
  l2) metric=0 ;;
  ip) metric=1 ;;
  *) echo "Usage: $0 <pq_size> <dim> <cos/l2/ip>" >&2; exit 1 ;;
esac
here="$(cd "$(dirname "$0")" && pwd)"
cat > "$here/song_params.h" <<PARAMS
#define SONG_PQ $1
#define SONG_DIM $2
#define SONG_METRIC $metric
PARAMS
make -C "$here/.." bin/song && cp "$here/../bin/song" . && cp "$here/build_graph.sh" "$here/test_query.sh" .
