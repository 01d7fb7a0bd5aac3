#!/bin/bash
# Search every query of a LIBSVM file against ./bfsg.data and ./bfsg.graph on the GPU
# (same usage as the reference test_query.sh).
if [ "$#" -lt 4 ]; then
  echo "Usage: $0 <query_data> <built_graph_row> <built_graph_dimension> <l2/ip/cos> [display top k]" >&2
  echo "For example: $0 letter.scale.t 15000 26 cos" >&2
  echo "Use display top 5: $0 letter.scale.t 15000 26 cos 5" >&2
  exit 1
fi
search=100  # pq when the binary was built without fill_parameters.sh
display=100
if [ "$#" -ge 5 ]; then
  display=$5
fi
"$(dirname "$0")/song" test 0 "$1" ${search} "$2" "$3" ${display} "$4"
