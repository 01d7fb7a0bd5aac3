// several-queries-per-wave search instantiations for the IP metric (one file per metric so they compile in parallel).
#include "group_inst.hpp"

namespace song {
SearchFn group_fn_ip(bool exact, int store, int shape, int lanes) {
    return group_pick_lanes<M_IP>(exact, store, shape, lanes);
}
}  // namespace song
