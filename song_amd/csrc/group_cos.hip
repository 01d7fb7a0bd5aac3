// several-queries-per-wave search instantiations for the COS metric (one file per metric so they compile in parallel).
#include "group_inst.hpp"

namespace song {
SearchFn group_fn_cos(bool exact, int store, int shape, int lanes) {
    return group_pick_lanes<M_COS>(exact, store, shape, lanes);
}
}  // namespace song
