// several-queries-per-wave search instantiations for the L2 metric (one file per metric so they compile in parallel).
#include "group_inst.hpp"

namespace song {
SearchFn group_fn_l2(bool exact, int store, int shape, int lanes) {
    return group_pick_lanes<M_L2>(exact, store, shape, lanes);
}
}  // namespace song
