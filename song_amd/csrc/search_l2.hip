// search kernel instantiations for the L2 metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_l2(bool exact, int store, int shape, bool persistent, bool bagl, bool bg) {
    return pick_kernel<M_L2>(exact, store, shape, persistent, bagl, bg);
}
SearchFn duo_fn_l2(int shape, bool bg) { return pick_duo<M_L2>(shape, bg); }
}  // namespace song
