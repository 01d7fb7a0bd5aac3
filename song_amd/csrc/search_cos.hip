// search kernel instantiations for the COS metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_cos(bool exact, int store, int shape, bool persistent, bool bagl, bool bg) {
    return pick_kernel<M_COS>(exact, store, shape, persistent, bagl, bg);
}
SearchFn duo_fn_cos(int shape, bool bg) { return pick_duo<M_COS>(shape, bg); }
}  // namespace song
