// search kernel instantiations for the COS metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_cos(bool exact, int store, int shape, bool persistent, bool bagl) {
    return pick_kernel<M_COS>(exact, store, shape, persistent, bagl);
}
SearchFn duo_fn_cos(int shape) { return pick_duo<M_COS>(shape); }
}  // namespace song
