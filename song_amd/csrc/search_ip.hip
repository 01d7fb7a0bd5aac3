// search kernel instantiations for the IP metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_ip(bool exact, int store, int shape, bool persistent, bool bagl, bool bg) {
    return pick_kernel<M_IP>(exact, store, shape, persistent, bagl, bg);
}
SearchFn duo_fn_ip(int shape, bool bg) { return pick_duo<M_IP>(shape, bg); }
}  // namespace song
