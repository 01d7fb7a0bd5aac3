// search_kernel instantiations for the IP metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_ip(bool exact, int shape) {
    return exact ? pick_shape<M_IP, true>(shape) : pick_shape<M_IP, false>(shape);
}
}  // namespace song
