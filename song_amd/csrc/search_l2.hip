// search_kernel instantiations for the L2 metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_l2(bool exact, int shape) {
    return exact ? pick_shape<M_L2, true>(shape) : pick_shape<M_L2, false>(shape);
}
}  // namespace song
