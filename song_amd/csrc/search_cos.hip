// search_kernel instantiations for the COS metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_cos(bool exact, int shape) {
    return exact ? pick_shape<M_COS, true>(shape) : pick_shape<M_COS, false>(shape);
}
}  // namespace song
