// search_kernel instantiations for the L2 metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_l2(bool exact, int si, int qi) {
    return exact ? pick_sub<M_L2, true>(si, qi) : pick_sub<M_L2, false>(si, qi);
}
}  // namespace song
