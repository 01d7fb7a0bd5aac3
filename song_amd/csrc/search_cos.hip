// search_kernel instantiations for the COS metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_cos(bool exact, int si, int qi) {
    return exact ? pick_sub<M_COS, true>(si, qi) : pick_sub<M_COS, false>(si, qi);
}
}  // namespace song
