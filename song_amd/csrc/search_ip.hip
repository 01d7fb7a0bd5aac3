// search_kernel instantiations for the IP metric (one file per metric so they compile in parallel).
#include "search_inst.hpp"

namespace song {
SearchFn search_fn_ip(bool exact, int si, int qi) {
    return exact ? pick_sub<M_IP, true>(si, qi) : pick_sub<M_IP, false>(si, qi);
}
}  // namespace song
