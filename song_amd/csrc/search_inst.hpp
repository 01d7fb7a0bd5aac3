// Instantiation table for search_kernel<METRIC, EXACT, SUB, QS>.
//   SUB: lanes that share one candidate row (4, 16, 64) -- picked from the row width.
//   QS : queue slots per lane (queue capacity QS*64) -- picked from pq + slack.
#pragma once

#include "search_kernel.hpp"

namespace song {

using SearchFn = void (*)(SearchParams);

constexpr int kSubs[] = {4, 16, 64};
constexpr int kQs[] = {2, 4, 8, 16, 32};
constexpr int kNumSub = 3, kNumQs = 5;

template <int METRIC, bool EXACT, int SUB>
inline SearchFn pick_qs(int qi) {
    switch (qi) {
    case 0: return search_kernel<METRIC, EXACT, SUB, 2>;
    case 1: return search_kernel<METRIC, EXACT, SUB, 4>;
    case 2: return search_kernel<METRIC, EXACT, SUB, 8>;
    case 3: return search_kernel<METRIC, EXACT, SUB, 16>;
    default: return search_kernel<METRIC, EXACT, SUB, 32>;
    }
}

template <int METRIC, bool EXACT>
inline SearchFn pick_sub(int si, int qi) {
    switch (si) {
    case 0: return pick_qs<METRIC, EXACT, 4>(qi);
    case 1: return pick_qs<METRIC, EXACT, 16>(qi);
    default: return pick_qs<METRIC, EXACT, 64>(qi);
    }
}

// defined in search_{l2,ip,cos}.hip
SearchFn search_fn_l2(bool exact, int si, int qi);
SearchFn search_fn_ip(bool exact, int si, int qi);
SearchFn search_fn_cos(bool exact, int si, int qi);

}  // namespace song
