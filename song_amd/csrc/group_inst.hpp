// Instantiation table of search_group_persistent<METRIC, EXACT, STORE, SHAPE, L> (several queries
// per wave): L = 16 or 32 lanes per query, SHAPE 0..3 (a row's lanes lie in one sub-group).
#pragma once

#include "search_group.hpp"
#include "search_inst.hpp"

namespace song {

template <int METRIC, bool EXACT, int STORE, int L>
inline SearchFn group_pick_shape(int shape) {
    switch (shape) {
    case 0: return search_group_persistent<METRIC, EXACT, STORE, 0, L>;
    case 1: return search_group_persistent<METRIC, EXACT, STORE, 1, L>;
    case 2: return search_group_persistent<METRIC, EXACT, STORE, 2, L>;
    default: return search_group_persistent<METRIC, EXACT, STORE, 3, L>;
    }
}

template <int METRIC, int L>
inline SearchFn group_pick(bool exact, int store, int shape) {
    if (exact) return group_pick_shape<METRIC, true, ST_F32, L>(shape);
    switch (store) {
    case ST_BF16: return group_pick_shape<METRIC, false, ST_BF16, L>(shape);
    case ST_FP8: return group_pick_shape<METRIC, false, ST_FP8, L>(shape);
    default: return group_pick_shape<METRIC, false, ST_F32, L>(shape);
    }
}

template <int METRIC>
inline SearchFn group_pick_lanes(bool exact, int store, int shape, int lanes) {
    return lanes == 16 ? group_pick<METRIC, 16>(exact, store, shape) : group_pick<METRIC, 32>(exact, store, shape);
}

}  // namespace song
