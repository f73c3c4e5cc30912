// Tile lists of a multi-group batch: host-only bookkeeping with no device
// work, kept apart from the HIP calls of the batch driver so it can be
// built and checked on the CPU (under AddressSanitizer in the tests).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "oslmi_groupargs.h"

namespace oslmi {

// The tiles of one batch: every item's points cut into tiles of tile_size
// points, the items of each part (kernel) together and, within a part, in
// the part's item order or (by_cost, with one cost per item) by measured
// cost per tile, most expensive first.
struct TileLayout {
    std::vector<BatchTile> tiles;
    std::vector<int> item_first, item_count;  // tile range of every item
    std::vector<int> part_first, part_ntiles; // tile range of every part
};

enum TileListStatus { kTilesOk = 0, kTilesBadInput, kTilesTooMany };

// counts[k]: points of item k (>= 0); part_items[q]: the items of part q,
// each item index in [0, counts.size()). tile_size >= 1. Fails (leaving
// out unspecified) on input outside those bounds or when the list would
// pass 2^31 - 1 tiles.
inline TileListStatus tile_list(const std::vector<int64_t>& counts,
                                const std::vector<std::vector<int>>& part_items, int64_t tile_size,
                                bool by_cost, const std::vector<double>& cost, TileLayout& out)
{
    const int ni = (int)counts.size();
    if (tile_size < 1)
        return kTilesBadInput;
    for (int64_t c : counts)
        if (c < 0)
            return kTilesBadInput;
    for (const std::vector<int>& items : part_items)
        for (int k : items)
            if (k < 0 || k >= ni)
                return kTilesBadInput;
    out.tiles.clear();
    out.item_first.assign(ni, 0);
    out.item_count.assign(ni, 0);
    out.part_first.assign(part_items.size(), 0);
    out.part_ntiles.assign(part_items.size(), 0);
    const bool sorted = by_cost && (int)cost.size() == ni;
    for (size_t q = 0; q < part_items.size(); ++q) {
        std::vector<int> order = part_items[q];
        if (sorted)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        out.part_first[q] = (int)out.tiles.size();
        for (int k : order) {
            out.item_first[k] = (int)out.tiles.size();
            for (int64_t b = 0; b < counts[k]; b += tile_size) {
                if (out.tiles.size() >= (size_t)0x7fffffff)
                    return kTilesTooMany;
                BatchTile t;
                t.item  = k;
                t.pad   = 0;
                t.begin = b;
                t.end   = std::min<int64_t>(b + tile_size, counts[k]);
                out.tiles.push_back(t);
            }
            out.item_count[k] = (int)out.tiles.size() - out.item_first[k];
        }
        out.part_ntiles[q] = (int)out.tiles.size() - out.part_first[q];
    }
    return kTilesOk;
}

}  // namespace oslmi
