// CPU check of the batch tile bookkeeping (oslmi_tiles.h), built with
// AddressSanitizer and UBSan by tests/test_tiles_host.py: random batches
// (item counts from 0 to a few tiles, parts of shuffled items, every tile
// size and order the autotuner tries) and the invariants the batch kernel
// relies on: each item's points covered once, in order, by tiles no larger
// than the tile size; each item's and each part's tiles contiguous; cost
// order non-increasing within a part. Prints "ok <cases>" on success.
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "oslmi_tiles.h"

using namespace oslmi;

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static uint32_t rnd()
{
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(g_state >> 33);
}

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return false;                                                \
        }                                                                \
    } while (0)

static bool check_case(int ni, int nparts, int64_t ts, bool by_cost)
{
    std::vector<int64_t> counts(ni);
    for (int k = 0; k < ni; ++k)
        counts[k] = (rnd() % 5 == 0) ? 0 : (int64_t)(rnd() % (4 * ts + 3));
    std::vector<int> perm(ni);
    std::iota(perm.begin(), perm.end(), 0);
    for (int k = ni - 1; k > 0; --k)
        std::swap(perm[k], perm[rnd() % (k + 1)]);
    std::vector<std::vector<int>> parts(nparts);
    for (int k = 0; k < ni; ++k)
        parts[k % nparts].push_back(perm[k]);
    std::vector<double> cost(ni);
    for (double& c : cost)
        c = (double)(rnd() % 7);  // ties keep the part's order
    TileLayout lay;
    CHECK(tile_list(counts, parts, ts, by_cost, cost, lay) == kTilesOk);
    CHECK((int)lay.item_first.size() == ni && (int)lay.item_count.size() == ni);
    CHECK((int)lay.part_first.size() == nparts && (int)lay.part_ntiles.size() == nparts);
    int64_t expect = 0;
    for (int k = 0; k < ni; ++k) {
        const int64_t nt = (counts[k] + ts - 1) / ts;
        expect += nt;
        CHECK(lay.item_count[k] == nt);
        int64_t next = 0;
        for (int t = lay.item_first[k]; t < lay.item_first[k] + lay.item_count[k]; ++t) {
            CHECK(t >= 0 && t < (int)lay.tiles.size());
            const BatchTile& b = lay.tiles[t];
            CHECK(b.item == k && b.begin == next && b.end > b.begin && b.end - b.begin <= ts);
            next = b.end;
        }
        CHECK(next == counts[k]);
    }
    CHECK((int64_t)lay.tiles.size() == expect);
    int at = 0;
    for (int q = 0; q < nparts; ++q) {
        CHECK(lay.part_first[q] == at);
        int covered = 0;
        double last = 1e300;
        for (int t = lay.part_first[q]; t < lay.part_first[q] + lay.part_ntiles[q]; ++t) {
            const int k = lay.tiles[t].item;
            bool mine = false;
            for (int j : parts[q])
                mine = mine || j == k;
            CHECK(mine);
            if (by_cost) {
                CHECK(cost[k] <= last);
                last = cost[k];
            }
        }
        for (int k : parts[q])
            covered += lay.item_count[k];
        CHECK(covered == lay.part_ntiles[q]);
        at += lay.part_ntiles[q];
    }
    CHECK(at == (int)lay.tiles.size());
    return true;
}

int main()
{
    int cases = 0;
    for (int rep = 0; rep < 200; ++rep)
        for (int64_t ts : { 64, 128, 256, 512, 1024 })
            for (int by_cost = 0; by_cost <= 1; ++by_cost) {
                const int ni     = (int)(rnd() % 70);
                const int nparts = 1 + (int)(rnd() % 9);
                if (!check_case(ni, nparts, ts, by_cost != 0))
                    return 1;
                ++cases;
            }
    // bad input is refused without touching the output
    TileLayout lay;
    std::vector<double> none;
    if (tile_list({ 5, -1 }, { { 0, 1 } }, 64, false, none, lay) != kTilesBadInput)
        return 2;
    if (tile_list({ 5 }, { { 0, 1 } }, 64, false, none, lay) != kTilesBadInput)
        return 3;
    if (tile_list({ 5 }, { { 0 } }, 0, false, none, lay) != kTilesBadInput)
        return 4;
    // an empty batch and parts without items
    if (tile_list({}, { {}, {} }, 64, true, none, lay) != kTilesOk || !lay.tiles.empty()
        || lay.part_ntiles.size() != 2)
        return 5;
    printf("ok %d\n", cases);
    return 0;
}
