// Multi-group batches: persistent launches that shade every point of many
// shader groups (host side of the emitted oslmi_batch_main kernels).
//
// The reference shades one group per BatchedExecutor<W>::execute call
// (src/include/OSL/oslexec.h:1000-1004); a renderer with many materials
// calls it group after group. On a 256-CU GPU that loop idles the machine
// whenever a group has fewer points than it takes to fill it, and pays one
// kernel's ramp-up and drain per group. A batch instead compiles the
// groups' per-point code into a few persistent kernels ("parts", up to
// groups_per_kernel groups each: one kernel over all groups would cost
// minutes of compile time, and its register budget is the maximum over
// every group, so small parts keep occupancy close to the groups' own).
// Each group's points are cut into tiles of block * ppt points; a part's
// tile list keeps each group's tiles together (group-sorted, so
// neighbouring workgroups run the same code) and is ordered either as given
// or by measured cost per tile, most expensive first (longest-processing-
// time order, which shortens the tail). A persistent grid of blocks_per_cu
// workgroups per CU pulls tiles from a device counter.
//
// The launch shape is tunable (block, ppt, blocks_per_cu, blackbody table
// in LDS or not, tile order, number of streams the kernels overlap on);
// oslmi_batch_autotune times candidates on the bound work and keeps the
// fastest. oslmi_batch_capture records the launches into a hipGraph; every
// counter resets itself at the end of its launch, so replays need no
// memset node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "oslmi.h"
#include "oslmi_color.h"
#include "oslmi_group.h"
#include "oslmi_internal.h"
#include "oslmi_tiles.h"

namespace oslmi {
namespace {

struct Part {
    std::vector<int> items;  // batch items (= group indices) of this kernel
    std::string source;
    std::vector<char> code;
    int rc = OSLMI_OK;
    std::string err;
    double seconds     = 0.0;
    hipModule_t mod    = nullptr;
    hipFunction_t fn   = nullptr;
    int max_blocks[3]  = { 0, 0, 0 };  // resident workgroups per CU at block 64/128/256
    int first_tile = 0, ntiles = 0;
};

int block_slot(int block) { return block == 64 ? 0 : block == 128 ? 1 : 2; }

struct Batch {
    ShadingSys* ss = nullptr;
    std::vector<Group*> groups;
    std::vector<Part> parts;
    double jit_seconds = 0.0;
    bool any_bb       = false;
    int num_cu        = 0;
    // bound work
    bool bound = false;
    std::vector<GroupArgs> hargs;
    GroupArgs* dargs   = nullptr;
    size_t dargs_cap   = 0;
    std::vector<BatchTile> tiles;
    std::vector<int> item_first, item_count;  // tile range of every item
    BatchTile* dtiles  = nullptr;
    size_t dtiles_cap  = 0;
    int32_t* dcounter  = nullptr;  // one per item (per-group launches), one per part
    size_t counter_cap = 0;
    const float* bb    = nullptr;
    std::vector<double> cost;  // measured ms per tile per item (order 1)
    oslmi_batch_config cfg{ 256, 1, 8, 1, 0, 1 };
    // side streams the kernels are spread over (cfg.streams > 1), with the
    // fork event and one join event per side stream
    std::vector<hipStream_t> side;
    hipEvent_t fork_ev = nullptr;
    std::vector<hipEvent_t> join_ev;
    // host staging of the no-alloc entry point (argument blocks, tiles)
    std::vector<GroupArgs> na_args;
    std::vector<BatchTile> na_tiles;
    // captured launches
    bool in_capture          = false;
    hipGraph_t graph         = nullptr;
    hipGraphExec_t gexec     = nullptr;
    hipStream_t graph_stream = nullptr;

    ~Batch()
    {
        drop_graph();
        for (hipStream_t s : side)
            (void)hipStreamDestroy(s);
        for (hipEvent_t e : join_ev)
            (void)hipEventDestroy(e);
        if (fork_ev)
            (void)hipEventDestroy(fork_ev);
        if (dargs)
            (void)hipFree(dargs);
        if (dtiles)
            (void)hipFree(dtiles);
        if (dcounter)
            (void)hipFree(dcounter);
        for (Part& p : parts)
            if (p.mod)
                (void)hipModuleUnload(p.mod);
    }

    void drop_graph()
    {
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (graph)
            (void)hipGraphDestroy(graph);
        gexec        = nullptr;
        graph        = nullptr;
        graph_stream = nullptr;
    }

    int load()
    {
        int dev = 0;
        if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
            return rc;
        if (int rc = check_hip(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev),
                               "hipDeviceGetAttribute"))
            return rc;
        for (Part& p : parts) {
            if (int rc = check_hip(hipModuleLoadData(&p.mod, p.code.data()), "hipModuleLoadData(batch)"))
                return rc;
            if (int rc = check_hip(hipModuleGetFunction(&p.fn, p.mod, "oslmi_batch_main"),
                                   "hipModuleGetFunction(oslmi_batch_main)"))
                return rc;
            for (int block : { 64, 128, 256 }) {
                int nb = 0;
                if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, p.fn, block, 0) != hipSuccess || nb < 1)
                    nb = 1;
                p.max_blocks[block_slot(block)] = nb;
            }
        }
        return OSLMI_OK;
    }

    // Tile list for the current config: each item's points in tiles of
    // block * ppt; within a part, items in the configured order.
    int build_tiles()
    {
        std::vector<int64_t> counts(hargs.size());
        for (size_t k = 0; k < hargs.size(); ++k)
            counts[k] = hargs[k].n;
        std::vector<int> part_first, part_ntiles;
        if (int rc = tile_list(counts, tiles, item_first, item_count, part_first, part_ntiles))
            return rc;
        for (size_t q = 0; q < parts.size(); ++q) {
            parts[q].first_tile = part_first[q];
            parts[q].ntiles     = part_ntiles[q];
        }
        if (tiles.size() > dtiles_cap) {
            if (dtiles)
                (void)hipFree(dtiles);
            dtiles     = nullptr;
            dtiles_cap = 0;
            if (int rc = check_hip(hipMalloc((void**)&dtiles, tiles.size() * sizeof(BatchTile)),
                                   "hipMalloc(batch tiles)"))
                return rc;
            dtiles_cap = tiles.size();
        }
        if (!tiles.empty())
            if (int rc = check_hip(hipMemcpy(dtiles, tiles.data(), tiles.size() * sizeof(BatchTile),
                                             hipMemcpyHostToDevice),
                                   "hipMemcpy(batch tiles)"))
                return rc;
        return OSLMI_OK;
    }

    // The tile list of items with `counts` points under the current config
    // (no device work; oslmi_tiles.h).
    int tile_list(const std::vector<int64_t>& counts, std::vector<BatchTile>& out,
                  std::vector<int>& first, std::vector<int>& count, std::vector<int>& part_first,
                  std::vector<int>& part_ntiles) const
    {
        std::vector<std::vector<int>> part_items(parts.size());
        for (size_t q = 0; q < parts.size(); ++q)
            part_items[q] = parts[q].items;
        TileLayout lay;
        lay.tiles.swap(out);
        const TileListStatus st =
            oslmi::tile_list(counts, part_items, (int64_t)cfg.block * cfg.ppt, cfg.order == 1, cost, lay);
        if (st == kTilesTooMany)
            return set_error(OSLMI_ERR_INVALID, "oslmi_batch: too many tiles");
        if (st != kTilesOk)
            return set_error(OSLMI_ERR_INVALID, "oslmi_batch: bad tile list input");
        out.swap(lay.tiles);
        first.swap(lay.item_first);
        count.swap(lay.item_count);
        part_first.swap(lay.part_first);
        part_ntiles.swap(lay.part_ntiles);
        return OSLMI_OK;
    }

    int part_of(int item) const
    {
        for (size_t p = 0; p < parts.size(); ++p)
            for (int k : parts[p].items)
                if (k == item)
                    return (int)p;
        return -1;
    }

    // One launch of part p's kernel over tiles [first, first + count) with
    // counter slot c.
    int launch_range(const Part& p, int first, int count, int c, hipStream_t stream)
    {
        return launch_range_on(p, first, count, dargs, dtiles, dcounter + c, bb, stream);
    }
    int launch_range_on(const Part& p, int first, int count, const GroupArgs* d_args,
                        const BatchTile* d_tiles, int32_t* d_counter, const float* d_bb,
                        hipStream_t stream)
    {
        if (count <= 0)
            return OSLMI_OK;
        int64_t grid = (int64_t)std::min(cfg.blocks_per_cu, p.max_blocks[block_slot(cfg.block)]) * num_cu;
        grid         = std::max<int64_t>(1, std::min<int64_t>(grid, count));
        BatchArgs B;
        memset(&B, 0, sizeof(B));
        B.args       = d_args;
        B.tiles      = d_tiles;
        B.tile_begin = first;
        B.tile_end   = first + count;
        B.counter    = d_counter;
        B.ppt        = cfg.ppt;
        B.bb_lds     = cfg.bb_lds;
        B.bb         = d_bb;
        void* args[] = { &B };
        return check_hip(hipModuleLaunchKernel(p.fn, (unsigned)grid, 1, 1, (unsigned)cfg.block, 1, 1, 0,
                                               stream, args, nullptr),
                         "hipModuleLaunchKernel(oslmi_batch_main)");
    }

    int ensure_side_streams(int ns)
    {
        if (!fork_ev)
            if (int rc = check_hip(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "hipEventCreate"))
                return rc;
        while ((int)side.size() < ns) {
            hipStream_t s = nullptr;
            hipEvent_t e  = nullptr;
            // default priority, non-blocking with respect to the null stream
            if (int rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))
                return rc;
            side.push_back(s);
            if (int rc = check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"))
                return rc;
            join_ev.push_back(e);
        }
        return OSLMI_OK;
    }

    // Every part's kernel over its tiles. With cfg.streams > 1 the kernels
    // are dealt round-robin to side streams that wait for the caller's
    // stream and are waited for by it. Under graph capture (in_capture) the
    // kernels go to the capturing stream one after another: a captured
    // fork/join replayed slower than eager launches, and a run that tuned
    // many such graphs ended in a host crash whose cause was not found, so
    // captured graphs keep to one stream.
    int launch_all(hipStream_t stream)
    {
        const int ni = (int)hargs.size();
        const int ns = in_capture ? 1 : std::min<int>(cfg.streams, (int)parts.size());
        if (ns <= 1) {
            for (size_t q = 0; q < parts.size(); ++q)
                if (int rc = launch_range(parts[q], parts[q].first_tile, parts[q].ntiles, ni + (int)q, stream))
                    return rc;
            return OSLMI_OK;
        }
        if (int rc = ensure_side_streams(ns))
            return rc;
        if (int rc = check_hip(hipEventRecord(fork_ev, stream), "hipEventRecord(fork)"))
            return rc;
        for (int s = 0; s < ns; ++s)
            if (int rc = check_hip(hipStreamWaitEvent(side[s], fork_ev, 0), "hipStreamWaitEvent(fork)"))
                return rc;
        for (size_t q = 0; q < parts.size(); ++q)
            if (int rc = launch_range(parts[q], parts[q].first_tile, parts[q].ntiles, ni + (int)q,
                                      side[q % ns]))
                return rc;
        for (int s = 0; s < ns; ++s) {
            if (int rc = check_hip(hipEventRecord(join_ev[s], side[s]), "hipEventRecord(join)"))
                return rc;
            if (int rc = check_hip(hipStreamWaitEvent(stream, join_ev[s], 0), "hipStreamWaitEvent(join)"))
                return rc;
        }
        return OSLMI_OK;
    }

    int launch_item(int k, hipStream_t stream)
    {
        return launch_range(parts[part_of(k)], item_first[k], item_count[k], k, stream);
    }

    // Records launch_all on `st` into a graph (replaces any previous one).
    int capture(hipStream_t st)
    {
        drop_graph();
        if (int rc = check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                               "hipStreamBeginCapture"))
            return rc;
        in_capture   = true;
        int lrc      = launch_all(st);
        in_capture   = false;
        hipGraph_t g = nullptr;
        int erc      = check_hip(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
        if (lrc || erc) {
            if (g)
                (void)hipGraphDestroy(g);
            return lrc ? lrc : erc;
        }
        graph = g;
        if (int rc = check_hip(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0), "hipGraphInstantiate")) {
            drop_graph();
            return rc;
        }
        graph_stream = st;
        return OSLMI_OK;
    }

    // Named transforms and renderer attributes changed since the last
    // upload reach the table the kernels read (the table's address never
    // changes, so captured graphs see the new values too). Skipped while
    // the stream is being captured: the upload waits for the device.
    int sync_renderer_table(hipStream_t stream)
    {
        bool any_xf = false;
        for (const Group* g : groups)
            any_xf = any_xf || g->uses_xforms;
        if (!any_xf)
            return OSLMI_OK;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (int rc = check_hip(hipStreamIsCapturing(stream, &cs), "hipStreamIsCapturing"))
            return rc;
        if (cs != hipStreamCaptureStatusNone)
            return OSLMI_OK;
        return xform_table_sync(*ss);
    }

    int launch_per_group(hipStream_t stream)
    {
        for (int k = 0; k < (int)hargs.size(); ++k)
            if (int rc = launch_item(k, stream))
                return rc;
        return OSLMI_OK;
    }
};

int validate(const oslmi_batch_config& c)
{
    if (c.block != 64 && c.block != 128 && c.block != 256)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: block must be 64, 128 or 256");
    if (c.ppt < 1 || c.ppt > 8)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: ppt must be in 1..8");
    if (c.blocks_per_cu < 1 || c.blocks_per_cu > 32)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: blocks_per_cu must be in 1..32");
    if (c.bb_lds != 0 && c.bb_lds != 1)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: bb_lds must be 0 or 1");
    if (c.order != 0 && c.order != 1)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: order must be 0 or 1");
    if (c.streams < 1 || c.streams > 4)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch: streams must be in 1..4");
    return OSLMI_OK;
}

// fn(0 .. n-1) on up to 16 host threads
template <class F>
void parallel_for(int n, F fn)
{
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next { 0 };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(hw, n); ++t)
        pool.emplace_back([&] {
            for (int k; (k = next.fetch_add(1)) < n;)
                fn(k);
        });
    for (std::thread& t : pool)
        t.join();
}

// VGPR + AGPR count of the (single) kernel in a gfx950 code object, read
// from its metadata note (msgpack map entries ".vgpr_count" and
// ".agpr_count"); 0 when absent.
int code_object_registers(const std::vector<char>& code)
{
    auto field = [&](const char* key) -> int {
        const size_t kl = strlen(key);
        for (size_t i = 0; i + kl + 3 < code.size(); ++i) {
            if (memcmp(code.data() + i, key, kl) != 0)
                continue;
            const unsigned char* v = (const unsigned char*)code.data() + i + kl;
            if (v[0] < 0x80)
                return v[0];  // positive fixint
            if (v[0] == 0xcc)
                return v[1];  // uint8
            if (v[0] == 0xcd)
                return (v[1] << 8) | v[2];  // uint16, big endian
        }
        return 0;
    };
    return field(".vgpr_count") + field(".agpr_count");
}

// mean ms of `reps` calls of f on stream (after one untimed call)
template <class F>
int time_ms(hipStream_t stream, int reps, F f, double& ms)
{
    hipEvent_t e0, e1;
    if (int rc = check_hip(hipEventCreate(&e0), "hipEventCreate"))
        return rc;
    if (int rc = check_hip(hipEventCreate(&e1), "hipEventCreate")) {
        (void)hipEventDestroy(e0);
        return rc;
    }
    int rc = f();
    if (!rc)
        rc = check_hip(hipEventRecord(e0, stream), "hipEventRecord");
    for (int r = 0; r < reps && !rc; ++r)
        rc = f();
    if (!rc)
        rc = check_hip(hipEventRecord(e1, stream), "hipEventRecord");
    if (!rc)
        rc = check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
    float t = 0.f;
    if (!rc)
        rc = check_hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    ms = (double)t / std::max(1, reps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace
}  // namespace oslmi

using namespace oslmi;

struct oslmi_batch {
    Batch b;
};

extern "C" {

oslmi_batch* oslmi_batch_create(oslmi_shadingsys* s, int ngroups, oslmi_group* const* groups,
                                int groups_per_kernel, int load)
{
    if (!s || ngroups <= 0 || !groups) {
        set_error(OSLMI_ERR_INVALID, "oslmi_batch_create: need a shading system and >= 1 group");
        return nullptr;
    }
    if (groups_per_kernel < 0) {
        set_error(OSLMI_ERR_INVALID, "oslmi_batch_create: groups_per_kernel < 0");
        return nullptr;
    }
    const int per = groups_per_kernel ? groups_per_kernel : 8;
    oslmi_batch* h = new oslmi_batch;
    Batch& b       = h->b;
    b.ss           = &s->ss;
    for (int k = 0; k < ngroups; ++k) {
        if (!groups[k] || !groups[k]->g->ended) {
            set_error(OSLMI_ERR_INVALID, "oslmi_batch_create: group " + std::to_string(k) + " is not ended");
            delete h;
            return nullptr;
        }
        for (int j = 0; j < k; ++j)
            if (groups[j]->g == groups[k]->g) {
                set_error(OSLMI_ERR_INVALID, "oslmi_batch_create: group " + std::to_string(k) + " listed twice");
                delete h;
                return nullptr;
            }
        b.groups.push_back(groups[k]->g);
    }
    // Register need of every group: its own kernel's VGPR + AGPR count
    // (compiled through the code-object cache). Groups of similar need
    // share a kernel, whose register budget is the maximum over its groups,
    // so light groups keep their occupancy.
    std::vector<int> regs(ngroups, 0);
    {
        std::vector<int> rcs(ngroups, OSLMI_OK);
        std::vector<std::string> errs(ngroups);
        parallel_for(ngroups, [&](int k) {
            Group& g = *b.groups[k];
            std::string src, err;
            std::vector<char> code;
            bool ok;
            {
                std::lock_guard<std::mutex> lock(g.mu);
                ok = emit_group(s->ss, g, src, err);
            }
            if (!ok) {
                rcs[k]  = OSLMI_ERR_COMPILE;
                errs[k] = "emit " + g.name + ": " + err;
                return;
            }
            if ((rcs[k] = build_code_object(s->ss, src, g.name, code, nullptr)) != OSLMI_OK) {
                errs[k] = oslmi_last_error();
                return;
            }
            regs[k] = code_object_registers(code);
        });
        for (int k = 0; k < ngroups; ++k)
            if (rcs[k]) {
                set_error(rcs[k], "oslmi_batch_create: " + errs[k]);
                delete h;
                return nullptr;
            }
    }
    std::vector<int> order(ngroups);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return regs[a] < regs[c]; });
    // emit (the emitter records each group's outputs, under its lock)
    for (int first = 0; first < ngroups; first += per) {
        Part p;
        std::vector<Group*> gs;
        for (int j = first; j < std::min(ngroups, first + per); ++j) {
            p.items.push_back(order[j]);
            gs.push_back(b.groups[order[j]]);
        }
        std::string err;
        bool ok;
        {
            std::vector<std::unique_lock<std::mutex>> locks;
            for (Group* g : gs)
                locks.emplace_back(g->mu);
            ok = emit_batch(s->ss, gs, p.items, p.source, err);
        }
        if (!ok) {
            set_error(OSLMI_ERR_COMPILE, "oslmi_batch_create: " + err);
            delete h;
            return nullptr;
        }
        b.any_bb = b.any_bb || p.source.find("s_bbl") != std::string::npos;
        b.parts.push_back(std::move(p));
    }
    // compile the parts side by side (distinct hiprtc programs)
    {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        size_t next = 0;
        while (next < b.parts.size()) {
            std::vector<std::thread> pool;
            for (unsigned t = 0; t < hw && next < b.parts.size(); ++t, ++next) {
                Part* p = &b.parts[next];
                pool.emplace_back([p, s, next] {
                    p->rc = build_code_object(s->ss, p->source, "oslmi_batch" + std::to_string(next), p->code,
                                              &p->seconds);
                    if (p->rc)
                        p->err = oslmi_last_error();
                });
            }
            for (std::thread& t : pool)
                t.join();
        }
        for (Part& p : b.parts) {
            b.jit_seconds = std::max(b.jit_seconds, p.seconds);
            if (p.rc) {
                set_error(p.rc, p.err);
                delete h;
                return nullptr;
            }
        }
    }
    if (load && b.load()) {
        delete h;
        return nullptr;
    }
    return h;
}

void oslmi_batch_destroy(oslmi_batch* h) { delete h; }

double oslmi_batch_jit_seconds(oslmi_batch* h) { return h ? h->b.jit_seconds : 0.0; }

int oslmi_batch_kernels(oslmi_batch* h) { return h ? (int)h->b.parts.size() : 0; }

int oslmi_batch_bind(oslmi_batch* h, int nitems, const oslmi_batch_item* items)
{
    if (!h || !items)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_bind: null argument");
    Batch& b = h->b;
    if (b.parts.empty() || !b.parts[0].fn)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_bind: batch was created without loading");
    if (nitems != (int)b.groups.size())
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_bind: need one item per group");
    std::vector<GroupArgs> args(nitems);
    for (int k = 0; k < nitems; ++k) {
        const oslmi_batch_item& it = items[k];
        std::string who = "oslmi_batch_bind item " + std::to_string(k);
        if (int rc = make_group_args(*b.ss, *b.groups[k], &it.globals, it.d_outputs, it.d_wo,
                                     it.d_wi, it.d_eval, who.c_str(), args[k]))
            return rc;
        if (it.globals.n > ((int64_t)1 << 40))
            return set_error(OSLMI_ERR_INVALID, who + ": n too large");
    }
    if (b.any_bb) {
        const ColorSysDev* cs = nullptr;
        if (int rc = device_colorsys(&cs))
            return rc;
        b.bb = (const float*)((const char*)cs + offsetof(ColorSysDev, bb));
    }
    // launches still in flight may read the buffers about to change
    if (int rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return rc;
    b.drop_graph();
    b.bound = false;
    if ((size_t)nitems > b.dargs_cap) {
        if (b.dargs)
            (void)hipFree(b.dargs);
        b.dargs     = nullptr;
        b.dargs_cap = 0;
        if (int rc = check_hip(hipMalloc((void**)&b.dargs, nitems * sizeof(GroupArgs)), "hipMalloc(batch args)"))
            return rc;
        b.dargs_cap = nitems;
    }
    if (int rc = check_hip(hipMemcpy(b.dargs, args.data(), nitems * sizeof(GroupArgs), hipMemcpyHostToDevice),
                           "hipMemcpy(batch args)"))
        return rc;
    const size_t ncount = (size_t)nitems + b.parts.size();
    if (ncount > b.counter_cap) {
        if (b.dcounter)
            (void)hipFree(b.dcounter);
        b.dcounter    = nullptr;
        b.counter_cap = 0;
        if (int rc = check_hip(hipMalloc((void**)&b.dcounter, ncount * sizeof(int32_t)), "hipMalloc(batch counters)"))
            return rc;
        b.counter_cap = ncount;
    }
    if (int rc = check_hip(hipMemset(b.dcounter, 0, ncount * sizeof(int32_t)), "hipMemset(batch counters)"))
        return rc;
    b.hargs = args;
    b.cost.clear();
    if (int rc = b.build_tiles())
        return rc;
    b.bound = true;
    return OSLMI_OK;
}

int oslmi_batch_set_config(oslmi_batch* h, const oslmi_batch_config* cfg)
{
    if (!h || !cfg)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_set_config: null argument");
    if (int rc = validate(*cfg))
        return rc;
    Batch& b = h->b;
    if (!b.bound) {
        b.cfg = *cfg;
        return OSLMI_OK;
    }
    if (int rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return rc;
    b.drop_graph();
    b.cfg = *cfg;
    return b.build_tiles();
}

int oslmi_batch_get_config(oslmi_batch* h, oslmi_batch_config* cfg)
{
    if (!h || !cfg)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_get_config: null argument");
    *cfg = h->b.cfg;
    return OSLMI_OK;
}

int oslmi_batch_launch(oslmi_batch* h, void* stream)
{
    if (!h || !h->b.bound)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch: batch not bound");
    Batch& b = h->b;
    if (int rc = b.sync_renderer_table((hipStream_t)stream))
        return rc;
    if (b.gexec && b.graph_stream == (hipStream_t)stream)
        return check_hip(hipGraphLaunch(b.gexec, (hipStream_t)stream), "hipGraphLaunch");
    return b.launch_all((hipStream_t)stream);
}

int oslmi_batch_launch_per_group(oslmi_batch* h, void* stream)
{
    if (!h || !h->b.bound)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch_per_group: batch not bound");
    if (int rc = h->b.sync_renderer_table((hipStream_t)stream))
        return rc;
    return h->b.launch_per_group((hipStream_t)stream);
}

int oslmi_batch_capture(oslmi_batch* h, void* stream)
{
    if (!h || !h->b.bound)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_capture: batch not bound");
    if (!stream)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_capture: capture needs a created stream, not the null stream");
    if (int rc = h->b.sync_renderer_table((hipStream_t)stream))
        return rc;
    return h->b.capture((hipStream_t)stream);
}

int oslmi_batch_autotune(oslmi_batch* h, void* stream, int reps, int graph,
                         oslmi_batch_config* best, double* best_ms)
{
    if (graph && !stream)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_autotune: graph tuning needs a created stream");
    if (!h || !h->b.bound)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_autotune: batch not bound");
    Batch& b       = h->b;
    hipStream_t st = (hipStream_t)stream;
    reps           = std::max(1, reps);
    if (int rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return rc;
    b.drop_graph();
    // cost of every item alone (block 256, one point per work-item), as ms
    // per tile with as many of its tiles resident as fit
    {
        b.cfg = oslmi_batch_config{ 256, 1, 32, 1, 0, 1 };
        if (int rc = b.build_tiles())
            return rc;
        const int ni = (int)b.hargs.size();
        std::vector<double> cost(ni, 0.0);
        for (int k = 0; k < ni; ++k) {
            if (b.item_count[k] == 0)
                continue;
            const Part& p       = b.parts[b.part_of(k)];
            const int64_t slots = (int64_t)p.max_blocks[2] * b.num_cu;
            double ms           = 0.0;
            if (int rc = time_ms(st, 1, [&] { return b.launch_item(k, st); }, ms))
                return rc;
            cost[k] = ms * (double)std::min<int64_t>(b.item_count[k], slots) / b.item_count[k];
        }
        b.cost = cost;
    }
    // two stages: the launch geometry (block, ppt, blocks_per_cu) with the
    // LDS table, cost order and all side streams; then those three
    // switches with the winning geometry
    // (captured graphs keep to one stream, see launch_all)
    const int max_streams = graph ? 1 : std::min<int>(4, (int)b.parts.size());
    oslmi_batch_config win{ 256, 1, 8, 1, 1, max_streams };
    double win_ms = 1e30;
    auto trial = [&](const oslmi_batch_config& c) -> int {
        b.cfg = c;
        if (int rc = b.build_tiles())
            return rc;
        double ms = 0.0;
        if (graph) {
            // timed the way it will run: as a replayed graph
            if (int rc = b.capture(st))
                return rc;
            if (int rc = time_ms(st, reps, [&] {
                    return check_hip(hipGraphLaunch(b.gexec, st), "hipGraphLaunch");
                }, ms))
                return rc;
            b.drop_graph();
        } else if (int rc = time_ms(st, reps, [&] { return b.launch_all(st); }, ms)) {
            return rc;
        }
        if (ms < win_ms) {
            win_ms = ms;
            win    = c;
        }
        return OSLMI_OK;
    };
    for (int block : { 64, 128, 256 }) {
        int maxb = 1;
        for (const Part& p : b.parts)
            maxb = std::max(maxb, p.max_blocks[block_slot(block)]);
        maxb = std::min(32, maxb);
        std::vector<int> bpcs;
        for (int v = 1; v < maxb; v *= 2)
            bpcs.push_back(v);
        bpcs.push_back(maxb);
        for (int ppt : { 1, 2, 4 })
            for (int bpc : bpcs)
                if (int rc = trial(oslmi_batch_config{ block, ppt, bpc, 1, 1, max_streams }))
                    return rc;
    }
    const oslmi_batch_config geo = win;
    for (int lds = 0; lds <= (b.any_bb ? 1 : 0); ++lds)
        for (int order = 0; order <= 1; ++order)
            for (int ns = 1; ns <= max_streams; ns *= 2)
                if (int rc = trial(oslmi_batch_config{ geo.block, geo.ppt, geo.blocks_per_cu,
                                                       b.any_bb ? lds : 1, order, ns }))
                    return rc;
    b.cfg = win;
    if (int rc = b.build_tiles())
        return rc;
    if (graph)
        if (int rc = b.capture(st))
            return rc;
    if (best)
        *best = win;
    if (best_ms)
        *best_ms = win_ms;
    return OSLMI_OK;
}

static_assert(sizeof(GroupArgs) <= kNoAllocArgsSlot, "no-alloc workspace: argument block slot too small");
static_assert(kNoAllocArgsOffset >= kNoAllocMaxCounters * sizeof(int32_t), "workspace layout");
static_assert(kNoAllocTilesOffset + kNoAllocMaxTiles * sizeof(BatchTile) <= kNoAllocScratchBytes,
              "workspace layout");

int oslmi_batch_launch_noalloc(oslmi_batch* h, int nitems, const oslmi_batch_item* items,
                               void* stream)
{
    if (!h || (nitems > 0 && !items))
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch_noalloc: null argument");
    Batch& b       = h->b;
    hipStream_t st = (hipStream_t)stream;
    if (nitems != (int)b.groups.size())
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch_noalloc: nitems must equal the batch's group count");
    if (b.parts.empty() || !b.parts[0].fn)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch_noalloc: batch kernels not loaded");
    const int ncount = nitems + (int)b.parts.size();
    if (nitems > kNoAllocMaxItems || ncount > kNoAllocMaxCounters)
        return set_error(OSLMI_ERR_INVALID, "oslmi_batch_launch_noalloc: more than "
                                                + std::to_string(kNoAllocMaxItems) + " items");
    unsigned char* base = nullptr;
    if (int rc = noalloc_scratch(&base))
        return rc;
    // argument blocks (tables must already be resident: nothing is allocated)
    std::vector<int64_t> counts(nitems);
    b.na_args.resize(nitems);
    for (int k = 0; k < nitems; ++k) {
        const oslmi_batch_item& it = items[k];
        std::string who = "oslmi_batch_launch_noalloc item " + std::to_string(k);
        if (int rc = make_group_args(*b.ss, *b.groups[k], &it.globals, it.d_outputs, it.d_wo,
                                     it.d_wi, it.d_eval, who.c_str(), b.na_args[k], false))
            return rc;
        if (it.globals.n > ((int64_t)1 << 40))
            return set_error(OSLMI_ERR_INVALID, who + ": n too large");
        counts[k] = it.globals.n;
    }
    const float* bbp = nullptr;
    if (b.any_bb) {
        const ColorSysDev* cs = nullptr;
        if (int rc = device_colorsys(&cs, false))
            return rc;
        bbp = (const float*)((const char*)cs + offsetof(ColorSysDev, bb));
    }
    std::vector<int> first, count, part_first, part_ntiles;
    if (int rc = b.tile_list(counts, b.na_tiles, first, count, part_first, part_ntiles))
        return rc;
    if ((int64_t)b.na_tiles.size() > kNoAllocMaxTiles)
        return set_error(OSLMI_ERR_INVALID,
                         "oslmi_batch_launch_noalloc: " + std::to_string(b.na_tiles.size())
                             + " tiles exceed the workspace's " + std::to_string(kNoAllocMaxTiles)
                             + " (larger block or ppt, or fewer points per call)");
    int32_t* d_counter   = (int32_t*)base;
    GroupArgs* d_args    = (GroupArgs*)(base + kNoAllocArgsOffset);
    BatchTile* d_tiles   = (BatchTile*)(base + kNoAllocTilesOffset);
    if (nitems > 0)
        if (int rc = check_hip(hipMemcpyAsync(d_args, b.na_args.data(), nitems * sizeof(GroupArgs),
                                              hipMemcpyHostToDevice, st),
                               "hipMemcpyAsync(no-alloc args)"))
            return rc;
    if (!b.na_tiles.empty())
        if (int rc = check_hip(hipMemcpyAsync(d_tiles, b.na_tiles.data(),
                                              b.na_tiles.size() * sizeof(BatchTile),
                                              hipMemcpyHostToDevice, st),
                               "hipMemcpyAsync(no-alloc tiles)"))
            return rc;
    // counters start at zero whatever an earlier call left
    if (int rc = noalloc_clear(d_counter, ncount, st))
        return rc;
    // the host staging buffers are reused by the next call
    if (int rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return rc;
    for (size_t q = 0; q < b.parts.size(); ++q)
        if (int rc = b.launch_range_on(b.parts[q], part_first[q], part_ntiles[q], d_args, d_tiles,
                                       d_counter + nitems + (int)q, bbp, st))
            return rc;
    return OSLMI_OK;
}

}  // extern "C"
