// dt_engine.cpp -- continuous batching of solves of uneven sizes.
//
// Every admitted instance (a volume with its parameters) owns a slot: an
// Inst record and a control block in slot-indexed device arrays, and a range
// of a device pool holding its solver state.  Instances are grouped by
// algorithm into lanes; a lane has its own stream, tile table and chunk
// pipeline, so the passes of one algorithm never wait for the other's.  A
// pass of the gradient lane is one launch of the fused gradient kernel over
// the tile table of all its instances, a pass of the greedy lane the fused
// greedy kernel over all of its; each block starts by advancing its
// instance's controller on the previous pass's sums (the blocks of an
// instance all reach the same state), so step lengths, line searches and
// stopping tests stay per instance and never wait for the host.  Per lane the host enqueues
// chunks of passes (two in flight).  Each controller stores its instance's
// completion word (and, once done, its control block) in host-coherent
// memory, so only pass kernels enter a lane's stream: the host follows the
// words, retires the instances that have stopped (result copy, slot and
// memory freed) and admits new ones between chunks; the tile table is
// rebuilt whenever the lane's set of instances changes.  A chunk can be
// replayed from a captured hipGraph.
//
// An instance's result does not depend on what else is in the batch, nor on
// its tiling: its sums are exact (fixed point) and its controller is its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "disptools_hip.h"
#include "dt_host.hpp"
#include "dt_match.hpp"

namespace dt {

namespace {

using Clock = std::chrono::steady_clock;

enum SlotState { SLOT_FREE = 0, SLOT_ACTIVE = 1, SLOT_RETIRING = 2 };

struct Slot {
    int state = SLOT_FREE;
    long long ticket = -1;
    int lane = 0;                 // DT_GRADIENT, DT_GREEDY or DT_MATCHING
    Plan P;
    dt_params params{};
    int host_io = 0;
    float *out_field = nullptr;   // where the result goes (caller's buffer)
    size_t pool_off = 0, pool_bytes = 0;
    long long launched = 0;       // passes launched over this instance
    int failed = 0;
    Ctrl final_ctrl{};
    MCtrl final_mctrl{};          // matching instances
    float *dev_field = nullptr;   // matching: the field on the device (positions -> field at the end)
    float spacing[3] = {1.0f, 1.0f, 1.0f};
    Clock::time_point t_submit;
    hipEvent_t ev_ready = nullptr;   // inputs uploaded (host_io)
    hipEvent_t ev_stop = nullptr;    // the lane's stream reached the retirement point
    hipEvent_t ev_done = nullptr;    // result in place
};

// An instance's share of a chunk: the chunk has run for it once its
// completion word counts `passes` passes or reports the end of its solve.
struct Need {
    int slot;
    long long ticket;
    long long passes;
};

struct Chunk {
    long long seq;
    long long order;  // engine-wide enqueue order
    std::vector<Need> needs;
};

// The instances of one algorithm and their pass pipeline.
struct Lane {
    int kind = DT_GRADIENT;
    hipStream_t st = nullptr;     // passes, instance uploads, result copies
    hipStream_t cap = nullptr;    // graph capture
    bool own_st = true;
    Tile *d_tiles = nullptr;
    size_t d_cap = 0;
    Tile *h_tiles[2] = {nullptr, nullptr};   // pinned staging, alternating per chunk
    hipEvent_t ev_tiles[2] = {nullptr, nullptr};   // upload from the staging buffer done
    bool tiles_pending[2] = {false, false};
    size_t h_cap = 0;
    int ntiles = 0;
    bool dirty = true;
    hipGraphExec_t graph = nullptr;
    int graph_ntiles = -1;
    Tile *graph_tiles = nullptr;
    std::deque<Chunk> inflight;
    long long seq = 0;
    int n_active = 0;
};

// First-fit allocator over [0, size) with coalescing.
struct Pool {
    std::map<size_t, size_t> free_;   // offset -> bytes
    size_t size = 0, used = 0, peak = 0;
    void init(size_t s)
    {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t bytes, size_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used += bytes;
                peak = std::max(peak, used);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr size_t POOL_GRAIN = (size_t)1 << 16;
constexpr int MAX_SLOTS = 4096;
constexpr int NLANES = 3;   // DT_GRADIENT, DT_GREEDY, DT_MATCHING

}  // namespace

}  // namespace dt

using namespace dt;

struct dt_engine {
    int device = 0;
    int precision = 32;
    int variant = 1;             // gradient pass variant
    int greedy_variant = 1;      // greedy pass variant
    int chunk = 2;
    int use_graphs = 0;
    int tz_policy = 0;
    int max_slots = 256;
    hipStream_t io = nullptr;    // host staging copies
    Lane lanes[NLANES];

    char *pool_mem = nullptr;
    Pool pool;

    Inst *d_insts = nullptr;
    Ctrl *d_ctrl = nullptr;            // two copies per slot (pass parity)
    Acc *d_acc = nullptr;              // ACC_SLOTS * ACC_COPIES per slot
    unsigned int *d_counter = nullptr; // matching instances' arrival counters
    Inst *h_insts = nullptr;     // pinned mirror, one per slot
    MInst *d_minsts = nullptr;   // matching instances, slot-indexed
    MCtrl *d_mctrl = nullptr;
    MInst *h_minsts = nullptr;   // pinned mirror
    MCtrl *h_mctrl0 = nullptr;   // pinned initial control blocks
    Ctrl *h_ctrl0 = nullptr;     // pinned initial control blocks (two copies per slot)
    int *h_zero = nullptr;       // pinned 0 (deactivates a stuck instance)
    // host-coherent, per slot: completion words, then final control blocks
    char *coherent = nullptr;
    char *coherent_dev = nullptr;
    size_t final_off = 0;
    size_t mfinal_off = 0;       // published matching control blocks
    bool status_words = true;    // the kernels publish them (DISPTOOLS_STATUS_WORD)
    std::chrono::steady_clock::time_point last_check{};   // of check_lanes

    std::vector<Slot> slots;
    std::vector<dt_engine_result> results;
    long long order = 0;
    long long next_ticket = 0;
    int n_active = 0, n_retiring = 0;
    dt_engine_stats stats{};
};

namespace dt {
namespace {

volatile unsigned long long *status_of(dt_engine *e, int s) { return (volatile unsigned long long *)e->coherent + s; }
Ctrl *final_of(dt_engine *e, int s) { return (Ctrl *)(e->coherent + e->final_off) + s; }
MCtrl *mfinal_of(dt_engine *e, int s) { return (MCtrl *)(e->coherent + e->mfinal_off) + s; }

// Tiles (blocks per pass) of the instance in slot s.
int slot_tiles(const dt_engine *e, int s)
{
    const Slot &sl = e->slots[s];
    return sl.lane == DT_MATCHING ? e->h_minsts[s].nblocks : sl.P.ntiles;
}

// Depth of balanced z-chunks of at most `most` planes (a multiple of 4), in
// whole groups of four planes.
int balanced_tz(int nz, int most)
{
    const int groups = (nz + 3) / 4;
    const int chunks = (groups + most / 4 - 1) / (most / 4);
    return std::min((groups + chunks - 1) / chunks * 4, (nz + 3) / 4 * 4);
}

// Depth at admission: fixed (tz_policy > 0), as a lone solve (-1), else
// balanced chunks of at most 16 planes (re-planned with the lane, below).
int engine_tz(const dt_engine *e, int nz)
{
    if (e->tz_policy > 0) return std::min(e->tz_policy, nz);
    if (e->tz_policy == -1) return 0;   // make_plan: as a lone solve
    return balanced_tz(nz, 16);
}

// Re-plans the z-chunk depth of every instance of a gradient or greedy lane
// (tz_policy 0) for the pass over the lane's current set: every instance is
// cut into balanced chunks of at most D planes, D chosen to minimise the
// estimated pass time.  A block of depth d costs d + O plane steps (O: the
// halo planes of the pipeline and the block's fixed work -- control step,
// sums -- in plane-step units); the CUs run the blocks in rounds, so a pass
// of up to two rounds of blocks of largest cost c takes ceil(blocks / CUs) * c,
// a longer one about ceil(W / (CUs * c)) * c for blocks of total cost W.
// Small D fills the CUs when few instances remain, large D saves halo planes
// when many do.  The results do not depend on the depths (exact sums), so an
// instance may be re-tiled between any two passes.
void retile_lane(dt_engine *e, Lane &L)
{
    const int cus = device_cus();
    const int O = L.kind == DT_GREEDY ? 7 : 6;
    int best_d = 16;
    double best_cost = 1e300;
    for (int D = 4; D <= 64; D += 4) {
        double work = 0.0;
        long long blocks = 0;
        int cmax = 0;
        for (int s = 0; s < (int)e->slots.size(); ++s) {
            const Slot &sl = e->slots[s];
            if (sl.state != SLOT_ACTIVE || sl.lane != L.kind) continue;
            const int d = balanced_tz(sl.P.nz, D);
            const long long tiles = (long long)sl.P.ntx * sl.P.nty * ((sl.P.nz + d - 1) / d);
            const int c = std::min(d, sl.P.nz) + O;
            work += (double)tiles * c;
            blocks += tiles;
            cmax = std::max(cmax, c);
        }
        if (cmax == 0) return;
        // a few rounds: whole rounds of the deepest block (257 blocks take
        // two rounds, not 1.004); many: the work spread over the CUs
        const double cost = blocks <= 2ll * cus ? (double)((blocks + cus - 1) / cus) * cmax
                                                : std::ceil(work / ((double)cus * cmax)) * cmax;
        if (cost < best_cost * 0.999) { best_cost = cost; best_d = D; }
    }
    for (int s = 0; s < (int)e->slots.size(); ++s) {
        Slot &sl = e->slots[s];
        if (sl.state != SLOT_ACTIVE || sl.lane != L.kind) continue;
        sl.P.tz = balanced_tz(sl.P.nz, best_d);
        sl.P.ntz = (sl.P.nz + sl.P.tz - 1) / sl.P.tz;
        sl.P.ntiles = sl.P.ntx * sl.P.nty * sl.P.ntz;
    }
}

Plan engine_plan(const dt_engine *e, int algorithm, int nx, int ny, int nz)
{
    return make_plan(algorithm, nx, ny, nz, e->precision, algorithm == DT_GRADIENT ? e->variant : e->greedy_variant,
                     engine_tz(e, nz));
}

size_t io_bytes(long long n)
{
    return align_up(3 * (size_t)n * sizeof(float), 256) + align_up((size_t)n * sizeof(float), 256) +
           align_up((size_t)n, 256);
}

size_t footprint(const dt_engine *e, const Plan &P, int host_io)
{
    (void)e;
    const size_t b = (host_io ? io_bytes(P.n) : 0) + scratch_bytes(P, 0, false);
    return align_up(b, POOL_GRAIN);
}

size_t match_footprint(int nx, int ny, int nz, int host_io)
{
    const long long n = (long long)nx * ny * nz;
    const size_t b = (host_io ? io_bytes(n) : 0) + match_scratch_bytes(nx, ny, nz);
    return align_up(b, POOL_GRAIN);
}

int ensure_tables(Lane &L, size_t need)
{
    if (L.d_cap < need) {
        // the current table may still be read by queued passes
        DT_HIP(hipStreamSynchronize(L.st));
        if (L.d_tiles) DT_HIP(hipFree(L.d_tiles));
        L.d_tiles = nullptr;
        const size_t cap = std::max<size_t>(need, 2 * L.d_cap);
        DT_HIP(hipMalloc(&L.d_tiles, cap * sizeof(Tile)));
        L.d_cap = cap;
    }
    if (L.h_cap < need) {
        DT_HIP(hipStreamSynchronize(L.st));
        L.tiles_pending[0] = L.tiles_pending[1] = false;
        const size_t cap = std::max<size_t>(need, 2 * L.h_cap);
        for (int s = 0; s < 2; ++s) {
            if (L.h_tiles[s]) DT_HIP(hipHostFree(L.h_tiles[s]));
            L.h_tiles[s] = nullptr;
            DT_HIP(hipHostMalloc((void **)&L.h_tiles[s], cap * sizeof(Tile), hipHostMallocDefault));
        }
        L.h_cap = cap;
    }
    return 0;
}

// Rebuilds and uploads the lane's tile table through staging buffer `stage`.
int rebuild_tables(dt_engine *e, Lane &L, int stage)
{
    if (e->tz_policy == 0 && L.kind != DT_MATCHING) retile_lane(e, L);
    std::vector<TileRun> seg;
    size_t count = 0;
    for (int s = 0; s < (int)e->slots.size(); ++s) {
        const Slot &sl = e->slots[s];
        if (sl.state != SLOT_ACTIVE || sl.lane != L.kind) continue;
        const int nt = slot_tiles(e, s);
        seg.push_back(sl.lane == DT_MATCHING ? TileRun{s, nt, 1, 1} : plan_run(s, sl.P));
        count += (size_t)nt;
    }
    if (ensure_tables(L, std::max<size_t>(count, 1))) return -1;
    std::vector<Tile> t;
    t.reserve(count);
    append_tiles_xcd(t, seg);
    L.ntiles = (int)t.size();
    if (!t.empty()) {
        // the staging buffer's previous upload must have been read
        if (L.tiles_pending[stage]) DT_HIP(hipEventSynchronize(L.ev_tiles[stage]));
        L.tiles_pending[stage] = false;
        std::memcpy(L.h_tiles[stage], t.data(), t.size() * sizeof(Tile));
        DT_HIP(hipMemcpyAsync(L.d_tiles, L.h_tiles[stage], t.size() * sizeof(Tile), hipMemcpyHostToDevice, L.st));
        DT_HIP(hipEventRecord(L.ev_tiles[stage], L.st));
        L.tiles_pending[stage] = true;
    }
    L.dirty = false;
    return 0;
}

int launch_passes(dt_engine *e, const Lane &L, hipStream_t s)
{
    // chunks have an even length, so every instance's passes alternate the
    // parity from its first one (instances join between chunks)
    for (int i = 0; i < e->chunk; ++i) {
        if (L.kind == DT_GRADIENT) {
            DT_HIP(launch_grad_pass(e->precision, e->variant, e->d_insts, L.d_tiles, L.ntiles, i & 1, s));
        } else if (L.kind == DT_GREEDY) {
            DT_HIP(launch_greedy_pass(e->precision, e->greedy_variant, e->d_insts, L.d_tiles, L.ntiles, i & 1, s));
        } else {
            DT_HIP(launch_match_batch(e->d_minsts, L.d_tiles, L.ntiles, s));
        }
    }
    return 0;
}

int build_graph(dt_engine *e, Lane &L)
{
    if (L.graph) {
        DT_HIP(hipGraphExecDestroy(L.graph));
        L.graph = nullptr;
    }
    hipGraph_t g = nullptr;
    DT_HIP(hipStreamBeginCapture(L.cap, hipStreamCaptureModeThreadLocal));
    const int rc = launch_passes(e, L, L.cap);
    const hipError_t ce = hipStreamEndCapture(L.cap, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return -1;
    }
    DT_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&L.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    DT_HIP(ie);
    L.graph_ntiles = L.ntiles;
    L.graph_tiles = L.d_tiles;
    e->stats.graph_builds++;
    return 0;
}

int enqueue_chunk(dt_engine *e, Lane &L)
{
    const int stage = (int)(L.seq & 1);
    if (L.dirty && rebuild_tables(e, L, stage)) return -1;
    if (e->use_graphs) {
        if (!L.graph || L.graph_ntiles != L.ntiles || L.graph_tiles != L.d_tiles)
            if (build_graph(e, L)) return -1;
        DT_HIP(hipGraphLaunch(L.graph, L.st));
    } else {
        if (launch_passes(e, L, L.st)) return -1;
    }
    Chunk c;
    c.seq = L.seq++;
    c.order = e->order++;
    for (int s = 0; s < (int)e->slots.size(); ++s) {
        Slot &sl = e->slots[s];
        if (sl.state != SLOT_ACTIVE || sl.lane != L.kind) continue;
        sl.launched += e->chunk;
        c.needs.push_back({s, sl.ticket, sl.launched});
    }
    L.inflight.push_back(std::move(c));
    e->stats.chunks++;
    e->stats.passes += e->chunk;
    return 0;
}

// The instance has stopped (or is declared stuck): queue its result copy.
int retire(dt_engine *e, int s, const Ctrl &fin, int failed)
{
    Slot &sl = e->slots[s];
    Lane &L = e->lanes[sl.lane];
    sl.final_ctrl = fin;
    sl.failed = failed;
    sl.state = SLOT_RETIRING;
    e->n_active--;
    L.n_active--;
    e->n_retiring++;
    L.dirty = true;
    if (sl.lane == DT_MATCHING) {
        // accepted node positions -> field on the device, behind the passes
        const MInst &h = e->h_minsts[s];
        DT_HIP(launch_match_out(h, e->d_minsts + s, sl.dev_field, sl.spacing[0], sl.spacing[1], sl.spacing[2],
                                L.st));
        const size_t mbytes = 3 * (size_t)h.n * sizeof(float);
        if (sl.host_io) {
            DT_HIP(hipEventRecord(sl.ev_stop, L.st));
            DT_HIP(hipStreamWaitEvent(e->io, sl.ev_stop, 0));
            DT_HIP(hipMemcpyAsync(sl.out_field, sl.dev_field, mbytes, hipMemcpyDeviceToHost, e->io));
            DT_HIP(hipEventRecord(sl.ev_done, e->io));
        } else {
            DT_HIP(hipEventRecord(sl.ev_done, L.st));
        }
        return 0;
    }
    const int rb = (fin.result_buf >= 0 && fin.result_buf < NBUF) ? fin.result_buf : 0;
    const size_t bytes = 3 * (size_t)sl.P.n * sizeof(float);
    float *src = e->h_insts[s].buf[rb];
    if (sl.host_io) {
        DT_HIP(hipEventRecord(sl.ev_stop, L.st));
        DT_HIP(hipStreamWaitEvent(e->io, sl.ev_stop, 0));
        DT_HIP(hipMemcpyAsync(sl.out_field, src, bytes, hipMemcpyDeviceToHost, e->io));
        DT_HIP(hipEventRecord(sl.ev_done, e->io));
    } else {
        if (src != sl.out_field)
            DT_HIP(hipMemcpyAsync(sl.out_field, src, bytes, hipMemcpyDeviceToDevice, L.st));
        DT_HIP(hipEventRecord(sl.ev_done, L.st));
    }
    return 0;
}

void finish(dt_engine *e, int s)
{
    Slot &sl = e->slots[s];
    dt_engine_result r{};
    r.ticket = sl.ticket;
    r.status = sl.failed ? -1 : 0;
    if (sl.lane == DT_MATCHING) {
        // matching: evaluations = cost + gradient passes, final_error = final cost
        const MCtrl &f = sl.final_mctrl;
        r.info.iterations = f.it;
        r.info.evaluations = f.cost_passes + f.grad_passes;
        r.info.rejections = 0;
        r.info.reason = sl.failed ? DT_RUNNING : f.reason;
        r.info.final_error = f.H;
        r.info.final_max_error = f.m;
        r.info.field_io = 0;
    } else {
        const Ctrl &f = sl.final_ctrl;
        r.info.iterations = f.it;
        r.info.evaluations = f.evaluations;
        r.info.rejections = f.rejections;
        r.info.reason = sl.failed ? DT_RUNNING : f.reason;
        r.info.final_error = f.E_result;
        r.info.final_max_error = f.m_result;
        r.info.field_io = f.field_io;
    }
    r.info.launches = sl.launched;
    r.info.seconds = std::chrono::duration<double>(Clock::now() - sl.t_submit).count();
    r.info.trace_len = 0;
    e->results.push_back(r);
    e->pool.release(sl.pool_off, sl.pool_bytes);
    sl.state = SLOT_FREE;
    sl.ticket = -1;
    e->n_retiring--;
    e->stats.finished++;
}

int reap(dt_engine *e, bool wait)
{
    for (int s = 0; s < (int)e->slots.size(); ++s) {
        Slot &sl = e->slots[s];
        if (sl.state != SLOT_RETIRING) continue;
        hipError_t q = wait ? hipEventSynchronize(sl.ev_done) : hipEventQuery(sl.ev_done);
        if (q == hipErrorNotReady) continue;
        DT_HIP(q);
        finish(e, s);
    }
    return 0;
}

// Retires every instance whose completion word reports the end of its solve;
// returns how many.
int retire_done(dt_engine *e)
{
    int k = 0;
    for (int s = 0; s < (int)e->slots.size(); ++s) {
        Slot &sl = e->slots[s];
        if (sl.state != SLOT_ACTIVE || !(*status_of(e, s) & 1ull)) continue;
        // the device stores the control block before the word
        std::atomic_thread_fence(std::memory_order_acquire);
        Ctrl f{};
        if (sl.lane == DT_MATCHING) std::memcpy(&sl.final_mctrl, mfinal_of(e, s), sizeof(MCtrl));
        else std::memcpy(&f, final_of(e, s), sizeof(Ctrl));
        if (retire(e, s, f, 0)) return -1;
        ++k;
    }
    return k;
}

bool chunk_ran(dt_engine *e, const Chunk &c)
{
    for (const Need &n : c.needs) {
        const Slot &sl = e->slots[n.slot];
        if (sl.state != SLOT_ACTIVE || sl.ticket != n.ticket) continue;
        const unsigned long long v = *status_of(e, n.slot);
        if ((v & 1ull) || (long long)(v >> 1) >= n.passes) continue;
        return false;
    }
    return true;
}

// A chunk has run: stop the instances that exceeded their pass budget and
// return their last accepted iterate.
int check_stuck(dt_engine *e, Lane &L, const Chunk &c)
{
    for (const Need &n : c.needs) {
        Slot &sl = e->slots[n.slot];
        if (sl.state != SLOT_ACTIVE || sl.ticket != n.ticket) continue;
        if (sl.lane == DT_MATCHING) {
            if (sl.launched <= match_max_passes(sl.params.it_max)) continue;
            DT_HIP(hipMemcpyAsync(&e->d_mctrl[n.slot].active, e->h_zero, sizeof(int), hipMemcpyHostToDevice, L.st));
            DT_HIP(hipStreamSynchronize(L.st));
            DT_HIP(hipMemcpy(&sl.final_mctrl, &e->d_mctrl[n.slot], sizeof(MCtrl), hipMemcpyDeviceToHost));
            if (retire(e, n.slot, Ctrl{}, sl.final_mctrl.phase != MPH_DONE)) return -1;
            continue;
        }
        if (sl.launched <= max_passes(sl.params)) continue;
        DT_HIP(hipMemcpyAsync(&e->d_ctrl[2 * n.slot].active, e->h_zero, sizeof(int), hipMemcpyHostToDevice, L.st));
        DT_HIP(hipMemcpyAsync(&e->d_ctrl[2 * n.slot + 1].active, e->h_zero, sizeof(int), hipMemcpyHostToDevice,
                              L.st));
        DT_HIP(hipStreamSynchronize(L.st));
        Ctrl fc[2];
        DT_HIP(hipMemcpy(fc, &e->d_ctrl[2 * n.slot], sizeof fc, hipMemcpyDeviceToHost));
        Ctrl f = latest_ctrl(fc);
        const int failed = f.phase != PH_DONE;
        if (failed) f.result_buf = f.base;
        if (retire(e, n.slot, f, failed)) return -1;
    }
    return 0;
}

// Pops the chunks that have run, in order per lane; returns how many.
int complete_ready(dt_engine *e)
{
    int k = 0;
    for (Lane &L : e->lanes) {
        while (!L.inflight.empty() && chunk_ran(e, L.inflight.front())) {
            const Chunk c = std::move(L.inflight.front());
            L.inflight.pop_front();
            if (check_stuck(e, L, c)) return -1;
            ++k;
        }
    }
    return k;
}

// Device errors; a lane whose stream drained while its words lag behind
// (host memory not updated by the device) is brought up to date from the
// control blocks in device memory.
int check_lanes(dt_engine *e)
{
    for (Lane &L : e->lanes) {
        if (L.inflight.empty()) continue;
        const hipError_t q = hipStreamQuery(L.st);
        if (q == hipErrorNotReady) continue;
        DT_HIP(q);
        if (chunk_ran(e, L.inflight.front())) continue;
        for (const Chunk &c : L.inflight)
            for (const Need &n : c.needs) {
                const Slot &sl = e->slots[n.slot];
                if (sl.state != SLOT_ACTIVE || sl.ticket != n.ticket) continue;
                if (sl.lane == DT_MATCHING) {
                    MCtrl m;
                    DT_HIP(hipMemcpy(&m, &e->d_mctrl[n.slot], sizeof(MCtrl), hipMemcpyDeviceToHost));
                    const bool mdone = m.phase == MPH_DONE;
                    if (mdone) std::memcpy(mfinal_of(e, n.slot), &m, sizeof(MCtrl));
                    *status_of(e, n.slot) =
                        ((unsigned long long)(m.cost_passes + m.grad_passes) << 1) | (mdone ? 1ull : 0ull);
                    continue;
                }
                Ctrl fc[2];
                DT_HIP(hipMemcpy(fc, &e->d_ctrl[2 * n.slot], sizeof fc, hipMemcpyDeviceToHost));
                const Ctrl &f = latest_ctrl(fc);
                const bool done = f.phase == PH_DONE;
                if (done) std::memcpy(final_of(e, n.slot), &f, sizeof(Ctrl));
                *status_of(e, n.slot) = ((unsigned long long)f.evaluations << 1) | (done ? 1ull : 0ull);
            }
    }
    return 0;
}

int fill(dt_engine *e)
{
    for (Lane &L : e->lanes)
        while ((int)L.inflight.size() < 2 && L.n_active > 0)
            if (enqueue_chunk(e, L)) return -1;
    return 0;
}

// check_lanes at most every 2 ms while nothing moves
int maybe_check_lanes(dt_engine *e)
{
    const auto now = Clock::now();
    if (now - e->last_check < std::chrono::milliseconds(2)) return 0;
    e->last_check = now;
    return check_lanes(e);
}

int progress(dt_engine *e, int wait)
{
    if (reap(e, false)) return -1;
    int moved = retire_done(e);
    if (moved < 0) return -1;
    if (fill(e)) return -1;
    const int c0 = complete_ready(e);
    if (c0 < 0) return -1;
    moved += c0;
    if (moved) e->last_check = Clock::now();
    else if (!wait && maybe_check_lanes(e)) return -1;
    if (!moved && wait) {
        bool pending = false;
        for (Lane &L : e->lanes) pending |= !L.inflight.empty();
        if (pending) {
            // follow the completion words of every lane (a blocking wait on
            // one lane would leave the other idle once its chunks have drained)
            e->last_check = Clock::now();
            for (unsigned spin = 1; !moved; ++spin) {
                const int r = retire_done(e);
                if (r < 0) return -1;
                const int c = complete_ready(e);
                if (c < 0) return -1;
                moved = r + c;
                if (moved) break;
                if ((spin & 63u) == 0 && maybe_check_lanes(e)) return -1;
                if (spin < 4096) cpu_relax();
                else std::this_thread::sleep_for(std::chrono::microseconds(10));
            }
        } else if (e->n_retiring > 0) {
            if (reap(e, true)) return -1;
        }
    }
    // keep the device busy while the host handles results
    if (fill(e)) return -1;
    if (reap(e, false)) return -1;
    return e->n_active + e->n_retiring;
}

void destroy(dt_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (Lane &L : e->lanes)
        if (L.st) (void)hipStreamSynchronize(L.st);
    if (e->io) (void)hipStreamSynchronize(e->io);
    for (Lane &L : e->lanes) {
        if (L.graph) (void)hipGraphExecDestroy(L.graph);
        if (L.d_tiles) (void)hipFree(L.d_tiles);
        for (int s = 0; s < 2; ++s) {
            if (L.h_tiles[s]) (void)hipHostFree(L.h_tiles[s]);
            if (L.ev_tiles[s]) (void)hipEventDestroy(L.ev_tiles[s]);
        }
        if (L.cap) (void)hipStreamDestroy(L.cap);
        if (L.own_st && L.st) (void)hipStreamDestroy(L.st);
    }
    for (auto &sl : e->slots) {
        if (sl.ev_ready) (void)hipEventDestroy(sl.ev_ready);
        if (sl.ev_stop) (void)hipEventDestroy(sl.ev_stop);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
    }
    if (e->h_insts) (void)hipHostFree(e->h_insts);
    if (e->h_ctrl0) (void)hipHostFree(e->h_ctrl0);
    if (e->h_zero) (void)hipHostFree(e->h_zero);
    if (e->h_minsts) (void)hipHostFree(e->h_minsts);
    if (e->h_mctrl0) (void)hipHostFree(e->h_mctrl0);
    if (e->d_minsts) (void)hipFree(e->d_minsts);
    if (e->d_mctrl) (void)hipFree(e->d_mctrl);
    if (e->coherent) (void)hipHostFree(e->coherent);
    if (e->d_insts) (void)hipFree(e->d_insts);
    if (e->d_ctrl) (void)hipFree(e->d_ctrl);
    if (e->d_acc) (void)hipFree(e->d_acc);
    if (e->d_counter) (void)hipFree(e->d_counter);
    if (e->pool_mem) (void)hipFree(e->pool_mem);
    if (e->io) (void)hipStreamDestroy(e->io);
    delete e;
}

int create(const dt_engine_config *cfg, void *stream, dt_engine *e)
{
    DT_HIP(hipGetDevice(&e->device));
    e->precision = precision_from_env();
    // gradient lane: 32 x 24 column tiles unless overridden.  Mixed sizes
    // leave fewer empty columns in 32-wide tiles than in 64-wide ones (64
    // volumes U{48..160}^3: 165 ms against 176 ms for 64 x 12, 171 for
    // 32 x 20; serving at 80 req/s: p99 41 ms against 49 ms).  The fields do
    // not depend on the tile (exact sums), so a lane's solves equal lone ones.
    e->variant = (cfg && cfg->variant >= 0 && cfg->variant < GRAD_VARIANTS) ? cfg->variant
                 : std::getenv("DISPTOOLS_GRAD_VARIANT")                  ? variant_from_env()
                                                                          : ENGINE_GRAD_VARIANT;
    // one tile shape per lane: the 32 x 16 fused pass unless overridden
    const int gv = greedy_variant_from_env();
    e->greedy_variant = gv >= 0 ? gv : GREEDY_FUSED;
    // short chunks: no host round trip separates them, and an admitted
    // instance joins at the next chunk (64 volumes U{48..160}^3: 2 passes per
    // chunk: 1 % faster than 4, 3.5 % than 8; serving percentiles unchanged)
    // (an even length: see launch_passes)
    e->chunk = (cfg && cfg->chunk > 0) ? std::min((cfg->chunk + 1) / 2 * 2, 64) : 2;
    e->use_graphs = (cfg && cfg->use_graphs) ? 1 : 0;
    e->status_words = status_words_from_env();
    e->tz_policy = cfg ? cfg->tz : 0;
    e->max_slots = (cfg && cfg->max_slots > 0) ? cfg->max_slots : 256;
    if (e->max_slots > MAX_SLOTS) DT_FAIL("max_slots is at most 4096");
    size_t budget = cfg ? cfg->hbm_budget : 0;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        DT_HIP(hipMemGetInfo(&fr, &tot));
        budget = fr / 2;
    }
    budget = budget / POOL_GRAIN * POOL_GRAIN;
    if (budget < POOL_GRAIN) DT_FAIL("hbm_budget too small");

    const int S = e->max_slots;
    for (int k = 0; k < NLANES; ++k) {
        Lane &L = e->lanes[k];
        L.kind = k;
        if (k == DT_GRADIENT && stream) {
            L.st = (hipStream_t)stream;
            L.own_st = false;
        } else {
            DT_HIP(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
        }
        DT_HIP(hipStreamCreateWithFlags(&L.cap, hipStreamNonBlocking));
        for (int s = 0; s < 2; ++s) DT_HIP(hipEventCreateWithFlags(&L.ev_tiles[s], hipEventDisableTiming));
    }
    DT_HIP(hipStreamCreateWithFlags(&e->io, hipStreamNonBlocking));

    if (hipMalloc(&e->pool_mem, budget) != hipSuccess) {
        (void)hipGetLastError();
        e->pool_mem = nullptr;
        DT_FAIL("cannot allocate the engine's device pool (hbm_budget)");
    }
    e->pool.init(budget);
    DT_HIP(hipMalloc(&e->d_insts, (size_t)S * sizeof(Inst)));
    DT_HIP(hipMalloc(&e->d_ctrl, (size_t)S * 2 * sizeof(Ctrl)));
    DT_HIP(hipMalloc(&e->d_acc, (size_t)S * ACC_SLOTS * ACC_COPIES * sizeof(Acc)));
    DT_HIP(hipMalloc(&e->d_counter, (size_t)S * sizeof(unsigned int)));
    DT_HIP(hipMemset(e->d_ctrl, 0, (size_t)S * 2 * sizeof(Ctrl)));
    DT_HIP(hipMemset(e->d_acc, 0, (size_t)S * ACC_SLOTS * ACC_COPIES * sizeof(Acc)));
    DT_HIP(hipMemset(e->d_counter, 0, (size_t)S * sizeof(unsigned int)));
    DT_HIP(hipHostMalloc((void **)&e->h_insts, (size_t)S * sizeof(Inst), hipHostMallocDefault));
    DT_HIP(hipHostMalloc((void **)&e->h_ctrl0, (size_t)S * 2 * sizeof(Ctrl), hipHostMallocDefault));
    DT_HIP(hipHostMalloc((void **)&e->h_zero, sizeof(int), hipHostMallocDefault));
    *e->h_zero = 0;
    DT_HIP(hipMalloc(&e->d_minsts, (size_t)S * sizeof(MInst)));
    DT_HIP(hipMalloc(&e->d_mctrl, (size_t)S * sizeof(MCtrl)));
    DT_HIP(hipMemset(e->d_mctrl, 0, (size_t)S * sizeof(MCtrl)));
    DT_HIP(hipHostMalloc((void **)&e->h_minsts, (size_t)S * sizeof(MInst), hipHostMallocDefault));
    DT_HIP(hipHostMalloc((void **)&e->h_mctrl0, (size_t)S * sizeof(MCtrl), hipHostMallocDefault));
    e->final_off = align_up((size_t)S * sizeof(unsigned long long), 256);
    e->mfinal_off = align_up(e->final_off + (size_t)S * sizeof(Ctrl), 256);
    const size_t coherent_bytes = e->mfinal_off + (size_t)S * sizeof(MCtrl);
    DT_HIP(host_words_alloc((void **)&e->coherent, coherent_bytes));
    DT_HIP(hipHostGetDevicePointer((void **)&e->coherent_dev, e->coherent, 0));
    std::memset(e->coherent, 0, coherent_bytes);
    e->slots.resize(S);
    for (auto &sl : e->slots) {
        DT_HIP(hipEventCreateWithFlags(&sl.ev_ready, hipEventDisableTiming));
        DT_HIP(hipEventCreateWithFlags(&sl.ev_stop, hipEventDisableTiming));
        DT_HIP(hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming));
    }
    e->stats.pool_bytes = budget;
    return 0;
}

// A matching solve (src/VolumeMatching3D.c:689-796): epsilon, delta, zeta
// and it_max of `params`; the mask is not read, as in the reference.
long long submit_matching(dt_engine *e, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                          const float *target, const dt_params *params, float *field, int host_io)
{
    if (!params) DT_FAIL("null params");
    if (matching_validate(nx, ny, nz, dx, dy, dz, params->epsilon, params->delta, params->zeta)) return -1;
    if (!target || !field) DT_FAIL("null buffer");
    int s = -1;
    for (int i = 0; i < (int)e->slots.size(); ++i)
        if (e->slots[i].state == SLOT_FREE) { s = i; break; }
    if (s < 0) return DT_ENGINE_FULL;
    const size_t bytes = match_footprint((int)nx, (int)ny, (int)nz, host_io);
    if (bytes > e->pool.size) DT_FAIL("instance larger than the engine's hbm_budget");
    size_t off = 0;
    if (!e->pool.alloc(bytes, off)) return DT_ENGINE_FULL;

    Lane &L = e->lanes[DT_MATCHING];
    Slot &sl = e->slots[s];
    sl.lane = DT_MATCHING;
    sl.P = Plan{};
    sl.params = *params;
    sl.host_io = host_io;
    sl.pool_off = off;
    sl.pool_bytes = bytes;
    sl.failed = 0;
    sl.t_submit = Clock::now();
    const size_t n = nx * ny * nz;
    Carver cv(e->pool_mem + off);
    const float *d_target;
    float *d_field;
    int rc = 0;
    if (host_io) {
        d_field = cv.take<float>(3 * n);
        float *t = cv.take<float>(n);
        if (hipMemcpyAsync(d_field, field, 3 * n * sizeof(float), hipMemcpyHostToDevice, e->io) != hipSuccess ||
            hipMemcpyAsync(t, target, n * sizeof(float), hipMemcpyHostToDevice, e->io) != hipSuccess ||
            hipEventRecord(sl.ev_ready, e->io) != hipSuccess ||
            hipStreamWaitEvent(L.st, sl.ev_ready, 0) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot stage the instance's inputs on the device");
            rc = -1;
        }
        d_target = t;
    } else {
        d_field = field;
        d_target = target;
    }
    sl.out_field = field;
    sl.dev_field = d_field;
    sl.spacing[0] = dx;
    sl.spacing[1] = dy;
    sl.spacing[2] = dz;
    if (!rc) {
        *status_of(e, s) = 0;
        MInst &h = e->h_minsts[s];
        h = match_make_inst((int)nx, (int)ny, (int)nz, d_target, params->epsilon, params->delta, params->zeta,
                            params->it_max, cv, &e->d_mctrl[s], &e->d_counter[s]);
        if (e->status_words) {
            h.status = (unsigned long long *)e->coherent_dev + s;
            h.final_ctrl = (MCtrl *)(e->coherent_dev + e->mfinal_off) + s;
        }
        e->h_mctrl0[s] = match_initial_ctrl();
        // field -> node positions, then the passes (same stream)
        if (hipMemcpyAsync(&e->d_minsts[s], &h, sizeof(MInst), hipMemcpyHostToDevice, L.st) != hipSuccess ||
            hipMemcpyAsync(&e->d_mctrl[s], &e->h_mctrl0[s], sizeof(MCtrl), hipMemcpyHostToDevice, L.st) !=
                hipSuccess ||
            hipMemsetAsync(&e->d_counter[s], 0, sizeof(unsigned int), L.st) != hipSuccess ||
            launch_match_in(h, d_field, dx, dy, dz, L.st) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot upload the instance record");
            rc = -1;
        }
    }
    if (rc) {
        e->pool.release(off, bytes);
        return -1;
    }
    sl.state = SLOT_ACTIVE;
    sl.ticket = e->next_ticket++;
    sl.launched = 0;
    e->n_active++;
    L.n_active++;
    L.dirty = true;
    e->stats.admitted++;
    return sl.ticket;
}

long long submit(dt_engine *e, int algorithm, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                 const float *target, const unsigned char *mask, const dt_params *params, float *field,
                 int host_io)
{
    if (algorithm == DT_MATCHING)
        return submit_matching(e, nx, ny, nz, dx, dy, dz, target, params, field, host_io);
    if (validate(algorithm, nx, ny, nz, dx, dy, dz, params)) return -1;
    if (!target || !mask || !field) DT_FAIL("null buffer");
    int s = -1;
    for (int i = 0; i < (int)e->slots.size(); ++i)
        if (e->slots[i].state == SLOT_FREE) { s = i; break; }
    if (s < 0) return DT_ENGINE_FULL;
    const Plan P = engine_plan(e, algorithm, (int)nx, (int)ny, (int)nz);
    const size_t bytes = footprint(e, P, host_io);
    if (bytes > e->pool.size) DT_FAIL("instance larger than the engine's hbm_budget");
    size_t off = 0;
    if (!e->pool.alloc(bytes, off)) return DT_ENGINE_FULL;

    Lane &L = e->lanes[algorithm];
    Slot &sl = e->slots[s];
    sl.lane = algorithm;
    sl.P = P;
    sl.params = *params;
    sl.host_io = host_io;
    sl.pool_off = off;
    sl.pool_bytes = bytes;
    sl.failed = 0;
    sl.t_submit = Clock::now();
    const long long n = P.n;
    Carver cv(e->pool_mem + off);
    const float *d_target;
    const uint8_t *d_mask;
    float *d_field;
    int rc = 0;
    if (host_io) {
        d_field = cv.take<float>(3 * (size_t)n);
        float *t = cv.take<float>((size_t)n);
        uint8_t *m = cv.take<uint8_t>((size_t)n);
        if (hipMemcpyAsync(d_field, field, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, e->io) !=
                hipSuccess ||
            hipMemcpyAsync(t, target, (size_t)n * sizeof(float), hipMemcpyHostToDevice, e->io) != hipSuccess ||
            hipMemcpyAsync(m, mask, (size_t)n, hipMemcpyHostToDevice, e->io) != hipSuccess ||
            hipEventRecord(sl.ev_ready, e->io) != hipSuccess ||
            hipStreamWaitEvent(L.st, sl.ev_ready, 0) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot stage the instance's inputs on the device");
            rc = -1;
        }
        d_target = t;
        d_mask = m;
    } else {
        d_field = field;
        d_target = target;
        d_mask = (const uint8_t *)mask;
    }
    sl.out_field = field;
    if (!rc) {
        *status_of(e, s) = 0;
        Acc *acc = e->d_acc + (size_t)s * ACC_SLOTS * ACC_COPIES;
        e->h_insts[s] = make_inst(P, dx, dy, dz, *params, d_target, d_mask, d_field, cv, nullptr, 0,
                                  &e->d_ctrl[2 * s], acc);
        if (e->status_words) {
            e->h_insts[s].status = (unsigned long long *)e->coherent_dev + s;
            e->h_insts[s].final_ctrl = (Ctrl *)(e->coherent_dev + e->final_off) + s;
        }
        e->h_ctrl0[2 * s] = e->h_ctrl0[2 * s + 1] = initial_ctrl(algorithm, *params);
        if (hipMemcpyAsync(&e->d_insts[s], &e->h_insts[s], sizeof(Inst), hipMemcpyHostToDevice, L.st) !=
                hipSuccess ||
            hipMemcpyAsync(&e->d_ctrl[2 * s], &e->h_ctrl0[2 * s], 2 * sizeof(Ctrl), hipMemcpyHostToDevice, L.st) !=
                hipSuccess ||
            hipMemsetAsync(acc, 0, ACC_SLOTS * ACC_COPIES * sizeof(Acc), L.st) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot upload the instance record");
            rc = -1;
        }
        // the greedy step never writes the outer face of its output buffer,
        // which starts at zero (src/disptools.c:61, displacement_field_greedy.c:43-65)
        if (!rc && algorithm == DT_GREEDY &&
            hipMemsetAsync(e->h_insts[s].buf[1], 0, 3 * (size_t)n * sizeof(float), L.st) != hipSuccess) {
            set_error(__LINE__, __FILE__, __func__, "cannot clear the greedy buffer");
            rc = -1;
        }
    }
    if (rc) {
        e->pool.release(off, bytes);
        return -1;
    }
    sl.state = SLOT_ACTIVE;
    sl.ticket = e->next_ticket++;
    sl.launched = 0;
    e->n_active++;
    L.n_active++;
    L.dirty = true;
    e->stats.admitted++;
    return sl.ticket;
}

}  // namespace
}  // namespace dt

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" dt_engine *dt_engine_create(const dt_engine_config *cfg, void *stream)
{
    disptools_clear_error();
    dt_engine *e = new (std::nothrow) dt_engine();
    if (!e) {
        set_error(__LINE__, __FILE__, __func__, "out of host memory");
        return nullptr;
    }
    if (create(cfg, stream, e)) {
        destroy(e);
        return nullptr;
    }
    return e;
}

extern "C" void dt_engine_destroy(dt_engine *e) { destroy(e); }

extern "C" size_t dt_engine_footprint(const dt_engine *e, int algorithm, size_t nx, size_t ny, size_t nz,
                                      int host_io)
{
    if (!e) return 0;
    (void)hipSetDevice(e->device);
    if (algorithm == DT_MATCHING)
        return nx < 2 || ny < 2 || nz < 2 ? 0 : match_footprint((int)nx, (int)ny, (int)nz, host_io);
    if (nx < 3 || ny < 3 || nz < 3) return 0;
    return footprint(e, engine_plan(e, algorithm, (int)nx, (int)ny, (int)nz), host_io);
}

extern "C" long long dt_engine_submit(dt_engine *e, int algorithm, size_t nx, size_t ny, size_t nz, float dx,
                                      float dy, float dz, const float *target, const unsigned char *mask,
                                      const dt_params *params, float *field, int host_io)
{
    disptools_clear_error();
    if (!e) DT_FAIL("null engine");
    DT_HIP(hipSetDevice(e->device));
    return submit(e, algorithm, nx, ny, nz, dx, dy, dz, target, mask, params, field, host_io ? 1 : 0);
}

extern "C" int dt_engine_progress(dt_engine *e, int wait)
{
    if (!e) DT_FAIL("null engine");
    DT_HIP(hipSetDevice(e->device));
    return progress(e, wait);
}

extern "C" int dt_engine_poll(dt_engine *e, dt_engine_result *out, int cap)
{
    if (!e || !out || cap <= 0) return 0;
    const int k = std::min<int>(cap, (int)e->results.size());
    std::copy(e->results.begin(), e->results.begin() + k, out);
    e->results.erase(e->results.begin(), e->results.begin() + k);
    return k;
}

extern "C" int dt_engine_drain(dt_engine *e)
{
    if (!e) DT_FAIL("null engine");
    DT_HIP(hipSetDevice(e->device));
    for (;;) {
        const int left = progress(e, 1);
        if (left < 0) return -1;
        if (left == 0) return 0;
    }
}

extern "C" int dt_engine_get_stats(const dt_engine *e, dt_engine_stats *out)
{
    if (!e || !out) return -1;
    *out = e->stats;
    out->active = e->n_active;
    out->retiring = e->n_retiring;
    out->pool_used = e->pool.used;
    out->pool_peak = e->pool.peak;
    return 0;
}
