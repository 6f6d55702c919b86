// runtime.hip -- context, device memory, launch plans and the C ABI of
// include/sister.h.
//
// The reference computes one rig per call on the host (compute_disparities,
// SisterMultiviewDisparities.hpp:26-119).  Here a call takes any number of
// rigs: the host builds one plan (rig table + per-stage block-start tables),
// uploads it, and issues one launch per stage for all rigs together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sister.h"
#include "kernels.h"
#include "pool.h"
#include "rig.h"

using namespace sister;

namespace {



// workspace layout of one rig (offsets from the workspace base)
struct WsLayout {
    size_t code, wta, med, mask, msum, part, nvol, vol, edge, scratch, vstate, sum, total;
};

inline int padded_depth(int D) { return (D + 15) & ~15; }

// keep_sum adds the debug copy of S.  separate: the per-chain SGM families,
// five volumes of Dp bytes per pixel (rows forward, columns, diagonals,
// anti-diagonals forward, rows backward); otherwise the fused vertical sweep,
// four (the rows family's forward bytes, its partial sums P in two, the rows
// family's backward bytes) or, in_place (plans short of memory; only for
// D = 16 KL CH and not in the debug sweep), two: P then replaces the rows
// bytes row by row.  The serial median's order statistics (prepass form)
// alias the volumes, which nothing writes before the SGM.  Footprints for
// admission and sister_workspace_bytes take the default, the largest form.
WsLayout ws_layout(int W, int H, int D, bool keep_sum = false, bool separate = true, bool in_place = false)
{
    const size_t Wp = (size_t)W + 2 * D, Hp = (size_t)H + 2 * D;
    const size_t npx = Wp * Hp, ncell = npx * padded_depth(D);
    WsLayout l;
    size_t o = 0;
    l.code = o;    o += align_up(8 * npx * sizeof(uint64_t));
    l.wta = o;     o += align_up(8 * npx * sizeof(uint16_t));
    l.med = o;     o += align_up(8 * npx * sizeof(uint16_t));
    l.mask = o;    o += align_up(4 * npx);
    l.msum = o;    o += align_up(ncell);
    l.nvol = separate ? 5 : (in_place && vsweep_in_place(D) && !keep_sum) ? 2 : 4;
    l.vol = align_up(ncell);
    l.part = o;    o += std::max(l.nvol * l.vol, align_up(8 * npx * sizeof(uint64_t)));
    l.scratch = l.part;
    l.edge = o;    o += align_up(4 * 2 * Wp * padded_depth(D) * sizeof(uint16_t));
    l.vstate = o;  o += align_up(2 * 3 * Wp * (padded_depth(D) + 16));
    l.sum = o;     if (keep_sum) o += align_up(ncell * sizeof(uint16_t));
    l.total = o;
    return l;
}

// host-visible staging of one rig: 5 views, 4 uint16 maps, 1 float map
size_t io_bytes(size_t WH)
{
    return 5 * align_up(WH) + 4 * align_up(WH * sizeof(uint16_t)) + align_up(WH * sizeof(float));
}

bool valid_dims(int W, int H, int D)
{
    return W >= 4 && H >= 4 && W % 4 == 0 && H % 4 == 0 && D >= 8 && D % 8 == 0 && D <= kMaxDisp &&
           W + 2 * D <= kMaxSide && H + 2 * D <= kMaxSide;
}

// Pinned host buffers for the serving path's staging, cached by size (a
// request's inputs are copied into one with the CPU, then sent with a single
// asynchronous copy; outputs come back the same way).  Freed at destroy.
struct PinnedPool {
    std::multimap<size_t, char*> idle;
    std::vector<char*> all;
    char* get(size_t bytes)
    {
        auto it = idle.lower_bound(bytes);
        if (it != idle.end() && it->first <= 2 * bytes) {
            char* p = it->second;
            idle.erase(it);
            return p;
        }
        char* p = nullptr;
        if (hipHostMalloc((void**)&p, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        all.push_back(p);
        sizes[p] = bytes;
        return p;
    }
    void put(char* p)
    {
        if (p) idle.emplace(sizes[p], p);
    }
    void release_all()
    {
        for (char* p : all) (void)hipHostFree(p);
        all.clear();
        idle.clear();
        sizes.clear();
    }
    std::unordered_map<char*, size_t> sizes;
};

// HIP events around every kernel, accumulated per kernel id.
struct EventTimer : KernelTimer {
    struct Pair { hipEvent_t a, b; int id; };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> free_events;
    double ms[kNumKernels] = {};
    int count[kNumKernels] = {};
    hipEvent_t cur_begin = nullptr;

    hipEvent_t ev()
    {
        if (!free_events.empty()) { hipEvent_t e = free_events.back(); free_events.pop_back(); return e; }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(int, hipStream_t st) override
    {
        cur_begin = ev();
        (void)hipEventRecord(cur_begin, st);
    }
    void end(int id, hipStream_t st) override
    {
        hipEvent_t e = ev();
        (void)hipEventRecord(e, st);
        pending.push_back({cur_begin, e, id});
    }
    void collect()
    {
        for (auto& p : pending) {
            float t = 0.f;
            if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
                ms[p.id] += t;
                count[p.id] += 1;
            }
            free_events.push_back(p.a);
            free_events.push_back(p.b);
        }
        pending.clear();
    }
    ~EventTimer() override
    {
        for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : free_events) (void)hipEventDestroy(e);
    }
};

// A prepared launch: rig table and block-start tables in device memory.
struct Plan {
    std::vector<RigDesc> host_rigs;   // table order (grouped by D)
    std::vector<int> slot;            // caller's rig index -> table index
    std::vector<Block> workspaces;
    Block table{nullptr, 0};
    LaunchSet ls{};
    bool in_scratch = false;          // carved from the scratch array, not the arena
};

// Bump allocation over a fixed device range (the scratch array below).
struct Carve {
    char* base;
    size_t size, used;
    int get(size_t n, Block& b)
    {
        const size_t o = align_up(used);
        if (o > size || n > size - o) return -1;
        b = {base + o, n};
        used = o + n;
        return 0;
    }
};

}  // namespace

// Device memory of the allocation-free entry point (sister_run_device_noalloc):
// the workspaces and launch tables of its rigs are carved from this array.
// Each GPU the library runs on holds its own copy until the process ends.
// In host code the name is only a stand-in for the symbol: the device address
// comes from hipGetSymbolAddress, once per device (scratch_address below).
constexpr size_t kScratchBytes = SISTER_NOALLOC_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// ---------------------------------------------------------------------------
// Launch forms chosen by the autotuner (sister_autotune).  A disparity group
// of a plan (its rigs of one D) is keyed by D and the octaves of its padded
// column total and tallest padded rig; a captured batch by its rig count, its
// distinct D count, the octave of its padded cells and its modes.  Every form
// gives the same outputs; only the time differs.
// ---------------------------------------------------------------------------
struct GroupForm {
    int fused = -1;           // -1 automatic, 0 separate chain families, 1 fused vertical sweep
    int kl[4] = {0, 0, 0, 0}; // lanes per chain of the separate families (0: automatic)
    int msum_split = 0;       // 1: the two-launch cost sum
    bool operator==(const GroupForm& o) const
    {
        return fused == o.fused && msum_split == o.msum_split && kl[0] == o.kl[0] && kl[1] == o.kl[1] &&
               kl[2] == o.kl[2] && kl[3] == o.kl[3];
    }
};
struct BatchForm {
    int parts = 0;            // independent graph parts (0: automatic)
    int by_d = 0;             // 1: one part per disparity count
    int stagger = 0;          // > 0: part h + 1 starts once part h's stages [0, stagger) are done
};
static int octave(double v) { return v <= 1 ? 0 : (int)std::floor(std::log2(v) * 2.0); }   // half octaves
struct GroupKey {
    int D, cols_b, hp_b;
    bool operator<(const GroupKey& o) const
    {
        return D != o.D ? D < o.D : cols_b != o.cols_b ? cols_b < o.cols_b : hp_b < o.hp_b;
    }
};
struct BatchKey {
    int n, nd, cells_b, modes;
    bool operator<(const BatchKey& o) const
    {
        return n != o.n ? n < o.n : nd != o.nd ? nd < o.nd : cells_b != o.cells_b ? cells_b < o.cells_b : modes < o.modes;
    }
};

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on `device`, which must be current; nullptr if
// the runtime cannot give it.
static char* scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[device] = (char*)p;
    }
    return g_scratch_ptr[device];
}

struct sister_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;           // last error of a caller thread (guarded by emu)
    mutable std::mutex emu;
    Pool pool;
    EventTimer timer;
    bool profiling = false;
    bool keep_sum = false;
    int vsweep = -1;              // vertical SGM families: -1 automatic, 0 separate, 1 fused
    bool p_in_place = false;      // debug: every plan keeps the fused sweep's P in place
    Plan last;                 // plan of the last sister_run_device (kept for debug fetch)
    Plan noalloc;              // plan of the last sister_run_device_noalloc (in the scratch array)
    std::map<GroupKey, GroupForm> gforms;   // autotuned group forms (under mu)
    std::map<BatchKey, BatchForm> bforms;   // autotuned batch forms (under mu)
    char* scratch = nullptr;   // this device's scratch array (scratch_address)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // a captured batch runs as independent parts on their own streams
    // (parallel branches of the graph): one part's latency-bound stages (the
    // serial median, launch tails) overlap another part's work
    static constexpr int kMaxParts = 8;
    Plan graph_plan[kMaxParts];
    hipStream_t side[kMaxParts] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxParts] = {}, ev_stage[kMaxParts] = {};
    FamStreams fam[kMaxParts] = {};   // per part: streams for concurrent SGM families
    std::mutex mu;             // guards pool / plans between the serving thread and callers

    // serving
    struct Request {
        uint64_t ticket;
        sister_rig rig;
        int modes;
        bool device;          // rig buffers are device memory (no staging)
        std::chrono::steady_clock::time_point t_submit;
        size_t footprint;
    };
    struct Inflight {
        std::vector<Request> reqs;
        Plan plan;
        std::vector<Block> io;         // device staging of host requests
        std::vector<char*> pinned;     // pinned host staging, per request (null for device requests)
        hipStream_t stream;
        hipEvent_t done;
    };
    PinnedPool pinned;
    std::thread server;
    std::mutex smu;
    std::condition_variable scv;
    std::deque<Request> queue;
    std::unordered_map<uint64_t, int> status;   // 0 queued, 1 running, 2 done, 3 failed
    std::deque<uint64_t> finished;               // finished tickets, oldest first
    std::string serve_error;                     // reason of the last failed request
    bool serving = false, stop = false;
    int max_inflight = 4;
    uint64_t next_ticket = 1;
    // latencies of the last kLatWindow completions (a ring)
    static constexpr size_t kLatWindow = 1 << 16;
    static constexpr size_t kKeepFinished = 1 << 16;   // finished tickets remembered
    std::vector<double> lat_ms;
    size_t lat_next = 0;
    uint64_t n_sub = 0, n_done = 0, n_rej = 0;
    std::chrono::steady_clock::time_point t_first{}, t_last{};
    bool have_first = false;
};

// The serving thread points this at its own string, so that a plan it fails
// to build never writes the context's error (which caller threads own).
static thread_local std::string* tl_err_sink = nullptr;
// The calling thread's last failure and the context it came from: what
// sister_last_error returns to that thread (like errno), so that callers on
// several threads each read their own failure.
static thread_local const sister_ctx* tl_err_ctx = nullptr;
static thread_local std::string tl_err;

static int fail(sister_ctx* c, int code, const std::string& msg)
{
    if (tl_err_sink) {
        *tl_err_sink = msg;
    } else if (c) {
        tl_err_ctx = c;
        tl_err = msg;
        std::lock_guard<std::mutex> lk(c->emu);
        c->err = msg;
    }
    return code;
}

// Under smu: a request's final state (2 done, 3 failed).  The table keeps the
// last kKeepFinished finished tickets; older ones are forgotten.
static void finish_ticket(sister_ctx* c, uint64_t t, int st)
{
    c->status[t] = st;
    c->finished.push_back(t);
    while (c->finished.size() > sister_ctx::kKeepFinished) {
        c->status.erase(c->finished.front());
        c->finished.pop_front();
    }
}

static void record_latency(sister_ctx* c, double ms)
{
    if (c->lat_ms.size() < sister_ctx::kLatWindow) {
        c->lat_ms.push_back(ms);
    } else {
        c->lat_ms[c->lat_next] = ms;
        c->lat_next = (c->lat_next + 1) % sister_ctx::kLatWindow;
    }
}

#define HIPCHK(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail(ctx, SISTER_EDEVICE, std::string(#expr ": ") + hipGetErrorString(_e));   \
    } while (0)

// Build the device plan for n rigs whose views/outputs are device pointers.
// The rig table is ordered by D so that the SGM launches can take each group
// of equal depth as one contiguous slice.
// p_in_place: fused groups keep their partial sums over the rows family's
// bytes (smaller workspaces, 0.7% slower): the retry of a plan that does not
// fit otherwise
static int build_plan(sister_ctx* c, const sister_rig* rigs, int n, Plan& p, Carve* carve = nullptr,
                      bool p_in_place = false)
{
    p.host_rigs.clear();
    p.workspaces.clear();
    p.in_scratch = carve != nullptr;
    auto get = [&](size_t bytes, Block& b) { return carve ? carve->get(bytes, b) : c->pool.get(bytes, b); };
    auto drop = [&]() {
        if (!carve)
            for (auto& w : p.workspaces) c->pool.put(w);
        p.workspaces.clear();
    };
    auto no_room = [&](const std::string& what) {
        drop();
        if (carve)
            return fail(c, SISTER_ENOMEM, what + " does not fit the " + std::to_string(kScratchBytes) +
                                              "-byte scratch of the allocation-free entry point (one 640x480 D=64 rig fits)");
        return fail(c, SISTER_ENOMEM, what + " exceeds the HBM budget");
    };
    p.slot.assign(n, 0);
    for (int i = 0; i < n; i++) {
        const sister_rig& r = rigs[i];
        if (!valid_dims(r.width, r.height, r.disp_count))
            return fail(c, SISTER_EINVAL, "rig " + std::to_string(i) + ": need W,H multiples of 4, D a multiple of 8 in [8,512], W+2D and H+2D at most 16384");
        for (int v = 0; v < 5; v++)
            if (!r.view[v]) return fail(c, SISTER_EINVAL, "rig " + std::to_string(i) + ": null view");
        if (r.stride < r.width) return fail(c, SISTER_EINVAL, "rig " + std::to_string(i) + ": stride < width");
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rigs[a].disp_count < rigs[b].disp_count; });
    int maxD = 8, maxCols = 8;
    for (int k = 0; k < n; k++) {
        const int i = order[k];
        p.slot[i] = k;
        const sister_rig& r = rigs[i];
        RigDesc g{};
        g.W = r.width; g.H = r.height; g.D = r.disp_count;
        g.Wp = g.W + 2 * g.D; g.Hp = g.H + 2 * g.D;
        g.Dp = padded_depth(g.D);
        g.stride = r.stride;
        for (int v = 0; v < 5; v++) g.view[v] = r.view[v];
        g.out[0] = r.out_multiview;
        g.out[1] = r.out_horizontal;
        g.out[2] = r.out_vertical;
        g.disp = r.out_disp;
        g.subpix = r.out_subpix;
        maxD = std::max(maxD, g.D);
        maxCols = std::max(maxCols, std::max(g.Wp, g.Hp));
        p.host_rigs.push_back(g);
    }
    // SGM groups: runs of equal D
    struct GroupHost { int i0, i1, KL[4], CH[4], vcols, maxHp, msum_split; };
    std::vector<GroupHost> groups;
    for (int k = 0; k < n;) {
        int e = k;
        long chains = 0;
        while (e < n && p.host_rigs[e].D == p.host_rigs[k].D) chains += p.host_rigs[e++].Wp;
        GroupHost G{k, e, {}, {}, 0, 0, 0};
        for (int j = k; j < e; j++) G.maxHp = std::max(G.maxHp, p.host_rigs[j].Hp);
        const int D = p.host_rigs[k].D;
        GroupForm form;
        const auto it = c->gforms.find(GroupKey{D, octave((double)chains), octave((double)G.maxHp)});
        if (it != c->gforms.end()) form = it->second;
        for (int f = 0; f < 4; f++) {
            const int kl = form.kl[f];
            const bool ok = (kl == 8 || kl == 16 || (kl == 4 && D <= 256));
            G.KL[f] = ok ? kl : sgm_pick_kl(chains, f, D);
            G.CH[f] = sgm_pick_ch(D, G.KL[f]);
        }
        // the debug switch wins over a tuned form
        G.vcols = vsweep_pick_cols(chains, D, c->vsweep >= 0 ? c->vsweep : form.fused);
        G.msum_split = form.msum_split;
        groups.push_back(G);
        k = e;
    }
    // workspaces: a rig of a fused group has four SGM volumes (two in place),
    // the others five
    std::vector<WsLayout> lay(n);
    for (const GroupHost& G : groups)
        for (int k = G.i0; k < G.i1; k++) {
            RigDesc& g = p.host_rigs[k];
            const WsLayout l = ws_layout(g.W, g.H, g.D, c->keep_sum, G.vcols == 0, p_in_place || c->p_in_place);
            lay[k] = l;
            Block b;
            if (get(l.total, b) != 0) return no_room("workspace of " + std::to_string(l.total) + " bytes");
            p.workspaces.push_back(b);
            const size_t npx = (size_t)g.Wp * g.Hp;
            for (int q = 0; q < 8; q++) {
                g.code[q] = (uint64_t*)(b.ptr + l.code) + q * npx;
                g.wta[q] = (uint16_t*)(b.ptr + l.wta) + q * npx;
                g.med[q] = (uint16_t*)(b.ptr + l.med) + q * npx;
            }
            for (int q = 0; q < 4; q++) g.mask[q] = (uint8_t*)(b.ptr + l.mask) + q * npx;
            g.msum = (uint8_t*)(b.ptr + l.msum);
            uint8_t* const vols = (uint8_t*)(b.ptr + l.part);
            for (int q = 0; q < 4; q++) g.part[q] = q < (int)l.nvol - 1 ? vols + q * l.vol : nullptr;
            g.rowsb = vols + (l.nvol - 1) * l.vol;
            g.edge = (uint16_t*)(b.ptr + l.edge);
            g.vstate = (uint32_t*)(b.ptr + l.vstate);
            g.sum = c->keep_sum ? (uint16_t*)(b.ptr + l.sum) : nullptr;
        }
    if ((int)groups.size() > kMaxDGroups) {
        drop();
        return fail(c, SISTER_EINVAL, "too many distinct disparity counts in one batch");
    }
    // table: rigs | scratch ptrs | bs[kNumStages][n+1] | per group kNumGroupKernels x (ng+1)
    const size_t rig_bytes = align_up(sizeof(RigDesc) * n);
    const size_t scr_bytes = align_up(sizeof(uint64_t*) * n);
    const size_t bs_bytes = align_up(sizeof(int) * (n + 1) * kNumStages);
    size_t gbs_ints = 0;
    for (auto& G : groups) gbs_ints += kNumGroupKernels * (size_t)(G.i1 - G.i0 + 1);
    const size_t gbs_bytes = align_up(sizeof(int) * gbs_ints);
    std::vector<char> host(rig_bytes + scr_bytes + bs_bytes + gbs_bytes, 0);
    std::memcpy(host.data(), p.host_rigs.data(), sizeof(RigDesc) * n);
    uint64_t** scr = (uint64_t**)(host.data() + rig_bytes);
    int* bs = (int*)(host.data() + rig_bytes + scr_bytes);
    int* gbs = (int*)(host.data() + rig_bytes + scr_bytes + bs_bytes);
    for (int k = 0; k < n; k++) scr[k] = (uint64_t*)(p.workspaces[k].ptr + lay[k].scratch);
    for (int st = 0; st < kNumStages; st++) p.ls.blocks[st] = 0;
    const bool wta_rows = wta_rows_ok(maxCols, maxD);
    for (int k = 0; k < n; k++) {
        int nb[kNumStages];
        stage_blocks(p.host_rigs[k], maxD, wta_rows, nb);
        for (int st = 0; st < kNumStages; st++) {
            bs[st * (n + 1) + k] = p.ls.blocks[st];
            p.ls.blocks[st] += nb[st];
        }
    }
    for (int st = 0; st < kNumStages; st++) bs[st * (n + 1) + n] = p.ls.blocks[st];
    size_t gpos = 0;
    std::vector<size_t> gofs;
    for (auto& G : groups) {
        const int ng = G.i1 - G.i0;
        gofs.push_back(gpos);
        int tot[kNumGroupKernels] = {};
        for (int k = 0; k < ng; k++) {
            const RigDesc& g = p.host_rigs[G.i0 + k];
            for (int f = 0; f < kNumGroupKernels; f++) {
                gbs[gpos + f * (ng + 1) + k] = tot[f];
                tot[f] += f < 4 ? sgm_family_blocks(g, G.KL[f], f)
                        : f == kGVSweep ? vsweep_blocks(g, G.vcols) : msum_blocks(g, f - kGMsumRows);
            }
        }
        for (int f = 0; f < kNumGroupKernels; f++) gbs[gpos + f * (ng + 1) + ng] = tot[f];
        gpos += kNumGroupKernels * (size_t)(ng + 1);
    }
    if (get(host.size(), p.table) != 0) return no_room("the launch tables of " + std::to_string(n) + " rigs");
    HIPCHK(c, hipMemcpy(p.table.ptr, host.data(), host.size(), hipMemcpyHostToDevice));
    p.ls.rigs = (const RigDesc*)p.table.ptr;
    p.ls.nrig = n;
    p.ls.scratch = (uint64_t* const*)(p.table.ptr + rig_bytes);
    for (int st = 0; st < kNumStages; st++)
        p.ls.bs[st] = (const int*)(p.table.ptr + rig_bytes + scr_bytes) + st * (n + 1);
    p.ls.maxD = maxD;
    p.ls.maxCols = maxCols;
    p.ls.wta_rows = wta_rows;
    p.ls.maxWp = 8;
    p.ls.maxHp = 8;
    for (const RigDesc& g : p.host_rigs) {
        p.ls.maxWp = std::max(p.ls.maxWp, g.Wp);
        p.ls.maxHp = std::max(p.ls.maxHp, g.Hp);
    }
    p.ls.ngroups = (int)groups.size();
    const int* dg = (const int*)(p.table.ptr + rig_bytes + scr_bytes + bs_bytes);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const GroupHost& G = groups[gi];
        const int ng = G.i1 - G.i0;
        DGroup& sg = p.ls.groups[gi];
        sg.rigs = p.ls.rigs + G.i0;
        sg.nrig = ng;
        sg.D = p.host_rigs[G.i0].D;
        sg.Dp = p.host_rigs[G.i0].Dp;
        for (int f = 0; f < kNumGroupKernels; f++) {
            sg.bs[f] = dg + gofs[gi] + f * (ng + 1);
            sg.blocks[f] = gbs[gofs[gi] + f * (ng + 1) + ng];
        }
        for (int f = 0; f < 4; f++) {
            sg.KL[f] = G.KL[f];
            sg.CH[f] = G.CH[f];
        }
        sg.vcols = G.vcols;
        sg.maxHp = G.maxHp;
        sg.msum_split = G.msum_split;
    }
    return SISTER_OK;
}

// build_plan for the batch entry points: when the arena's free memory is
// split into chunks too small for the plan (earlier plans of other shapes),
// its entirely free chunks go back to the device and the plan is built again
// once.  (hipFree synchronises the device; this runs only on that failure.)
// A plan that does not fit: first return the arena's entirely free chunks
// (its free memory may sit in chunks shaped for earlier plans) and build
// again, then build with the fused sweep's partial sums in place.
static int build_plan_compact(sister_ctx* c, const sister_rig* rigs, int n, Plan& p)
{
    int rc = build_plan(c, rigs, n, p);
    if (rc != SISTER_ENOMEM) return rc;
    c->pool.trim();
    rc = build_plan(c, rigs, n, p);
    if (rc != SISTER_ENOMEM) return rc;
    return build_plan(c, rigs, n, p, nullptr, true);
}

static void free_plan(sister_ctx* c, Plan& p)
{
    if (!p.in_scratch) {
        for (auto& w : p.workspaces) c->pool.put(w);
        c->pool.put(p.table);
    }
    p.workspaces.clear();
    p.table = {nullptr, 0};
    p.host_rigs.clear();
    p.in_scratch = false;
}

// enqueue every stage of a plan on `st`
// Stages of a plan, in order: 0 census, 1 WTA, 2 median, 3 masks, then per
// mode m 4 + 2m the cost sum and 5 + 2m the SGM (with the winner).
constexpr int kNumPlanStages = 10;

// enqueue stages [from, to) of a plan on `st`
static void enqueue_range(sister_ctx* c, const Plan& p, int modes, hipStream_t st, const FamStreams* fs,
                          KernelTimer* t, int from, int to)
{
    const LaunchSet& ls = p.ls;
    auto in = [&](int k) { return k >= from && k < to; };
    if (in(0)) launch_census(ls, st, t);
    if (in(1)) launch_wta(ls, st, t);
    if (in(2)) launch_median(ls, st, t);
    int fmask = 0;
    if (modes & SISTER_MODE_MULTIVIEW) fmask |= 0xF;
    if (modes & SISTER_MODE_HORIZONTAL) fmask |= 0x5;
    if (modes & SISTER_MODE_VERTICAL) fmask |= 0xA;
    if (in(3)) launch_mask(ls, fmask, st, t);
    static const int fuse[3] = {0xF, 0x5, 0xA};   // hpp:264-275: frames summed per mode
    for (int m = 0; m < 3; m++) {
        if (!(modes & (1 << m))) continue;
        if (in(4 + 2 * m)) launch_msum(ls, fuse[m], st, t);
        if (in(5 + 2 * m)) launch_sgm(ls, m, c->keep_sum ? 1 : 0, st, fs, t);
    }
}

// enqueue every stage of a plan on `st`
static void enqueue(sister_ctx* c, const Plan& p, int modes, hipStream_t st, const FamStreams* fs, KernelTimer* t)
{
    enqueue_range(c, p, modes, st, fs, t, 0, kNumPlanStages);
}

static int check_modes(sister_ctx* c, const sister_rig* rigs, int n, int modes)
{
    if (modes <= 0 || modes > 7) return fail(c, SISTER_EINVAL, "modes must be a non-empty subset of SISTER_MODE_ALL");
    for (int i = 0; i < n; i++) {
        if ((modes & 1) && !rigs[i].out_multiview) return fail(c, SISTER_EINVAL, "out_multiview is NULL");
        if ((modes & 2) && !rigs[i].out_horizontal) return fail(c, SISTER_EINVAL, "out_horizontal is NULL");
        if ((modes & 4) && !rigs[i].out_vertical) return fail(c, SISTER_EINVAL, "out_vertical is NULL");
    }
    return SISTER_OK;
}

extern "C" {

int sister_create(const sister_config* cfg, sister_ctx** out)
{
    if (!out) return SISTER_EINVAL;
    *out = nullptr;
    sister_ctx* c = new sister_ctx();
    c->device = cfg ? cfg->device : 0;
    c->pool.dev_alloc = [](char** p, size_t n) {
        if (hipMalloc((void**)p, n) == hipSuccess) return 0;
        (void)hipGetLastError();
        return -1;
    };
    c->pool.dev_free = [](char* p) { (void)hipFree(p); };
    if (cfg && cfg->hbm_budget) c->pool.budget = cfg->hbm_budget;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || c->device < 0) {
        (void)hipGetLastError();
        delete c;
        return SISTER_EDEVICE;
    }
    bool ok = hipSetDevice(c->device) == hipSuccess &&
              hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 1; ok && k < sister_ctx::kMaxParts; k++)
        ok = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < sister_ctx::kMaxParts; k++)
        ok = hipEventCreateWithFlags(&c->ev_stage[k], hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < sister_ctx::kMaxParts; k++) {
        FamStreams& f = c->fam[k];
        ok = hipEventCreateWithFlags(&f.fork, hipEventDisableTiming) == hipSuccess;
        for (int j = 0; ok && j < 2; j++)
            ok = hipStreamCreateWithFlags(&f.side[j], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&f.join[j], hipEventDisableTiming) == hipSuccess;
    }
    ok = ok && kernels_setup(c->device) == 0;
    // a failed look-up leaves the allocation-free entry point reporting it
    if (ok) c->scratch = scratch_address(c->device);
    if (!ok) {
        delete c;
        return SISTER_EDEVICE;
    }
    if (!(cfg && cfg->hbm_budget)) {
        // default: 90% of the HBM free at creation
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
            c->pool.budget = free_b / 10 * 9;
    }
    *out = c;
    return SISTER_OK;
}

void sister_destroy(sister_ctx* c)
{
    if (!c) return;
    if (tl_err_ctx == c) tl_err_ctx = nullptr;
    sister_serve_stop(c);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    free_plan(c, c->last);
    for (auto& gp : c->graph_plan) free_plan(c, gp);
    c->timer.collect();
    c->pool.release_all();
    c->pinned.release_all();
    (void)hipEventDestroy(c->ev_fork);
    for (int k = 0; k < sister_ctx::kMaxParts; k++) {
        if (c->ev_stage[k]) (void)hipEventDestroy(c->ev_stage[k]);
        if (k == 0) continue;   // part 0 runs on the context stream
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    }
    for (auto& f : c->fam) {
        (void)hipEventDestroy(f.fork);
        for (int j = 0; j < 2; j++) {
            (void)hipEventDestroy(f.join[j]);
            (void)hipStreamDestroy(f.side[j]);
        }
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* sister_last_error(const sister_ctx* c)
{
    // the calling thread's own last failure on this context; otherwise a
    // copy of the context's last failure (from any thread).  Either pointer
    // stays valid until this thread's next call.
    static thread_local std::string copy;
    if (!c) return "null context";
    if (tl_err_ctx == c) return tl_err.c_str();
    std::lock_guard<std::mutex> lk(c->emu);
    copy = c->err;
    return copy.c_str();
}

void* sister_stream(sister_ctx* c) { return c ? (void*)c->stream : nullptr; }

size_t sister_workspace_bytes(int W, int H, int D)
{
    if (!valid_dims(W, H, D)) return 0;
    return ws_layout(W, H, D).total;
}

int sister_outside_reference_domain(int D) { return D > 256 ? 1 : 0; }

int sister_run_device(sister_ctx* c, const sister_rig* rigs, int n, int modes)
{
    if (!c || !rigs || n <= 0) return fail(c, SISTER_EINVAL, "bad arguments");
    int rc = check_modes(c, rigs, n, modes);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // the previous plan may still be in use
    free_plan(c, c->last);
    rc = build_plan_compact(c, rigs, n, c->last);
    if (rc) return rc;
    // profiling times every kernel on its own: the SGM families then run in sequence
    enqueue(c, c->last, modes, c->stream, c->profiling ? nullptr : &c->fam[0], c->profiling ? &c->timer : nullptr);
    HIPCHK(c, hipGetLastError());
    return SISTER_OK;
}

int sister_run_device_noalloc(sister_ctx* c, const sister_rig* rigs, int n, int modes)
{
    if (!c || !rigs || n <= 0) return fail(c, SISTER_EINVAL, "bad arguments");
    int rc = check_modes(c, rigs, n, modes);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->scratch) return fail(c, SISTER_EDEVICE, "the scratch array's device address is unavailable");
    // this context's previous call may still be reading the scratch array
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_plan(c, c->noalloc);
    Carve carve{c->scratch, kScratchBytes, 0};
    rc = build_plan(c, rigs, n, c->noalloc, &carve);
    if (rc == SISTER_ENOMEM) {
        carve = Carve{c->scratch, kScratchBytes, 0};
        rc = build_plan(c, rigs, n, c->noalloc, &carve, true);
    }
    if (rc) return rc;
    enqueue(c, c->noalloc, modes, c->stream, c->profiling ? nullptr : &c->fam[0], c->profiling ? &c->timer : nullptr);
    HIPCHK(c, hipGetLastError());
    return SISTER_OK;
}

int sister_sync(sister_ctx* c)
{
    if (!c) return SISTER_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->profiling) c->timer.collect();
    return SISTER_OK;
}

int sister_compute_rig(sister_ctx* c, const sister_rig* r, int modes)
{
    if (!c || !r) return fail(c, SISTER_EINVAL, "bad arguments");
    int rc = check_modes(c, r, 1, modes);
    if (rc) return rc;
    if (!valid_dims(r->width, r->height, r->disp_count))
        return fail(c, SISTER_EINVAL, "need W,H multiples of 4, D a multiple of 8 in [8,512], W+2D and H+2D at most 16384");
    if (r->stride < r->width) return fail(c, SISTER_EINVAL, "stride < width");
    const size_t WH = (size_t)r->width * r->height;
    Block io;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        HIPCHK(c, hipSetDevice(c->device));
        if (c->pool.get(io_bytes(WH), io) != 0)
            return fail(c, SISTER_ENOMEM, "staging allocation failed");
    }
    char* p = io.ptr;
    sister_rig d = *r;
    d.stride = r->width;
    for (int v = 0; v < 5; v++) {
        if (!r->view[v]) { c->pool.put(io); return fail(c, SISTER_EINVAL, "null view"); }
        hipError_t e = hipMemcpy2DAsync(p, r->width, r->view[v], r->stride, r->width, r->height, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) { c->pool.put(io); return fail(c, SISTER_EDEVICE, hipGetErrorString(e)); }
        d.view[v] = (const uint8_t*)p;
        p += align_up(WH);
    }
    uint16_t* dm = (uint16_t*)p; p += align_up(WH * 2);
    uint16_t* dh = (uint16_t*)p; p += align_up(WH * 2);
    uint16_t* dv = (uint16_t*)p; p += align_up(WH * 2);
    uint16_t* dd = (uint16_t*)p; p += align_up(WH * 2);
    float* dsp = (float*)p;
    d.out_multiview = (modes & 1) ? dm : nullptr;
    d.out_horizontal = (modes & 2) ? dh : nullptr;
    d.out_vertical = (modes & 4) ? dv : nullptr;
    d.out_disp = r->out_disp ? dd : nullptr;
    d.out_subpix = r->out_subpix ? dsp : nullptr;
    rc = sister_run_device(c, &d, 1, modes);
    if (rc == SISTER_OK) {
        hipError_t e = hipSuccess;
        if (modes & 1) e = hipMemcpyAsync(r->out_multiview, dm, WH * 2, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && (modes & 2)) e = hipMemcpyAsync(r->out_horizontal, dh, WH * 2, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && (modes & 4)) e = hipMemcpyAsync(r->out_vertical, dv, WH * 2, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && r->out_disp) e = hipMemcpyAsync(r->out_disp, dd, WH * 2, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && r->out_subpix) e = hipMemcpyAsync(r->out_subpix, dsp, WH * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, SISTER_EDEVICE, hipGetErrorString(e));
        if (c->profiling) c->timer.collect();
    }
    {
        std::lock_guard<std::mutex> lk(c->mu);
        c->pool.put(io);
    }
    return rc;
}

}  // extern "C"

static BatchKey batch_key(const sister_rig* rigs, int n, int modes)
{
    double cells = 0;
    std::vector<int> ds;
    for (int i = 0; i < n; i++) {
        const sister_rig& r = rigs[i];
        cells += (double)(r.width + 2 * r.disp_count) * (r.height + 2 * r.disp_count) * r.disp_count;
        ds.push_back(r.disp_count);
    }
    std::sort(ds.begin(), ds.end());
    const int nd = (int)(std::unique(ds.begin(), ds.end()) - ds.begin());
    return BatchKey{n, nd, octave(cells), modes};
}

// The independent parts of a captured batch (parallel branches of its graph),
// under mu.  By default (or as tuned) the rigs are balanced by padded cells
// over the parts, largest first to the lightest part; the tuned form may
// instead give each disparity count a part of its own, so that groups of
// different D run side by side.  SISTER_GRAPH_PARTS (1..8) overrides the
// default part count (comparison runs).
static int partition_batch(sister_ctx* c, const sister_rig* rigs, int n, int modes,
                           std::vector<sister_rig> (&part)[sister_ctx::kMaxParts], int* stagger = nullptr)
{
    static const int env_parts = [] {
        const char* e = getenv("SISTER_GRAPH_PARTS");
        const int v = e ? atoi(e) : 0;
        return v < 0 ? 0 : (v > sister_ctx::kMaxParts ? sister_ctx::kMaxParts : v);
    }();
    // SISTER_GRAPH_STAGGER (1..9) overrides the stagger of the parts (comparison runs)
    static const int env_stagger = [] {
        const char* e = getenv("SISTER_GRAPH_STAGGER");
        const int v = e ? atoi(e) : -1;
        return v < 0 || v >= kNumPlanStages ? -1 : v;
    }();
    BatchForm form;
    const auto it = c->bforms.find(batch_key(rigs, n, modes));
    if (it != c->bforms.end()) form = it->second;
    if (stagger) *stagger = env_stagger >= 0 ? env_stagger : form.stagger;
    for (auto& p : part) p.clear();
    if (form.by_d) {
        std::vector<int> ds;
        for (int i = 0; i < n; i++) ds.push_back(rigs[i].disp_count);
        std::sort(ds.begin(), ds.end());
        ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
        if ((int)ds.size() <= sister_ctx::kMaxParts) {
            for (int i = 0; i < n; i++)
                part[std::lower_bound(ds.begin(), ds.end(), rigs[i].disp_count) - ds.begin()].push_back(rigs[i]);
            return (int)ds.size();
        }
    }
    const int want = env_parts ? env_parts : form.parts > 0 ? form.parts : 2;
    const int nparts = std::max(1, std::min(std::min(want, (int)sister_ctx::kMaxParts), n));
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    auto cells = [&](int i) {
        const sister_rig& r = rigs[i];
        return (double)(r.width + 2 * r.disp_count) * (r.height + 2 * r.disp_count) * r.disp_count;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    double load[sister_ctx::kMaxParts] = {};
    for (int i : order) {
        const int h = (int)(std::min_element(load, load + nparts) - load);
        part[h].push_back(rigs[i]);
        load[h] += cells(i);
    }
    return nparts;
}

extern "C" {

int sister_graph_capture(sister_ctx* c, const sister_rig* rigs, int n, int modes)
{
    if (!c || !rigs || n <= 0) return fail(c, SISTER_EINVAL, "bad arguments");
    int rc = check_modes(c, rigs, n, modes);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    for (auto& gp : c->graph_plan) free_plan(c, gp);
    std::vector<sister_rig> part[sister_ctx::kMaxParts];
    int stagger = 0;
    const int nparts = partition_batch(c, rigs, n, modes, part, &stagger);
    auto build_parts = [&](bool in_place) {
        for (int h = 0; h < nparts; h++) {
            const int r = build_plan(c, part[h].data(), (int)part[h].size(), c->graph_plan[h], nullptr, in_place);
            if (r) {
                for (auto& gp : c->graph_plan) free_plan(c, gp);
                return r;
            }
        }
        return (int)SISTER_OK;
    };
    rc = build_parts(false);
    if (rc == SISTER_ENOMEM) {
        // the arena's free memory may sit in chunks shaped for earlier plans:
        // return the entirely free ones to the device and build again, then
        // with the fused sweep's partial sums in place
        c->pool.trim();
        rc = build_parts(false);
        if (rc == SISTER_ENOMEM) rc = build_parts(true);
    }
    if (rc) return rc;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    (void)hipEventRecord(c->ev_fork, c->stream);
    for (int h = 1; h < nparts; h++) (void)hipStreamWaitEvent(c->side[h], c->ev_fork, 0);
    // the SGM families of a part run in sequence inside a captured graph
    // (concurrent families are used by eager runs only: a capture that forked
    // them from the part streams ended in a host-side crash in the runtime)
    auto part_stream = [&](int h) { return h ? c->side[h] : c->stream; };
    if (stagger > 0 && nparts > 1) {
        // part h + 1's first stages wait for part h's: one part's SGM then
        // runs beside the next part's census, WTA and cost sum
        for (int h = 0; h < nparts; h++) {
            if (h) (void)hipStreamWaitEvent(part_stream(h), c->ev_stage[h - 1], 0);
            enqueue_range(c, c->graph_plan[h], modes, part_stream(h), nullptr, nullptr, 0, stagger);
            (void)hipEventRecord(c->ev_stage[h], part_stream(h));
        }
        for (int h = 0; h < nparts; h++)
            enqueue_range(c, c->graph_plan[h], modes, part_stream(h), nullptr, nullptr, stagger, kNumPlanStages);
    } else {
        for (int h = 0; h < nparts; h++) enqueue(c, c->graph_plan[h], modes, part_stream(h), nullptr, nullptr);
    }
    for (int h = 1; h < nparts; h++) {
        (void)hipEventRecord(c->ev_join[h], c->side[h]);
        (void)hipStreamWaitEvent(c->stream, c->ev_join[h], 0);
    }
    hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
    if (e != hipSuccess) return fail(c, SISTER_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
    HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    return SISTER_OK;
}

int sister_graph_launch(sister_ctx* c)
{
    if (!c || !c->graph_exec) return fail(c, SISTER_EINVAL, "no captured graph");
    HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    return SISTER_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// autotuning
// ---------------------------------------------------------------------------
static std::string forms_text(const sister_ctx* c)
{
    std::string t = "# sister launch forms v1\n";
    char line[256];
    for (const auto& kv : c->bforms) {
        std::snprintf(line, sizeof line, "batch %d %d %d %d parts %d by_d %d stagger %d\n", kv.first.n,
                      kv.first.nd, kv.first.cells_b, kv.first.modes, kv.second.parts, kv.second.by_d,
                      kv.second.stagger);
        t += line;
    }
    for (const auto& kv : c->gforms) {
        const GroupForm& f = kv.second;
        std::snprintf(line, sizeof line, "group %d %d %d fused %d kl %d %d %d %d msum_split %d\n", kv.first.D,
                      kv.first.cols_b, kv.first.hp_b, f.fused, f.kl[0], f.kl[1], f.kl[2], f.kl[3], f.msum_split);
        t += line;
    }
    return t;
}

// Entries of a forms file into the tables (existing keys are replaced).
// Returns false on a malformed line.
static bool forms_parse(sister_ctx* c, FILE* fp)
{
    char line[512];
    while (std::fgets(line, sizeof line, fp)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        BatchKey bk;
        BatchForm bf;
        GroupKey gk;
        GroupForm gf;
        if (std::sscanf(line, "batch %d %d %d %d parts %d by_d %d stagger %d", &bk.n, &bk.nd, &bk.cells_b,
                        &bk.modes, &bf.parts, &bf.by_d, &bf.stagger) >= 6) {
            c->bforms[bk] = bf;
        } else if (std::sscanf(line, "group %d %d %d fused %d kl %d %d %d %d msum_split %d", &gk.D, &gk.cols_b,
                               &gk.hp_b, &gf.fused, &gf.kl[0], &gf.kl[1], &gf.kl[2], &gf.kl[3],
                               &gf.msum_split) == 9) {
            c->gforms[gk] = gf;
        } else {
            return false;
        }
    }
    return true;
}

// Group keys of the plans a capture of this batch would build (under mu).
static std::vector<GroupKey> batch_group_keys(sister_ctx* c, const sister_rig* rigs, int n, int modes)
{
    std::vector<sister_rig> part[sister_ctx::kMaxParts];
    const int np = partition_batch(c, rigs, n, modes, part);
    std::vector<GroupKey> keys;
    for (int h = 0; h < np; h++) {
        std::map<int, std::pair<long, int>> g;   // D -> (columns, tallest)
        for (const sister_rig& r : part[h]) {
            auto& e = g[r.disp_count];
            e.first += r.width + 2 * r.disp_count;
            e.second = std::max(e.second, r.height + 2 * r.disp_count);
        }
        for (const auto& kv : g) {
            const GroupKey k{kv.first, octave((double)kv.second.first), octave((double)kv.second.second)};
            if (std::find_if(keys.begin(), keys.end(), [&](const GroupKey& o) { return !(o < k) && !(k < o); }) ==
                keys.end())
                keys.push_back(k);
        }
    }
    return keys;
}

extern "C" {

int sister_autotune(sister_ctx* c, const sister_rig* rigs, int n, int modes, const char* path)
{
    if (!c || !rigs || n <= 0) return fail(c, SISTER_EINVAL, "bad arguments");
    int rc = check_modes(c, rigs, n, modes);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (path) {
            FILE* fp = std::fopen(path, "r");
            if (fp) {
                const bool ok = forms_parse(c, fp);
                std::fclose(fp);
                if (!ok) return fail(c, SISTER_EINVAL, std::string("malformed launch-forms file ") + path);
            }
        }
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        free_plan(c, c->last);   // room for the captures
    }
    // one candidate: capture, one warm replay, then the mean of three; a form
    // whose workspaces do not fit the budget is skipped (infinite time)
    auto time_graph = [&](double& ms) -> int {
        int r = sister_graph_capture(c, rigs, n, modes);
        if (r == SISTER_ENOMEM) {
            ms = 1e30;
            return SISTER_OK;
        }
        if (r) return r;
        HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < 3; i++) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 3;
        return SISTER_OK;
    };
    const BatchKey bk = batch_key(rigs, n, modes);
    bool tuned_batch;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        tuned_batch = c->bforms.count(bk) != 0;
    }
    // 1. the batch's graph parts: 1 to 4 balanced parts, or a part per D
    if (!tuned_batch) {
        std::vector<BatchForm> cand;
        for (int p = 1; p <= std::min(4, n); p++) cand.push_back(BatchForm{p, 0});
        if (bk.nd >= 2 && bk.nd <= sister_ctx::kMaxParts) cand.push_back(BatchForm{0, 1});
        double best = 1e30;
        BatchForm bestf;
        for (const BatchForm& f : cand) {
            {
                std::lock_guard<std::mutex> lk(c->mu);
                c->bforms[bk] = f;
            }
            double ms = 0;
            if ((rc = time_graph(ms))) return rc;
            if (ms < best) { best = ms; bestf = f; }
        }
        // staggered starts of the chosen parts: after the masks, or after the
        // first cost sum
        if (bestf.parts != 1 || bestf.by_d) {
            for (int st : {4, 5}) {
                BatchForm f = bestf;
                f.stagger = st;
                {
                    std::lock_guard<std::mutex> lk(c->mu);
                    c->bforms[bk] = f;
                }
                double ms = 0;
                if ((rc = time_graph(ms))) return rc;
                if (ms < best) { best = ms; bestf = f; }
            }
        }
        std::lock_guard<std::mutex> lk(c->mu);
        c->bforms[bk] = bestf;
    }
    // 2. each disparity group of those parts: vertical families fused or
    // separate, lanes per chain, cost-sum form (the other groups keep theirs)
    std::vector<GroupKey> keys;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        keys = batch_group_keys(c, rigs, n, modes);
    }
    for (const GroupKey& k : keys) {
        {
            std::lock_guard<std::mutex> lk(c->mu);
            if (c->gforms.count(k)) continue;
        }
        std::vector<GroupForm> cand;
        GroupForm f;
        cand.push_back(f);                                       // automatic
        f.fused = 0; cand.push_back(f);                          // separate families, automatic lanes
        if (vsweep_pick_cols(1, k.D, 1) > 0) {                   // fused vertical sweep, where it reaches
            f.fused = 1;
            cand.push_back(f);
        }
        f = GroupForm(); f.fused = 0;
        for (int q = 0; q < 4; q++) f.kl[q] = 16;
        cand.push_back(f);                                       // separate, 16 lanes everywhere
        f.kl[2] = f.kl[3] = 8; cand.push_back(f);                // separate, diagonals on 8
        f = GroupForm(); f.msum_split = 1; cand.push_back(f);    // two-launch cost sum
        double best = 1e30;
        GroupForm bestf;
        for (const GroupForm& g : cand) {
            {
                std::lock_guard<std::mutex> lk(c->mu);
                c->gforms[k] = g;
            }
            double ms = 0;
            if ((rc = time_graph(ms))) return rc;
            if (ms < best) { best = ms; bestf = g; }
        }
        std::lock_guard<std::mutex> lk(c->mu);
        c->gforms[k] = bestf;
    }
    // leave the batch captured in its tuned form
    if ((rc = sister_graph_capture(c, rigs, n, modes))) return rc;
    if (path) {
        std::lock_guard<std::mutex> lk(c->mu);
        const std::string t = forms_text(c);
        FILE* fp = std::fopen(path, "w");
        if (!fp || std::fwrite(t.data(), 1, t.size(), fp) != t.size()) {
            if (fp) std::fclose(fp);
            return fail(c, SISTER_EINVAL, std::string("cannot write ") + path);
        }
        std::fclose(fp);
    }
    return SISTER_OK;
}

int sister_tune_describe(sister_ctx* c, char* buf, size_t len)
{
    if (!c || !buf || len == 0) return SISTER_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    const std::string t = forms_text(c);
    const size_t k = std::min(len - 1, t.size());
    std::memcpy(buf, t.data(), k);
    buf[k] = 0;
    return t.size() < len ? SISTER_OK : fail(c, SISTER_EINVAL, "buffer too small for the forms table");
}

int sister_profile_enable(sister_ctx* c, int on)
{
    if (!c) return SISTER_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timer.collect();
    if (on) {
        for (int i = 0; i < kNumKernels; i++) { c->timer.ms[i] = 0; c->timer.count[i] = 0; c->timer.dispatched[i] = 0; }
    }
    c->profiling = on != 0;
    return SISTER_OK;
}

int sister_profile_read(sister_ctx* c, const char** names, double* ms, int* launches, int max)
{
    if (!c) return 0;
    (void)hipStreamSynchronize(c->stream);
    c->timer.collect();
    int n = std::min(max, (int)kNumKernels);
    for (int i = 0; i < n; i++) {
        if (names) names[i] = kKernelNames[i];
        if (ms) ms[i] = c->timer.ms[i];
        if (launches) launches[i] = (int)c->timer.dispatched[i];
    }
    return n;
}

int sister_debug_fetch(sister_ctx* c, int index, const char* what, int which, void* dst, size_t bytes)
{
    if (!c || !what || !dst) return fail(c, SISTER_EINVAL, "bad arguments");
    if (index < 0 || index >= (int)c->last.host_rigs.size()) return fail(c, SISTER_EINVAL, "no such rig in the last run");
    const RigDesc& g = c->last.host_rigs[c->last.slot[index]];
    const size_t npx = (size_t)g.Wp * g.Hp;
    const void* src = nullptr;
    size_t n = 0;
    std::string w(what);
    if (w == "census" && which >= 0 && which < 8) { src = g.code[which]; n = npx * 8; }
    else if (w == "wta" && which >= 0 && which < 8) { src = g.wta[which]; n = npx * 2; }
    else if (w == "med" && which >= 0 && which < 8) { src = g.med[which]; n = npx * 2; }
    else if (w == "mask" && which >= 0 && which < 4) { src = g.mask[which]; n = npx; }
    else if (w == "msum") { src = g.msum; n = npx * g.Dp; }
    else if (w == "sum") {
        if (!g.sum) return fail(c, SISTER_EINVAL, "S is only kept after sister_debug_keep_sum(ctx, 1)");
        src = g.sum; n = npx * g.Dp * 2;
    }
    else return fail(c, SISTER_EINVAL, "unknown intermediate " + w);
    if (bytes < n) return fail(c, SISTER_EINVAL, "destination too small");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
    if (w == "census") {
        // rows of the deep padding bands are stored once, in each band's first row
        const int F = which / 2, R = frame_rows(g, F), C = frame_cols(g, F);
        uint64_t* d = (uint64_t*)dst;
        for (int r = 0; r < R; r++) {
            const int b = band_row(r, R, g.D);
            if (b != r) std::memcpy(d + (size_t)r * C, d + (size_t)b * C, (size_t)C * 8);
        }
    }
    return SISTER_OK;
}

// debug: keep S of crop cells (otherwise the final sweep consumes it)
int sister_debug_keep_sum(sister_ctx* c, int on)
{
    if (!c) return SISTER_EINVAL;
    c->keep_sum = on != 0;
    return SISTER_OK;
}

// debug: form of the vertical SGM families in plans built from now on
int sister_debug_vsweep(sister_ctx* c, int mode)
{
    if (!c || mode < -1 || mode > 1) return SISTER_EINVAL;
    c->vsweep = mode;
    return SISTER_OK;
}

// debug: the fused sweep's partial sums in place in plans built from now on
int sister_debug_sweep_in_place(sister_ctx* c, int on)
{
    if (!c) return SISTER_EINVAL;
    c->p_in_place = on != 0;
    return SISTER_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// serving
// ---------------------------------------------------------------------------
static size_t rig_footprint(const sister_rig& r, bool device)
{
    const size_t WH = (size_t)r.width * r.height;
    return ws_layout(r.width, r.height, r.disp_count).total + (device ? 0 : io_bytes(WH)) + 16 * kAlign;
}

static void serve_loop(sister_ctx* c)
{
    std::string plan_error;   // this thread's failures (see fail())
    tl_err_sink = &plan_error;
    (void)hipSetDevice(c->device);
    std::vector<sister_ctx::Inflight*> inflight;
    std::vector<hipStream_t> streams;
    for (int i = 0; i < c->max_inflight; i++) {
        hipStream_t s;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        streams.push_back(s);
    }
    size_t reserved = 0;   // footprint of admitted requests
    // Admission counts footprints against the budget; the arena may still not
    // place a batch (chunk rounding, fragmentation).  Such a batch goes back to
    // the queue, and when nothing else is running the next one is halved.
    size_t admit_cap = (size_t)-1;
    for (;;) {
        // retire finished batches
        for (size_t i = 0; i < inflight.size();) {
            sister_ctx::Inflight* f = inflight[i];
            hipError_t q = hipEventQuery(f->done);
            if (q == hipErrorNotReady) { i++; continue; }
            // host requests: outputs from the pinned staging to the caller
            for (size_t k = 0; k < f->reqs.size(); k++) {
                char* hp = f->pinned[k];
                if (!hp || q != hipSuccess) continue;
                const sister_rig& h = f->reqs[k].rig;
                const int modes = f->reqs[k].modes;
                const size_t WH = (size_t)h.width * h.height;
                const char* o = hp + 5 * align_up(WH);
                const size_t m2 = align_up(WH * 2);
                if (h.out_multiview && (modes & 1)) std::memcpy(h.out_multiview, o, WH * 2);
                if (h.out_horizontal && (modes & 2)) std::memcpy(h.out_horizontal, o + m2, WH * 2);
                if (h.out_vertical && (modes & 4)) std::memcpy(h.out_vertical, o + 2 * m2, WH * 2);
                if (h.out_disp && (modes & 1)) std::memcpy(h.out_disp, o + 3 * m2, WH * 2);
                if (h.out_subpix && (modes & 1)) std::memcpy(h.out_subpix, o + 4 * m2, WH * 4);
            }
            auto now = std::chrono::steady_clock::now();
            {
                std::lock_guard<std::mutex> lk(c->mu);
                free_plan(c, f->plan);
                for (auto& b : f->io) c->pool.put(b);
            }
            for (char* hp : f->pinned) c->pinned.put(hp);
            {
                std::lock_guard<std::mutex> lk(c->smu);
                for (auto& r : f->reqs) {
                    finish_ticket(c, r.ticket, (q == hipSuccess) ? 2 : 3);
                    if (q != hipSuccess) c->serve_error = hipGetErrorString(q);
                    record_latency(c, std::chrono::duration<double, std::milli>(now - r.t_submit).count());
                    reserved -= r.footprint;
                    c->n_done++;
                }
                c->t_last = now;
            }
            c->scv.notify_all();
            (void)hipEventDestroy(f->done);
            streams.push_back(f->stream);
            delete f;
            inflight.erase(inflight.begin() + i);
        }
        // admit: every queued request whose footprint fits, into one batch
        std::vector<sister_ctx::Request> batch;
        {
            std::unique_lock<std::mutex> lk(c->smu);
            if (c->stop && c->queue.empty() && inflight.empty()) break;
            if (!streams.empty()) {
                while (!c->queue.empty() && batch.size() < admit_cap) {
                    auto& r = c->queue.front();
                    if (reserved + r.footprint > c->pool.budget) {
                        if (inflight.empty() && batch.empty()) {   // can never fit
                            c->serve_error = "footprint exceeds the HBM budget";
                            finish_ticket(c, r.ticket, 3);
                            c->n_rej++;
                            c->queue.pop_front();
                            c->scv.notify_all();
                            continue;
                        }
                        break;
                    }
                    reserved += r.footprint;
                    c->status[r.ticket] = 1;
                    batch.push_back(r);
                    c->queue.pop_front();
                }
            }
            if (batch.empty()) {
                if (inflight.empty())
                    c->scv.wait_for(lk, std::chrono::milliseconds(2));
                else
                    lk.unlock();
            }
        }
        if (batch.empty()) {
            if (!inflight.empty()) std::this_thread::sleep_for(std::chrono::microseconds(100));
            continue;
        }
        // stage inputs, build one plan for the batch, launch on a free stream
        auto* f = new sister_ctx::Inflight();
        f->reqs = batch;
        f->stream = streams.back();
        streams.pop_back();
        std::vector<sister_rig> dev(batch.size());
        bool ok = true;
        int modes = 0;
        for (auto& r : batch) modes |= r.modes;
        bool no_memory = false;
        std::string fail_reason;
        {
            std::lock_guard<std::mutex> lk(c->mu);
            f->pinned.assign(batch.size(), nullptr);
            for (size_t i = 0; i < batch.size() && ok; i++) {
                const sister_rig& h = batch[i].rig;
                if (batch[i].device) {   // computed in place
                    dev[i] = h;
                    continue;
                }
                const size_t WH = (size_t)h.width * h.height;
                Block b;
                if (c->pool.get(io_bytes(WH), b) != 0) { ok = false; no_memory = true; break; }
                f->io.push_back(b);
                char* hp = c->pinned.get(io_bytes(WH));
                if (!hp) { ok = false; break; }
                f->pinned[i] = hp;
                // views into the pinned buffer (CPU), then one asynchronous copy
                for (int v = 0; v < 5; v++) {
                    char* dst = hp + v * align_up(WH);
                    if (h.stride == h.width) {
                        std::memcpy(dst, h.view[v], WH);
                    } else {
                        for (int y = 0; y < h.height; y++)
                            std::memcpy(dst + (size_t)y * h.width, h.view[v] + (size_t)y * h.stride, h.width);
                    }
                }
                (void)hipMemcpyAsync(b.ptr, hp, 5 * align_up(WH), hipMemcpyHostToDevice, f->stream);
                char* p = b.ptr;
                sister_rig d = h;
                d.stride = h.width;
                for (int v = 0; v < 5; v++) {
                    d.view[v] = (const uint8_t*)p;
                    p += align_up(WH);
                }
                d.out_multiview = (uint16_t*)p; p += align_up(WH * 2);
                d.out_horizontal = (uint16_t*)p; p += align_up(WH * 2);
                d.out_vertical = (uint16_t*)p; p += align_up(WH * 2);
                d.out_disp = (uint16_t*)p; p += align_up(WH * 2);
                d.out_subpix = (float*)p;
                dev[i] = d;
            }
            if (ok) {
                int rc = build_plan(c, dev.data(), (int)dev.size(), f->plan);
                if (rc == SISTER_ENOMEM) rc = build_plan(c, dev.data(), (int)dev.size(), f->plan, nullptr, true);
                ok = rc == SISTER_OK;
                no_memory = rc == SISTER_ENOMEM;
                if (!ok) fail_reason = plan_error;
            } else {
                fail_reason = "staging allocation failed";
            }
            if (ok) enqueue(c, f->plan, modes, f->stream, nullptr, nullptr);
        }
        // outputs of host requests back into their pinned staging (copied to
        // the caller's buffers when the batch retires)
        for (size_t i = 0; ok && i < batch.size(); i++) {
            if (batch[i].device) continue;
            const sister_rig& h = batch[i].rig;
            const sister_rig& d = dev[i];
            const size_t WH = (size_t)h.width * h.height;
            char* o = f->pinned[i] + 5 * align_up(WH);
            const size_t m2 = align_up(WH * 2);
            const int rm = batch[i].modes;
            if (h.out_multiview && (rm & 1)) (void)hipMemcpyAsync(o, d.out_multiview, WH * 2, hipMemcpyDeviceToHost, f->stream);
            if (h.out_horizontal && (rm & 2)) (void)hipMemcpyAsync(o + m2, d.out_horizontal, WH * 2, hipMemcpyDeviceToHost, f->stream);
            if (h.out_vertical && (rm & 4)) (void)hipMemcpyAsync(o + 2 * m2, d.out_vertical, WH * 2, hipMemcpyDeviceToHost, f->stream);
            if (h.out_disp && (rm & 1)) (void)hipMemcpyAsync(o + 3 * m2, d.out_disp, WH * 2, hipMemcpyDeviceToHost, f->stream);
            if (h.out_subpix && (rm & 1)) (void)hipMemcpyAsync(o + 4 * m2, d.out_subpix, WH * 4, hipMemcpyDeviceToHost, f->stream);
        }
        (void)hipEventCreateWithFlags(&f->done, hipEventDisableTiming);
        (void)hipEventRecord(f->done, f->stream);
        if (!ok) {
            (void)hipStreamSynchronize(f->stream);
            for (char* hp : f->pinned) c->pinned.put(hp);
            std::lock_guard<std::mutex> lk(c->mu);
            free_plan(c, f->plan);
            for (auto& b : f->io) c->pool.put(b);
            std::lock_guard<std::mutex> lk2(c->smu);
            if (no_memory && (!inflight.empty() || f->reqs.size() > 1)) {
                // wait for running batches to return memory, or retry smaller
                if (inflight.empty()) admit_cap = std::max<size_t>(1, f->reqs.size() / 2);
                for (auto it = f->reqs.rbegin(); it != f->reqs.rend(); ++it) {
                    c->status[it->ticket] = 0;
                    reserved -= it->footprint;
                    c->queue.push_front(*it);
                }
            } else {
                c->serve_error = fail_reason;
                for (auto& r : f->reqs) { finish_ticket(c, r.ticket, 3); reserved -= r.footprint; c->n_rej++; }
            }
            c->scv.notify_all();
            (void)hipEventDestroy(f->done);
            streams.push_back(f->stream);
            delete f;
            continue;
        }
        admit_cap = (size_t)-1;
        inflight.push_back(f);
    }
    for (auto s : streams) (void)hipStreamDestroy(s);
}

extern "C" {

int sister_serve_start(sister_ctx* c, int max_inflight)
{
    if (!c) return SISTER_EINVAL;
    std::lock_guard<std::mutex> lk(c->smu);
    if (c->serving) return SISTER_OK;
    c->max_inflight = std::max(1, std::min(max_inflight, 16));
    c->stop = false;
    c->serving = true;
    c->lat_ms.clear();
    c->lat_next = 0;
    c->n_sub = c->n_done = c->n_rej = 0;
    c->have_first = false;
    c->server = std::thread(serve_loop, c);
    return SISTER_OK;
}

static int submit(sister_ctx* c, const sister_rig* r, int modes, uint64_t* ticket, bool device)
{
    if (!c || !r || !ticket) return fail(c, SISTER_EINVAL, "bad arguments");
    if (!valid_dims(r->width, r->height, r->disp_count)) return fail(c, SISTER_EINVAL, "invalid rig shape or D");
    if (modes <= 0 || modes > 7) return fail(c, SISTER_EINVAL, "bad modes");
    if (r->stride < r->width) return fail(c, SISTER_EINVAL, "stride < width");
    for (int v = 0; v < 5; v++)
        if (!r->view[v]) return fail(c, SISTER_EINVAL, "null view");
    if (device) {
        const int rc = check_modes(c, r, 1, modes);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(c->smu);
    if (!c->serving) return fail(c, SISTER_EINVAL, "serving not started");
    sister_ctx::Request q;
    q.ticket = c->next_ticket++;
    q.rig = *r;
    q.modes = modes;
    q.device = device;
    q.t_submit = std::chrono::steady_clock::now();
    q.footprint = rig_footprint(*r, device);
    if (q.footprint > c->pool.budget) {
        c->n_rej++;
        return fail(c, SISTER_EBUSY, "rig footprint exceeds the HBM budget");
    }
    if (!c->have_first) { c->t_first = q.t_submit; c->have_first = true; }
    c->status[q.ticket] = 0;
    c->queue.push_back(q);
    c->n_sub++;
    *ticket = q.ticket;
    c->scv.notify_all();
    return SISTER_OK;
}

int sister_submit(sister_ctx* c, const sister_rig* r, int modes, uint64_t* ticket)
{
    return submit(c, r, modes, ticket, false);
}

int sister_submit_device(sister_ctx* c, const sister_rig* r, int modes, uint64_t* ticket)
{
    return submit(c, r, modes, ticket, true);
}

int sister_poll(sister_ctx* c, uint64_t t, int* done)
{
    if (!c || !done) return SISTER_EINVAL;
    std::lock_guard<std::mutex> lk(c->smu);
    auto it = c->status.find(t);
    if (it == c->status.end()) return fail(c, SISTER_ENOTFOUND, "unknown ticket (never issued, or forgotten)");
    *done = it->second >= 2;
    return it->second == 3 ? fail(c, SISTER_EDEVICE, "request failed: " + c->serve_error) : SISTER_OK;
}

int sister_wait(sister_ctx* c, uint64_t t)
{
    if (!c) return SISTER_EINVAL;
    std::unique_lock<std::mutex> lk(c->smu);
    auto it = c->status.find(t);
    if (it == c->status.end()) return fail(c, SISTER_ENOTFOUND, "unknown ticket");
    c->scv.wait(lk, [&] { return c->status[t] >= 2; });
    return c->status[t] == 3 ? fail(c, SISTER_EDEVICE, "request failed: " + c->serve_error) : SISTER_OK;
}

int sister_serve_stop(sister_ctx* c)
{
    if (!c) return SISTER_EINVAL;
    {
        std::lock_guard<std::mutex> lk(c->smu);
        if (!c->serving) return SISTER_OK;
        c->stop = true;
    }
    c->scv.notify_all();
    if (c->server.joinable()) c->server.join();
    std::lock_guard<std::mutex> lk(c->smu);
    c->serving = false;
    return SISTER_OK;
}

int sister_stats_get(sister_ctx* c, sister_stats* s)
{
    if (!c || !s) return SISTER_EINVAL;
    std::memset(s, 0, sizeof(*s));
    std::vector<double> v;
    {
        std::lock_guard<std::mutex> lk(c->smu);
        s->submitted = c->n_sub;
        s->completed = c->n_done;
        s->rejected = c->n_rej;
        v = c->lat_ms;
        if (c->have_first && c->n_done > 0) {
            double span = std::chrono::duration<double>(c->t_last - c->t_first).count();
            if (span > 0) s->rigs_per_s = c->n_done / span;
        }
    }
    // percentiles outside the lock the serving thread needs
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { size_t i = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5)); return v[i]; };
        s->p50_ms = pct(0.50);
        s->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s->mean_ms = sum / v.size();
    }
    std::lock_guard<std::mutex> lk(c->mu);   // the arena is mutated under mu
    s->hbm_in_use = c->pool.in_use;
    s->hbm_peak = c->pool.peak;
    return SISTER_OK;
}

}  // extern "C"
