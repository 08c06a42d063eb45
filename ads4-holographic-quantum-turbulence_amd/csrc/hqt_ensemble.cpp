// Batched ensemble: many independent evolutions (contexts of uneven nx, ny, nz, mu, drive and
// precision) advanced together.  Each kernel of the step sequence is launched once per kernel
// class (same template instance: precision and Q = Mp/64 for the line passes, nz for the
// radial kernels) over all members of that class; a member owns a contiguous block range of
// the launch and reads its own descriptor (buffers, sizes, flags) from a device array.  The
// members are split into up to kLanes lanes of about equal point counts, members of one shape
// together where the balance allows (their launches merge); each lane runs its own class launches
// on its own stream.  The lanes share no member, so they meet only at the fork and the join
// around a replay, and a lane's steps need not wait for the other lanes: the steps are captured
// into two hipGraphs, one step and kGraphSteps steps, and step_n replays the long one while it
// can, which overlaps one lane's small or tail launches with another lane's work across steps.
//
// The per-member arithmetic is that of hqt_step_n on the same context (same kernels, same
// operands), so results are identical to stepping each member alone.
#include "hqt_internal.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <utility>
#include <vector>

namespace {

enum LKind { K_LINE, K_ZSTAGE, K_ZPOST, K_WPART, K_WAPPLY, K_NAN };

struct Launch {
    int kind = 0;
    int dir = 0, lkind = 0, Q = 1, nt = 1, ngroups = 1, nz = 0, fp32 = 0, wpe = 2, blocks = 0, nmem = 0, fold = 0, order = 0;
    int lane = 0;                    // stream: 0 = the ensemble's own, l > 0 = side stream l - 1
    size_t lds = 0;
    size_t desc_off = 0, boff_off = 0;   // offsets into the device plan buffer
};

// Host image of the device plan: descriptor arrays and block offsets, 256-byte aligned.
struct PlanImage {
    std::vector<unsigned char> bytes;
    size_t put(const void* p, size_t n) {
        const size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + n);
        std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};

constexpr int kLanes = 4;            // the ensemble's own stream + 3 side streams
constexpr int kGraphSteps = 8;       // steps per replay of the long graph

}  // namespace

struct hqt_ensemble {
    std::vector<hqt_ctx*> mem;       // in the caller's order
    hipStream_t stream = nullptr;
    hipStream_t side[kLanes - 1] = {};
    hipEvent_t fork = nullptr, join[kLanes - 1] = {};
    int lanes = 1;                   // lanes in use
    std::vector<Launch> plan;
    unsigned char* dplan = nullptr;
    hipGraphExec_t graph = nullptr;  // one step
    hipGraphExec_t graphk = nullptr; // kGraphSteps steps, captured on the first step_n that can use it
    bool graphs = false;             // replay graphs (else launch directly)
};

namespace {

inline int efail(int code, const std::string& msg) { return hqt_fail(code, msg); }

// Members grouped by key, largest first inside each group (large members' blocks are
// dispatched first, which shortens the tail of a launch).
template <class Key, class KeyFn, class Pred>
std::map<Key, std::vector<hqt_ctx*>> classes(const std::vector<hqt_ctx*>& mem, KeyFn key, Pred use) {
    std::map<Key, std::vector<hqt_ctx*>> out;
    for (hqt_ctx* c : mem)
        if (use(c)) out[key(c)].push_back(c);
    for (auto& kv : out)
        std::stable_sort(kv.second.begin(), kv.second.end(),
                         [](const hqt_ctx* a, const hqt_ctx* b) { return a->npts > b->npts; });
    return out;
}

template <class D, class DescFn, class BlocksFn>
Launch make_launch(PlanImage& img, int kind, const std::vector<hqt_ctx*>& ms, DescFn desc, BlocksFn blocks) {
    Launch L;
    L.kind = kind;
    L.nmem = (int)ms.size();
    std::vector<D> d;
    std::vector<int32_t> boff(1, 0);
    for (hqt_ctx* c : ms) {
        d.push_back(desc(c));
        boff.push_back(boff.back() + blocks(c));
    }
    L.blocks = boff.back();
    L.desc_off = img.put(d.data(), d.size() * sizeof(D));
    L.boff_off = img.put(boff.data(), boff.size() * sizeof(int32_t));
    return L;
}

void add_line(std::vector<Launch>& plan, PlanImage& img, const std::vector<hqt_ctx*>& mem, int pass, int dir, int name,
              bool (*use)(const hqt_ctx*)) {
    auto cls = classes<std::tuple<int, int, int, int>>(
        mem, [&](hqt_ctx* c) {
            return std::make_tuple(pass_fp32(c, pass), c->hp[pass].Mp / 64, c->hp[pass].lb / 16, c->hp[pass].fold);
        },
        [&](hqt_ctx* c) { return use(c); });
    for (auto& kv : cls) {
        Launch L = make_launch<LPass>(img, K_LINE, kv.second, [&](hqt_ctx* c) { return c->hp[pass]; },
                                      [&](hqt_ctx* c) { return line_blocks(c->hp[pass]); });
        L.dir = dir;
        L.fp32 = std::get<0>(kv.first);
        L.Q = std::get<1>(kv.first);
        L.nt = std::get<2>(kv.first);
        L.fold = std::get<3>(kv.first);
        // block order: the members' tuned choice by majority (ties: XCD order); it never changes a result
        int votes = 0;
        for (hqt_ctx* c : kv.second) votes += c->hp[pass].xcd ? 1 : -1;
        L.order = L.fold && votes >= 0 ? 1 : 0;
        L.lkind = (name == KN_LX_FILT || name == KN_LY_FILT) ? 2 : (name == KN_LX_CONS || name == KN_LY_CONS) ? 1 : 0;
        for (hqt_ctx* c : kv.second) {
            L.ngroups = std::max(L.ngroups, c->hp[pass].ngroups);
            L.lds = std::max(L.lds, line_pass_lds_bytes(c->hp[pass], pass_fp32(c, pass)));
        }
        plan.push_back(L);
    }
}

bool all_members(const hqt_ctx*) { return true; }
bool with_winding(const hqt_ctx* c) { return c->p.winding != 0; }
bool with_filter(const hqt_ctx* c) { return c->p.filter != 0; }
bool without_filter(const hqt_ctx* c) { return c->p.filter == 0; }

template <class D, class DescFn, class BlocksFn>
void add_z(std::vector<Launch>& plan, PlanImage& img, const std::vector<hqt_ctx*>& mem, int kind, bool by_nz,
           bool (*use)(const hqt_ctx*), DescFn desc, BlocksFn blocks) {
    auto cls = classes<std::tuple<int, int, int>>(
        mem, [&](hqt_ctx* c) { return std::make_tuple(c->fp32, by_nz ? c->nz : 0, kind == K_ZSTAGE ? c->zwpe : 2); },
        [&](hqt_ctx* c) { return use(c); });
    for (auto& kv : cls) {
        Launch L = make_launch<D>(img, kind, kv.second, desc, blocks);
        L.fp32 = std::get<0>(kv.first);
        L.nz = std::get<1>(kv.first);
        L.wpe = std::get<2>(kv.first);
        plan.push_back(L);
    }
}

// The step sequence of enqueue_step (hqt_host.cpp), class by class, for the members M of one lane.
void build_plan(hqt_ensemble* e, PlanImage& img, const std::vector<hqt_ctx*>& M, int lane) {
    auto& P = e->plan;
    const size_t first = P.size();
    auto zb = [](hqt_ctx* c) { return zblocks(c->ncol); };
    for (int s = 0; s < 4; ++s) {
        add_line(P, img, M, PX1 + s, 0, KN_LX_STAGE, all_members);
        add_line(P, img, M, PY1 + s, 1, KN_LY_STAGE, all_members);
        add_z<ZStage>(P, img, M, K_ZSTAGE, true, all_members, [s](hqt_ctx* c) { return c->zs[s]; }, zb);
    }
    add_line(P, img, M, PXC, 0, KN_LX_CONS, all_members);
    add_line(P, img, M, PYC, 1, KN_LY_CONS, all_members);
    add_z<ZPost>(P, img, M, K_ZPOST, true, all_members, [](hqt_ctx* c) { return c->zp; }, zb);
    add_z<Wind>(P, img, M, K_WPART, false, with_winding, [](hqt_ctx* c) { return c->wind; },
                [](hqt_ctx* c) { return c->ny; });
    add_z<Wind>(P, img, M, K_WAPPLY, false, with_winding, [](hqt_ctx* c) { return c->wind; },
                [](hqt_ctx* c) { return wind_apply_blocks(c->ncol, c->nz); });
    add_line(P, img, M, PXF, 0, KN_LX_FILT, with_filter);
    add_line(P, img, M, PYF, 1, KN_LY_FILT, with_filter);
    add_z<NanCheck>(P, img, M, K_NAN, false, without_filter, [](hqt_ctx* c) { return c->nan; },
                    [](hqt_ctx* c) { return nancheck_blocks(c->nvol); });
    for (size_t i = first; i < P.size(); ++i) P[i].lane = lane;
}

// Lanes: members of one shape (precision, nx, ny, nz) form a group, split into pieces of at most a
// lane's share of the points; the pieces by decreasing point count, each to the lane with the
// fewest points so far.
void build_lanes(hqt_ensemble* e, PlanImage& img) {
    std::map<std::tuple<int, int, int, int>, std::vector<hqt_ctx*>> shape;
    int64_t total = 0;
    for (hqt_ctx* c : e->mem) {
        shape[std::make_tuple(c->fp32, c->nx, c->ny, c->nz)].push_back(c);
        total += c->npts;
    }
    const int nl = std::min<int>(kLanes, (int)e->mem.size());
    const int64_t share = (total + nl - 1) / nl;
    std::vector<std::pair<int64_t, std::vector<hqt_ctx*>>> pieces;
    for (auto& kv : shape) {
        std::vector<hqt_ctx*> cur;
        int64_t pts = 0;
        for (hqt_ctx* c : kv.second) {
            if (!cur.empty() && pts + c->npts > share) {
                pieces.emplace_back(pts, cur);
                cur.clear();
                pts = 0;
            }
            cur.push_back(c);
            pts += c->npts;
        }
        pieces.emplace_back(pts, cur);
    }
    std::stable_sort(pieces.begin(), pieces.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<std::vector<hqt_ctx*>> lane(nl);
    std::vector<int64_t> load(nl, 0);
    for (auto& pc : pieces) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lane[l].insert(lane[l].end(), pc.second.begin(), pc.second.end());
        load[l] += pc.first;
    }
    e->lanes = nl;
    for (int l = 0; l < nl; ++l) build_plan(e, img, lane[l], l);
}

hipError_t issue(const hqt_ensemble* e, const Launch& L, hipStream_t s) {
    const void* d = e->dplan + L.desc_off;
    const Members M{reinterpret_cast<const int32_t*>(e->dplan + L.boff_off), L.nmem};
    switch (L.kind) {
        case K_LINE:
            return launch_line_pass(static_cast<const LPass*>(d), M, L.blocks, L.ngroups, L.lds, L.Q, L.nt, L.dir, L.fp32,
                                    L.lkind, s, L.fold, L.order);
        case K_ZSTAGE: return launch_zstage(static_cast<const ZStage*>(d), M, L.blocks, L.nz, L.fp32, L.wpe, s);
        case K_ZPOST: return launch_zpost(static_cast<const ZPost*>(d), M, L.blocks, L.nz, L.fp32, s);
        case K_WPART: return launch_wind_partial(static_cast<const Wind*>(d), M, L.blocks, s);
        case K_WAPPLY: return launch_wind_apply(static_cast<const Wind*>(d), M, L.blocks, L.fp32, s);
        case K_NAN: return launch_nancheck(static_cast<const NanCheck*>(d), M, L.blocks, L.fp32, s);
    }
    return hipErrorInvalidValue;
}

// nsteps steps of every lane between one fork and one join.
int enqueue(const hqt_ensemble* e, int nsteps) {
    auto lane = [&](int l) { return l == 0 ? e->stream : e->side[l - 1]; };
    hipError_t r = hipSuccess;
    if (e->lanes > 1) {
        r = hipEventRecord(e->fork, e->stream);
        for (int l = 1; l < e->lanes && r == hipSuccess; ++l) r = hipStreamWaitEvent(e->side[l - 1], e->fork, 0);
    }
    for (int i = 0; i < nsteps; ++i)
        for (size_t k = 0; k < e->plan.size() && r == hipSuccess; ++k) r = issue(e, e->plan[k], lane(e->plan[k].lane));
    for (int l = 1; l < e->lanes && r == hipSuccess; ++l) {
        r = hipEventRecord(e->join[l - 1], e->side[l - 1]);
        if (r == hipSuccess) r = hipStreamWaitEvent(e->stream, e->join[l - 1], 0);
    }
    if (r != hipSuccess) return efail(HQT_ERR_HIP, std::string("ensemble launch: ") + hipGetErrorString(r));
    return HQT_OK;
}

int build_graph(hqt_ensemble* e, int nsteps, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(e, nsteps);
    const hipError_t r = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (r != hipSuccess) return efail(HQT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
    const hipError_t ri = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ri != hipSuccess) { *out = nullptr; return efail(HQT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ri)); }
    return HQT_OK;
}

void release(hqt_ensemble* e) {
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    if (e->graphk) (void)hipGraphExecDestroy(e->graphk);
    if (e->dplan) (void)hipFree(e->dplan);
    for (int l = 0; l < kLanes - 1; ++l) {
        if (e->side[l]) { (void)hipStreamSynchronize(e->side[l]); (void)hipStreamDestroy(e->side[l]); }
        if (e->join[l]) (void)hipEventDestroy(e->join[l]);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    for (hqt_ctx* c : e->mem)
        if (c && c->owner == e->stream) c->owner = nullptr;
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // namespace

extern "C" {

int hqt_ensemble_create(hqt_ctx* const* members, int32_t n, int32_t flags, hqt_ensemble** out) {
    if (!out || !members || n < 1) return efail(HQT_ERR_ARG, "hqt_ensemble_create: need n >= 1 members and out");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        hqt_ctx* c = members[i];
        if (!c) return efail(HQT_ERR_ARG, "hqt_ensemble_create: NULL member " + std::to_string(i));
        if (c->owner) return efail(HQT_ERR_STATE, "hqt_ensemble_create: member " + std::to_string(i) + " is in another ensemble");
        if (!c->have_state) return efail(HQT_ERR_STATE, "hqt_ensemble_create: member " + std::to_string(i) + " has no state");
        for (int j = 0; j < i; ++j)
            if (members[j] == c) return efail(HQT_ERR_ARG, "hqt_ensemble_create: member listed twice");
    }
    hqt_ensemble* e = new (std::nothrow) hqt_ensemble();
    if (!e) return efail(HQT_ERR_OOM, "host allocation failed");
    e->mem.assign(members, members + n);
    hipError_t r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r != hipSuccess) { e->stream = nullptr; release(e); return efail(HQT_ERR_HIP, "hipStreamCreate"); }
    for (int l = 0; l < kLanes - 1; ++l) {
        // default priority, like the ensemble's own stream
        if (hipStreamCreateWithFlags(&e->side[l], hipStreamNonBlocking) != hipSuccess) e->side[l] = nullptr;
        if (hipEventCreateWithFlags(&e->join[l], hipEventDisableTiming) != hipSuccess) e->join[l] = nullptr;
        if (!e->side[l] || !e->join[l]) { release(e); return efail(HQT_ERR_HIP, "ensemble side stream/event"); }
    }
    if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
        e->fork = nullptr; release(e); return efail(HQT_ERR_HIP, "ensemble fork event");
    }
    PlanImage img;
    build_lanes(e, img);
    r = hipMalloc(&e->dplan, img.bytes.size());
    if (r != hipSuccess) { e->dplan = nullptr; release(e); return efail(HQT_ERR_OOM, "hipMalloc(ensemble plan)"); }
    r = hipMemcpy(e->dplan, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice);
    if (r != hipSuccess) { release(e); return efail(HQT_ERR_HIP, "hipMemcpy(ensemble plan)"); }
    for (hqt_ctx* c : e->mem) {
        if (int rc = hqt_sync_streams(c)) { release(e); return rc; }
        c->owner = e->stream;
    }
    e->graphs = !(flags & HQT_NO_GRAPH);
    if (e->graphs) {
        if (int rc = build_graph(e, 1, &e->graph)) { release(e); return rc; }
    }
    *out = e;
    return HQT_OK;
}

int hqt_ensemble_destroy(hqt_ensemble* e) {
    if (e) release(e);
    return HQT_OK;
}

int hqt_ensemble_launches(hqt_ensemble* e, int32_t* n) {
    if (!e || !n) return efail(HQT_ERR_ARG, "NULL argument");
    *n = (int32_t)e->plan.size();
    return HQT_OK;
}

int hqt_ensemble_step_n(hqt_ensemble* e, int32_t nsteps) {
    if (!e) return efail(HQT_ERR_ARG, "NULL ensemble");
    if (nsteps < 0) return efail(HQT_ERR_ARG, "nsteps < 0");
    for (hqt_ctx* c : e->mem) HIPC(hipStreamSynchronize(c->stream));   // members' own queued work first
    if (!e->graphs) return nsteps > 0 ? enqueue(e, nsteps) : HQT_OK;
    if (nsteps >= kGraphSteps && !e->graphk) {
        if (int rc = build_graph(e, kGraphSteps, &e->graphk)) return rc;
    }
    for (; nsteps >= kGraphSteps; nsteps -= kGraphSteps) HIPC(hipGraphLaunch(e->graphk, e->stream));
    for (; nsteps > 0; --nsteps) HIPC(hipGraphLaunch(e->graph, e->stream));
    return HQT_OK;
}

int hqt_ensemble_sync(hqt_ensemble* e, int32_t* nan_steps) {
    if (!e) return efail(HQT_ERR_ARG, "NULL ensemble");
    HIPC(hipStreamSynchronize(e->stream));
    int bad = 0, first = -1;
    for (size_t i = 0; i < e->mem.size(); ++i) {
        int32_t f[2];
        HIPC(hipMemcpy(f, e->mem[i]->flags, sizeof(f), hipMemcpyDeviceToHost));
        if (nan_steps) nan_steps[i] = f[0];
        if (f[0] != 0) { ++bad; if (first < 0) first = (int)i; }
    }
    if (bad)
        return efail(HQT_ERR_NAN, std::to_string(bad) + " ensemble member(s) hit NaN (first: member " + std::to_string(first) + ")");
    return HQT_OK;
}

}  // extern "C"
