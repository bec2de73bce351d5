// The C boundary of libswp_core (include/sweep_amd.h): entry table,
// dispatch, stream, layout probe, timing.  Every function catches everything
// and answers with a return code plus a message, like the reference's
// guarded() (csrc/cuda/common/capi.cu:230-242).
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <string>

#include "acoustic2d.h"
#include "acoustic3d.h"
#include "elastic2d.h"
#include "records.h"
#include "revolve.h"
#include "unet.h"

namespace swp {

namespace {
thread_local hipStream_t t_stream = nullptr;
std::atomic<int> g_tuning[SWP_TUNE_COUNT];

// ---- time-loop graphs
// A call's whole launch sequence (every step of the forward or backward
// loop: a few kernels per step, thousands of steps) is captured once into a
// HIP graph and replayed when the same call comes again: same entry,
// device, problem (pointers and scalars, the host coefficient values),
// io pointers, tuning and kernel-family settings.  Replays are bit-identical
// to the eager launches (the same kernels with the same arguments in the same
// stream order).  A problem whose buffers move between calls never hits: after
// kGraphMissLimit captures of one shape without a replay it runs eagerly.
std::atomic<int> g_graph_mode{-1};   // -1: SWP_GRAPHS (default on), 0 off, 1 on
std::mutex g_graph_mu;
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
};
struct ShapeStats {
    int captures = 0, replays = 0;
};
std::map<std::string, GraphEntry> g_graphs;
std::map<std::string, ShapeStats> g_graph_shapes;
uint64_t g_graph_clock = 0;
int64_t g_graph_counts[3] = {0, 0, 0};   // captures, replays, eager fallbacks
constexpr size_t kGraphCacheMax = 16;
constexpr int kGraphMissLimit = 3;
thread_local std::map<int, hipStream_t> t_capture_streams;

std::atomic<bool> g_census_on{false};
std::mutex g_census_mu;
std::map<std::string, std::pair<int64_t, int64_t>> g_census;   // site -> (launches, workgroups)

struct Entry {
    const char *name;
    int kind;        // 0 forward, 1 backward
    int equation;    // 0 acoustic2d, 1 acoustic3d, 2 elastic2d
    int memory;      // the memory mode a backward entry serves (0: forward)
};

// <C_NAME>_<op>, the names equation._C() resolves (equations/base.py:295-324).
const Entry kEntries[] = {
    {"acoustic2d_forward", 0, 0, 0},
    {"acoustic2d_backward", 1, 0, SWP_MEM_FULL},
    {"acoustic2d_backward_bs", 1, 0, SWP_MEM_BOUNDARY},
    {"acoustic2d_backward_ckpt", 1, 0, SWP_MEM_CKPT},
    {"acoustic2d_backward_recursive_ckpt", 1, 0, SWP_MEM_BINOMIAL},
    {"acoustic3d_forward", 0, 1, 0},
    {"acoustic3d_backward", 1, 1, SWP_MEM_FULL},
    {"acoustic3d_backward_bs", 1, 1, SWP_MEM_BOUNDARY},
    {"acoustic3d_backward_ckpt", 1, 1, SWP_MEM_CKPT},
    {"acoustic3d_backward_recursive_ckpt", 1, 1, SWP_MEM_BINOMIAL},
    {"elastic2d_forward", 0, 2, 0},
    {"elastic2d_backward", 1, 2, SWP_MEM_FULL},
    {"elastic2d_backward_bs", 1, 2, SWP_MEM_BOUNDARY},
    {"elastic2d_backward_ckpt", 1, 2, SWP_MEM_CKPT},
    {"elastic2d_backward_recursive_ckpt", 1, 2, SWP_MEM_BINOMIAL},
};
constexpr int kNumEntries = int(sizeof(kEntries) / sizeof(kEntries[0]));

void copy_err(const char *msg, char *err, int cap)
{
    if (!err || cap <= 0) return;
    std::strncpy(err, msg, size_t(cap) - 1);
    err[cap - 1] = '\0';
}

template <class F>
int guarded(char *err, int cap, F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        copy_err(e.what(), err, cap);
    } catch (...) {
        copy_err("unknown error in libswp_core", err, cap);
    }
    return 1;
}

const Entry &entry_at(int entry)
{
    SWP_CHECK(entry >= 0 && entry < kNumEntries, "entry ", entry, " out of range [0, ", kNumEntries, ")");
    return kEntries[entry];
}

void check_problem(const Entry &e, const swp_problem &p)
{
    SWP_CHECK(p.B >= 1 && p.nt >= 0, "problem: need B >= 1 and nt >= 0");
    SWP_CHECK(p.nsrc >= 0 && p.nrec >= 0, "problem: negative point count");
    SWP_CHECK(p.nsrc == 0 || (p.source && p.sources_loc), "problem: sources missing");
    SWP_CHECK(p.nrec == 0 || p.receivers_loc, "problem: receivers missing");
    if (e.kind == 1)
        SWP_CHECK(p.memory == e.memory, e.name, ": the forward ran with memory mode ", p.memory,
                  ", this entry serves mode ", e.memory);
}

}  // namespace

hipStream_t stream() { return t_stream; }

int tuning(int key, const char *env, int def)
{
    const int v = (key >= 0 && key < SWP_TUNE_COUNT) ? g_tuning[key].load() : 0;
    if (v > 0) return v;
    const char *e = env ? std::getenv(env) : nullptr;
    return e ? std::max(1, std::atoi(e)) : def;
}

bool eq_available(int eq) { return eq == 0 || (eq == 1 && a3d::available()) || (eq == 2 && e2d::available()); }

bool census_on() { return g_census_on.load(std::memory_order_relaxed); }

void census_add(const char *site, int64_t blocks)
{
    if (!census_on()) return;
    std::lock_guard<std::mutex> lk(g_census_mu);
    auto &e = g_census[site];
    e.first += 1;
    e.second += blocks;
}

namespace {

// Every launch of one call, on the thread's stream.
void run_entry(const Entry &e, const swp_problem &p, void *io)
{
    if (e.kind == 0) {
        auto *f = static_cast<swp_forward_io *>(io);
        if (e.equation == 0) a2d::forward(a2d::make_ctx(p), p, f->record);
        if (e.equation == 1) a3d::forward(p, f->record);
        if (e.equation == 2) e2d::forward(p, f->record);
    } else {
        auto *b = static_cast<swp_backward_io *>(io);
        if (e.equation == 0) a2d::backward(a2d::make_ctx(p), p, *b);
        if (e.equation == 1) a3d::backward(p, *b);
        if (e.equation == 2) e2d::backward(p, *b);
    }
}

bool graphs_on()
{
    int m = g_graph_mode.load();
    if (m < 0) {
        const char *e = std::getenv("SWP_GRAPHS");
        m = (e && std::string(e) == "0") ? 0 : 1;
    }
    return m == 1;
}

template <class T>
void put(std::string &k, const T &v)
{
    k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

void put_env(std::string &k, const char *name)
{
    const char *v = std::getenv(name);
    k += name;
    k += '=';
    k += v ? v : "";
    k += ';';
}

// (shape key, full key) of a call: everything the launches depend on, field
// by field (no struct padding).  The shape key leaves out the buffer
// addresses.
std::pair<std::string, std::string> graph_keys(int entry, int dev, const swp_problem &p, const void *io, int kind)
{
    std::string k;
    for (int v : {entry, dev, p.ndim, p.nz, p.ny, p.nx, p.B, p.M, p.nt, p.abcn, p.model_batched, p.nsrc, p.nrec,
                  p.memory, p.ckpt_interval, p.flags, p.nsrc_fields, p.nrec_fields})
        put(k, v);
    for (int v : p.pad) put(k, v);
    for (int v : p.src_fields) put(k, v);
    for (int v : p.rec_fields) put(k, v);
    for (float v : {p.dt, p.dz, p.dy, p.dx}) put(k, v);
    put(k, p.tape_bytes);
    put(k, p.workspace_bytes);
    for (int i = 0; i <= p.M && i < 16; ++i) {
        put(k, p.lap_coes ? p.lap_coes[i] : 0.f);
        put(k, p.grad_coes ? p.grad_coes[i] : 0.f);
    }
    for (int i = 0; i < SWP_TUNE_COUNT; ++i) put(k, g_tuning[i].load());
    for (const char *v : {"SWP_KERNELS", "SWP_SHELL3", "SWP_TZ", "SWP_TZ_BAND", "SWP_TZ_RECON", "SWP_TZ3",
                          "SWP_REGION_WAVES"})
        put_env(k, v);
    std::string full = k;
    for (const void *v : {static_cast<const void *>(p.source), static_cast<const void *>(p.sources_loc),
                          static_cast<const void *>(p.receivers_loc), static_cast<const void *>(p.tape),
                          static_cast<const void *>(p.workspace)})
        put(full, v);
    for (const float *v : p.models) put(full, v);
    for (const float *v : p.pml) put(full, v);
    if (kind == 0) {
        put(full, static_cast<const swp_forward_io *>(io)->record);
    } else {
        const auto *b = static_cast<const swp_backward_io *>(io);
        put(full, b->adjoint_source);
        put(full, b->grad_source);
        for (const float *v : b->grad_models) put(full, v);
        put(full, b->illum_source);
        put(full, b->illum_receiver);
    }
    return {k, full};
}

hipStream_t capture_stream(int dev)
{
    auto it = t_capture_streams.find(dev);
    if (it != t_capture_streams.end()) return it->second;
    hipStream_t s = nullptr;
    SWP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    t_capture_streams[dev] = s;
    return s;
}

// Run the call through the graph cache: replay a captured graph of the same
// call, or capture this one (on a side stream, nothing runs while capturing)
// and launch it.  Returns false when the call should run eagerly instead.
// Time steps a call runs (re-runs included): a captured graph holds a few
// kernel nodes per step in one dependency chain, and a chain of ~10^5 nodes
// (a binomial reversal with one snapshot slot re-runs O(nt^2) steps) is more
// than graph instantiation handles, so such calls run eagerly.
constexpr int64_t kGraphMaxSteps = 4096;

int64_t call_steps(const Entry &e, const swp_problem &p)
{
    if (e.kind == 0) return p.nt;
    if (p.memory == SWP_MEM_CKPT) return 2 * int64_t(p.nt);
    if (p.memory == SWP_MEM_BINOMIAL) return int64_t(revolve_schedule(p.nt, std::max(1, p.ckpt_interval)).size());
    return p.nt;
}

bool run_graphed(int entry, const Entry &e, const swp_problem &p, void *io)
{
    if (!(p.flags & SWP_FLAG_GRAPH) || !graphs_on() || timer().on || census_on()) return false;
    if (call_steps(e, p) > kGraphMaxSteps) return false;
    int dev = 0;
    SWP_HIP(hipGetDevice(&dev));
    const auto keys = graph_keys(entry, dev, p, io, e.kind);
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        auto it = g_graphs.find(keys.second);
        if (it != g_graphs.end()) {
            it->second.used = ++g_graph_clock;
            g_graph_shapes[keys.first].replays += 1;
            g_graph_counts[1] += 1;
            exec = it->second.exec;
        } else {
            const ShapeStats &st = g_graph_shapes[keys.first];
            if (st.replays == 0 && st.captures >= kGraphMissLimit) {
                g_graph_counts[2] += 1;
                return false;
            }
        }
    }
    if (!exec) {
        const hipStream_t user = stream();
        const hipStream_t cap = capture_stream(dev);
        hipGraph_t graph = nullptr;
        t_stream = cap;
        SWP_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            run_entry(e, p, io);
        } catch (...) {
            hipStreamEndCapture(cap, &graph);
            if (graph) hipGraphDestroy(graph);
            t_stream = user;
            throw;
        }
        t_stream = user;
        SWP_HIP(hipStreamEndCapture(cap, &graph));
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        SWP_CHECK(ie == hipSuccess, "hipGraphInstantiate: ", hipGetErrorString(ie));
        std::lock_guard<std::mutex> lk(g_graph_mu);
        if (g_graphs.size() >= kGraphCacheMax) {
            auto old = g_graphs.begin();
            for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
                if (it->second.used < old->second.used) old = it;
            hipGraphExecDestroy(old->second.exec);
            g_graphs.erase(old);
        }
        g_graphs[keys.second] = GraphEntry{exec, ++g_graph_clock};
        g_graph_shapes[keys.first].captures += 1;
        g_graph_counts[0] += 1;
    }
    SWP_HIP(hipGraphLaunch(exec, stream()));
    return true;
}

}  // namespace

}  // namespace swp

using namespace swp;

extern "C" {

SWP_CAPI int swp_core_abi_version(void) { return SWP_ABI_VERSION; }

#define SWP_FIELD(S, f) int64_t(offsetof(S, f))
static const int64_t kLayout[] = {
    int64_t(sizeof(swp_problem)),
    SWP_FIELD(swp_problem, ndim), SWP_FIELD(swp_problem, nz), SWP_FIELD(swp_problem, ny),
    SWP_FIELD(swp_problem, nx), SWP_FIELD(swp_problem, B), SWP_FIELD(swp_problem, M),
    SWP_FIELD(swp_problem, nt), SWP_FIELD(swp_problem, abcn), SWP_FIELD(swp_problem, pad),
    SWP_FIELD(swp_problem, dt), SWP_FIELD(swp_problem, dz), SWP_FIELD(swp_problem, dy),
    SWP_FIELD(swp_problem, dx), SWP_FIELD(swp_problem, lap_coes), SWP_FIELD(swp_problem, grad_coes),
    SWP_FIELD(swp_problem, pml), SWP_FIELD(swp_problem, models), SWP_FIELD(swp_problem, model_batched),
    SWP_FIELD(swp_problem, nsrc), SWP_FIELD(swp_problem, nrec), SWP_FIELD(swp_problem, source),
    SWP_FIELD(swp_problem, sources_loc), SWP_FIELD(swp_problem, receivers_loc),
    SWP_FIELD(swp_problem, memory), SWP_FIELD(swp_problem, ckpt_interval), SWP_FIELD(swp_problem, flags),
    SWP_FIELD(swp_problem, reserved), SWP_FIELD(swp_problem, nsrc_fields), SWP_FIELD(swp_problem, nrec_fields),
    SWP_FIELD(swp_problem, src_fields), SWP_FIELD(swp_problem, rec_fields),
    SWP_FIELD(swp_problem, tape), SWP_FIELD(swp_problem, tape_bytes),
    SWP_FIELD(swp_problem, workspace), SWP_FIELD(swp_problem, workspace_bytes),
    int64_t(sizeof(swp_forward_io)), SWP_FIELD(swp_forward_io, record),
    int64_t(sizeof(swp_backward_io)), SWP_FIELD(swp_backward_io, adjoint_source),
    SWP_FIELD(swp_backward_io, grad_source), SWP_FIELD(swp_backward_io, grad_models),
    SWP_FIELD(swp_backward_io, illum_source), SWP_FIELD(swp_backward_io, illum_receiver),
};
#undef SWP_FIELD

SWP_CAPI int64_t swp_layout_count(void) { return int64_t(sizeof(kLayout) / sizeof(kLayout[0])); }

SWP_CAPI int64_t swp_layout_probe(int64_t i)
{
    if (i < 0 || i >= swp_layout_count()) return -1;
    return kLayout[i];
}

SWP_CAPI int swp_entry_count(void) { return kNumEntries; }

SWP_CAPI const char *swp_entry_name(int entry)
{
    return (entry >= 0 && entry < kNumEntries) ? kEntries[entry].name : nullptr;
}

SWP_CAPI int swp_entry_kind(int entry)
{
    return (entry >= 0 && entry < kNumEntries) ? kEntries[entry].kind : -1;
}

SWP_CAPI int swp_entry_available(int entry)
{
    if (entry < 0 || entry >= kNumEntries) return -1;
    return eq_available(kEntries[entry].equation) ? 1 : 0;
}

SWP_CAPI const char *swp_core_device(void) { return "cuda"; }

SWP_CAPI void swp_set_stream(void *s) { t_stream = static_cast<hipStream_t>(s); }

SWP_CAPI void *swp_get_stream(void) { return static_cast<void *>(t_stream); }

SWP_CAPI int swp_query(int entry, const swp_problem *in, int64_t *workspace_bytes, int64_t *tape_bytes,
                       char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        SWP_CHECK(in != nullptr, "swp_query: null problem");
        const Entry &e = entry_at(entry);
        check_problem(e, *in);
        int64_t ws = 0, tp = 0;
        switch (e.equation) {
        case 0:
            ws = a2d::workspace_bytes(*in, e.kind);
            tp = a2d::tape_bytes(*in);
            break;
        case 1:
            ws = a3d::workspace_bytes(*in, e.kind);
            tp = a3d::tape_bytes(*in);
            break;
        case 2:
            ws = e2d::workspace_bytes(*in, e.kind);
            tp = e2d::tape_bytes(*in);
            break;
        }
        if (workspace_bytes) *workspace_bytes = ws;
        if (tape_bytes) *tape_bytes = tp;
    });
}


SWP_CAPI int swp_call(int entry, const swp_problem *in, void *io, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        SWP_CHECK(in != nullptr && io != nullptr, "swp_call: null problem or io");
        const Entry &e = entry_at(entry);
        SWP_CHECK(eq_available(e.equation), e.name, " is not available in this core");
        check_problem(e, *in);
        const swp_problem &p = *in;
        int64_t ws = 0, tp = 0;
        if (e.equation == 0) { ws = a2d::workspace_bytes(p, e.kind); tp = a2d::tape_bytes(p); }
        if (e.equation == 1) { ws = a3d::workspace_bytes(p, e.kind); tp = a3d::tape_bytes(p); }
        if (e.equation == 2) { ws = e2d::workspace_bytes(p, e.kind); tp = e2d::tape_bytes(p); }
        SWP_CHECK(p.workspace && p.workspace_bytes >= ws, e.name, ": workspace needs ", ws, " bytes, got ",
                  p.workspace_bytes);
        SWP_CHECK(p.memory == SWP_MEM_NONE || (p.tape && p.tape_bytes >= tp), e.name, ": tape needs ", tp,
                  " bytes, got ", p.tape_bytes);
        if (e.kind == 0)
            SWP_CHECK(p.nrec == 0 || static_cast<swp_forward_io *>(io)->record, e.name, ": record buffer missing");
        else
            SWP_CHECK(p.nrec == 0 || static_cast<swp_backward_io *>(io)->adjoint_source, e.name,
                      ": adjoint source missing");
        if (!run_graphed(entry, e, p, io)) run_entry(e, p, io);
    });
}

SWP_CAPI void swp_set_graphs(int mode)
{
    g_graph_mode.store(mode < 0 ? -1 : (mode ? 1 : 0));
}

SWP_CAPI int swp_graph_stats(int64_t *captures, int64_t *replays, int64_t *eager, int64_t *cached)
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    if (captures) *captures = g_graph_counts[0];
    if (replays) *replays = g_graph_counts[1];
    if (eager) *eager = g_graph_counts[2];
    if (cached) *cached = int64_t(g_graphs.size());
    return 0;
}

SWP_CAPI void swp_graph_clear(void)
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto &kv : g_graphs) hipGraphExecDestroy(kv.second.exec);
    g_graphs.clear();
    g_graph_shapes.clear();
    for (auto &c : g_graph_counts) c = 0;
}

SWP_CAPI void swp_set_timing(int stride)
{
    KTimer &t = timer();
    t.on = stride > 0;
    t.stride = stride > 0 ? stride : 1;
    if (t.on) t.reset();
}

SWP_CAPI int swp_last_kernel_stats(int which, double *ms_total, int64_t *launches)
{
    char err[8];
    return guarded(err, int(sizeof(err)), [&] {
        SWP_CHECK(which >= 0 && which < K_NCLASS, "kernel class out of range");
        double ms = 0.0;
        int64_t n = 0;
        timer().stats(which, &ms, &n);
        if (ms_total) *ms_total = ms;
        if (launches) *launches = n;
    });
}

SWP_CAPI void swp_census_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_census_mu);
    if (on) g_census.clear();
    g_census_on.store(on != 0);
}

SWP_CAPI int64_t swp_census(char *buf, int64_t cap)
{
    try {
        std::string s;
        {
            std::lock_guard<std::mutex> lk(g_census_mu);
            for (const auto &kv : g_census)
                s += kv.first + " " + std::to_string(kv.second.first) + " " + std::to_string(kv.second.second) + "\n";
        }
        if (buf && cap > 0) {
            const size_t n = std::min<size_t>(s.size(), size_t(cap - 1));
            std::memcpy(buf, s.data(), n);
            buf[n] = '\0';
        }
        return int64_t(s.size());
    } catch (...) {
        return -1;
    }
}

SWP_CAPI int64_t swp_revolve_schedule(int32_t nt, int32_t slots, int32_t *ops, int64_t cap)
{
    if (nt < 0 || slots < 1) return -1;
    try {
        const std::vector<RevOp> v = revolve_schedule(nt, slots);
        for (int64_t k = 0; ops && k < std::min<int64_t>(cap, int64_t(v.size())); ++k) {
            ops[2 * k] = v[size_t(k)].code;
            ops[2 * k + 1] = v[size_t(k)].arg;
        }
        return int64_t(v.size());
    } catch (...) {
        return -1;
    }
}

SWP_CAPI int64_t swp_revolve_spine(int32_t nt, int32_t slots, int32_t *pts, int64_t cap)
{
    if (nt < 0 || slots < 1) return -1;
    try {
        const std::vector<SpinePt> v = revolve_spine(nt, slots);
        for (int64_t k = 0; pts && k < std::min<int64_t>(cap, int64_t(v.size())); ++k) {
            pts[2 * k] = v[size_t(k)].slot;
            pts[2 * k + 1] = v[size_t(k)].step;
        }
        return int64_t(v.size());
    } catch (...) {
        return -1;
    }
}

SWP_CAPI int swp_set_tuning(int32_t key, int32_t value)
{
    if (key < 0 || key >= SWP_TUNE_COUNT || value < 0) return 1;
    g_tuning[key].store(value);
    return 0;
}

SWP_CAPI int32_t swp_get_tuning(int32_t key)
{
    if (key < 0 || key >= SWP_TUNE_COUNT) return -1;
    return g_tuning[key].load();
}

SWP_CAPI int swp_record_to_canonical(const float *raw, float *canonical, int32_t nfield, int32_t B,
                                     int32_t nrec, int32_t nt, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { rec::to_canonical(raw, canonical, nfield, B, nrec, nt); });
}

SWP_CAPI int swp_record_to_raw(const float *canonical, float *raw, int32_t nfield, int32_t B, int32_t nrec,
                               int32_t nt, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { rec::to_raw(canonical, raw, nfield, B, nrec, nt); });
}

SWP_CAPI int swp_l2_misfit(const float *pred, const float *obs, const float *mask, int32_t nfield, int32_t B,
                           int32_t nrec, int32_t nt, float *residual, double *loss, double *partials,
                           char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        rec::l2_misfit(pred, obs, mask, nfield, B, nrec, nt, residual, loss, partials);
    });
}

SWP_CAPI int swp_l2_misfit_noalloc(const float *pred, const float *obs, const float *mask, int32_t nfield, int32_t B,
                                   int32_t nrec, int32_t nt, float *residual, double *loss, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { rec::l2_misfit_noalloc(pred, obs, mask, nfield, B, nrec, nt, residual, loss); });
}

SWP_CAPI int swp_nn_conv_fwd(const uint16_t *in, int32_t N, int32_t H, int32_t W, int32_t cin, int32_t in_stride,
                             const uint16_t *w, int32_t cout, int32_t ks, const float *bias, int32_t relu,
                             void *out, int32_t out_stride, int32_t out_f32, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        nn::conv_fwd(in, N, H, W, cin, in_stride, w, cout, ks, bias, relu != 0, out, out_stride, out_f32 != 0);
    });
}

SWP_CAPI int64_t swp_nn_conv_wgrad_scratch(int32_t N, int32_t H, int32_t W, int32_t cin, int32_t cout, int32_t ks)
{
    if (N < 1 || H < 1 || W < 1 || cin < 1 || cout < 1 || (ks != 1 && ks != 3)) return -1;
    return nn::conv_wgrad_partials(N, H, W, cin, cout, ks);
}

SWP_CAPI int swp_nn_conv_wgrad(const uint16_t *in, int32_t N, int32_t H, int32_t W, int32_t cin, int32_t in_stride,
                               const uint16_t *dout, int32_t cout, int32_t dout_stride, int32_t ks, float *dw,
                               float *db, float *scratch, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        nn::conv_wgrad(in, N, H, W, cin, in_stride, dout, cout, dout_stride, ks, dw, db, scratch);
    });
}

SWP_CAPI int swp_nn_conv_flip(const uint16_t *w, int32_t cout, int32_t cin, int32_t ks, uint16_t *wt, char *err,
                              int err_cap)
{
    return guarded(err, err_cap, [&] { nn::conv_flip(w, cout, cin, ks, wt); });
}

SWP_CAPI int swp_nn_relu_bwd(uint16_t *d, int32_t d_stride, const uint16_t *act, int32_t act_stride, int64_t P,
                             int32_t C, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::relu_bwd(d, d_stride, act, act_stride, P, C); });
}

SWP_CAPI int swp_nn_pool2_fwd(const uint16_t *in, int32_t in_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                              uint16_t *out, int32_t out_stride, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::pool2_fwd(in, in_stride, N, H, W, C, out, out_stride); });
}

SWP_CAPI int swp_nn_pool2_bwd_acc(const uint16_t *dout, int32_t dout_stride, int32_t N, int32_t H, int32_t W,
                                  int32_t C, uint16_t *din, int32_t din_stride, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::pool2_bwd_acc(dout, dout_stride, N, H, W, C, din, din_stride); });
}

SWP_CAPI int swp_nn_up2_fwd(const uint16_t *in, int32_t in_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                            uint16_t *out, int32_t out_stride, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::up2_fwd(in, in_stride, N, H, W, C, out, out_stride); });
}

SWP_CAPI int swp_nn_up2_bwd(const uint16_t *dout, int32_t dout_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                            uint16_t *din, int32_t din_stride, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::up2_bwd(dout, dout_stride, N, H, W, C, din, din_stride); });
}

SWP_CAPI int swp_nn_pack_input(const float *const *planes, const float *scale, int32_t nsrc, int64_t P, uint16_t *x,
                               int32_t stride, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        SWP_CHECK(planes && scale, "pack_input: null plane or scale array");
        nn::pack_input(planes, scale, nsrc, P, x, stride);
    });
}

SWP_CAPI int swp_nn_mse(const float *out, int32_t out_stride, const float *target, const float *mask, int64_t P,
                        double count, uint16_t *dout, int32_t dout_stride, double *loss, double *partials,
                        char *err, int err_cap)
{
    return guarded(err, err_cap, [&] {
        nn::mse(out, out_stride, target, mask, P, count, dout, dout_stride, loss, partials);
    });
}

SWP_CAPI int swp_nn_adam(float *w, const float *g, float *m, float *v, uint16_t *wbf, int64_t n, float lr, float b1,
                         float b2, float eps, int32_t step, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::adam(w, g, m, v, wbf, n, lr, b1, b2, eps, step); });
}

SWP_CAPI int swp_nn_to_bf16(const float *src, uint16_t *dst, int64_t n, char *err, int err_cap)
{
    return guarded(err, err_cap, [&] { nn::to_bf16(src, dst, n); });
}

}  // extern "C"
