// The reference's C ABI, version 5 (csrc/core/capi.h:21-268 of
// DeepWave-KAUST/sweep), over libswp_core: the symbols its ctypes loader
// declares (backend/c/loader.py:52-79), the layout probe its ABI guard
// compares (loader.py:91-118), and sweep_call taking its ForwardInputCore /
// BackwardInputCore (input_core.h:10-132), translated onto one swp_problem
// and run by swp_call.  Entries are this core's fifteen
// <C_NAME>_<op> callables; the SEM, stepped-runner, boundary-session, visco
// FFT and nvCOMP functions are exported so the loader binds, and refuse or
// answer "none".
//
// Buffers: the core still never allocates.  It carves its wavefields out of
// one workspace and keeps the forward->backward state in one tape, so a
// caller binds, instead of the reference's per-field lists, the workspace as
// forward_workspace[0] (adjoint_workspace[0] for a backward) and the tape
// where the reference keeps that memory mode's state -- u_allt_out /
// u_forward, boundary_gpu[0] or checkpoints[0] -- sized by sweep_query_bytes.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "core.h"
#include "sweep_abi5.h"

namespace swp {
namespace abi5 {
namespace {

constexpr int64_t kRefLayout[] = {
#include "sweep_abi5_layout.inc"
};
constexpr int64_t kRefLayoutCount = int64_t(sizeof(kRefLayout) / sizeof(kRefLayout[0]));

#define S(T) int64_t(sizeof(T))
#define O(T, f) int64_t(offsetof(T, f))
// This mirror's sizes and offsets, in the probe's order, for the structs
// sweep_call reads and writes (the probe's first entries).
constexpr int64_t kMine[] = {
    S(Buf),
    O(Buf, data_),
    O(Buf, sizes_),
    O(Buf, strides_),
    O(Buf, numel_),
    O(Buf, ndim_),
    O(Buf, elem_size_),
    O(Buf, dtype_),
    O(Buf, defined_),
    O(Buf, is_cuda_),
    O(Buf, device_),
    S(BufList),
    O(BufList, p),
    O(BufList, n),
    S(IntSpan),
    O(IntSpan, p),
    O(IntSpan, n),
    S(FloatSpan),
    O(FloatSpan, p),
    O(FloatSpan, n),
    S(CStrList),
    O(CStrList, p),
    O(CStrList, n),
    S(GradList),
    O(GradList, items),
    O(GradList, n),
    S(ForwardOutputCore),
    O(ForwardOutputCore, wavefield),
    O(ForwardOutputCore, last_two),
    O(ForwardOutputCore, record),
    S(RTMOutputCore),
    O(RTMOutputCore, source_illumination),
    O(RTMOutputCore, receiver_illumination),
    O(RTMOutputCore, adcig),
    S(BackwardOutputCore),
    O(BackwardOutputCore, grads),
    O(BackwardOutputCore, source_illumination),
    O(BackwardOutputCore, receiver_illumination),
    O(BackwardOutputCore, adcig),
    S(ForwardInputCore),
    O(ForwardInputCore, models),
    O(ForwardInputCore, source),
    O(ForwardInputCore, lap_coes),
    O(ForwardInputCore, grad_coes),
    O(ForwardInputCore, M),
    O(ForwardInputCore, abcn),
    O(ForwardInputCore, sources_loc),
    O(ForwardInputCore, receivers_loc),
    O(ForwardInputCore, source_field_indices),
    O(ForwardInputCore, receiver_field_indices),
    O(ForwardInputCore, pml_vals),
    O(ForwardInputCore, wavefields),
    O(ForwardInputCore, forward_workspace),
    O(ForwardInputCore, derived_models),
    O(ForwardInputCore, last_two),
    O(ForwardInputCore, boundary_cpu),
    O(ForwardInputCore, boundary_gpu),
    O(ForwardInputCore, boundary_staging),
    O(ForwardInputCore, boundary_compress),
    O(ForwardInputCore, boundary_disk_files),
    O(ForwardInputCore, checkpoints),
    O(ForwardInputCore, checkpoint_steps),
    O(ForwardInputCore, save_all_wavefields),
    O(ForwardInputCore, use_boundary_saving),
    O(ForwardInputCore, use_checkpoint),
    O(ForwardInputCore, use_recursive_checkpoint),
    O(ForwardInputCore, checkpoint_on_cpu),
    O(ForwardInputCore, boundary_on_cpu),
    O(ForwardInputCore, boundary_on_disk),
    O(ForwardInputCore, boundary_disk_async_read),
    O(ForwardInputCore, use_pinned_memory),
    O(ForwardInputCore, free_surface),
    O(ForwardInputCore, fs_faces),
    O(ForwardInputCore, pad_lo),
    O(ForwardInputCore, pad_hi),
    O(ForwardInputCore, eq_aux),
    O(ForwardInputCore, topo_rows),
    O(ForwardInputCore, has_topo),
    O(ForwardInputCore, topo_category),
    O(ForwardInputCore, use_apm),
    O(ForwardInputCore, nt),
    O(ForwardInputCore, dt),
    O(ForwardInputCore, spacing),
    O(ForwardInputCore, transfer_interval),
    O(ForwardInputCore, boundary_ring_buffers),
    O(ForwardInputCore, boundary_tail_steps),
    O(ForwardInputCore, boundary_session),
    O(ForwardInputCore, checkpoint_interval),
    O(ForwardInputCore, checkpoint_count),
    O(ForwardInputCore, it_begin),
    O(ForwardInputCore, it_end),
    O(ForwardInputCore, record_out),
    O(ForwardInputCore, u_allt_out),
    O(ForwardInputCore, step_phase),
    O(ForwardInputCore, cut_face_mask),
    S(BackwardInputCore),
    O(BackwardInputCore, u_forward),
    O(BackwardInputCore, u_boundary),
    O(BackwardInputCore, u_last_two),
    O(BackwardInputCore, checkpoints),
    O(BackwardInputCore, checkpoint_replay),
    O(BackwardInputCore, checkpoint_steps),
    O(BackwardInputCore, adjoint_wavefields),
    O(BackwardInputCore, forward_wavefields),
    O(BackwardInputCore, adjoint_workspace),
    O(BackwardInputCore, derived_models),
    O(BackwardInputCore, boundary_cpu),
    O(BackwardInputCore, boundary_gpu),
    O(BackwardInputCore, boundary_staging),
    O(BackwardInputCore, boundary_compress),
    O(BackwardInputCore, boundary_disk_files),
    O(BackwardInputCore, models),
    O(BackwardInputCore, adjoint_source),
    O(BackwardInputCore, forward_source),
    O(BackwardInputCore, lap_coes),
    O(BackwardInputCore, grad_coes),
    O(BackwardInputCore, M),
    O(BackwardInputCore, abcn),
    O(BackwardInputCore, adjoint_sources_loc),
    O(BackwardInputCore, forward_sources_loc),
    O(BackwardInputCore, source_field_indices),
    O(BackwardInputCore, receiver_field_indices),
    O(BackwardInputCore, pml_vals),
    O(BackwardInputCore, eq_aux),
    O(BackwardInputCore, nt),
    O(BackwardInputCore, dt),
    O(BackwardInputCore, spacing),
    O(BackwardInputCore, free_surface),
    O(BackwardInputCore, fs_faces),
    O(BackwardInputCore, pad_lo),
    O(BackwardInputCore, pad_hi),
    O(BackwardInputCore, topo_rows),
    O(BackwardInputCore, has_topo),
    O(BackwardInputCore, topo_category),
    O(BackwardInputCore, use_apm),
    O(BackwardInputCore, checkpoint_on_cpu),
    O(BackwardInputCore, boundary_on_cpu),
    O(BackwardInputCore, boundary_on_disk),
    O(BackwardInputCore, boundary_disk_async_read),
    O(BackwardInputCore, use_pinned_memory),
    O(BackwardInputCore, transfer_interval),
    O(BackwardInputCore, boundary_ring_buffers),
    O(BackwardInputCore, boundary_tail_steps),
    O(BackwardInputCore, boundary_session),
    O(BackwardInputCore, checkpoint_interval),
    O(BackwardInputCore, checkpoint_count),
    O(BackwardInputCore, compute_illumination),
    O(BackwardInputCore, bw_it_begin),
    O(BackwardInputCore, bw_it_end),
    O(BackwardInputCore, step_phase),
    O(BackwardInputCore, grads_out),
    O(BackwardInputCore, illum_out),
    O(BackwardInputCore, adcig_out),
    O(BackwardInputCore, grad_split_iii_out),
    O(BackwardInputCore, cut_face_mask),
    O(BackwardInputCore, compute_adcig),
    O(BackwardInputCore, adcig_max_lag)
};
#undef S
#undef O
constexpr int64_t kMineCount = int64_t(sizeof(kMine) / sizeof(kMine[0]));

constexpr bool mirror_matches()
{
    for (int64_t i = 0; i < kMineCount; ++i)
        if (kMine[i] != kRefLayout[i]) return false;
    return true;
}
static_assert(kMineCount <= kRefLayoutCount, "mirror longer than the probe");
static_assert(mirror_matches(), "the ForwardInputCore/BackwardInputCore mirror disagrees with the ABI 5 layout table");
static_assert(kRefLayout[kRefLayoutCount - 4] == kFP32 && kRefLayout[kRefLayoutCount - 3] == kFP16 &&
                  kRefLayout[kRefLayoutCount - 2] == kBF16 && kRefLayout[kRefLayoutCount - 1] == kINT8,
              "BoundaryDtype enumerators");

void put_err(const std::string &m, char *err, int cap)
{
    if (!err || cap <= 0) return;
    std::strncpy(err, m.c_str(), size_t(cap) - 1);
    err[cap - 1] = '\0';
}

int64_t buf_bytes(const Buf &b) { return b.defined_ ? b.numel_ * int64_t(b.elem_size_) : 0; }

const Buf &first(const BufList &l, const char *what)
{
    SWP_CHECK(l.n >= 1 && l.p && l.p[0].defined_, "sweep_call: ", what, "[0] is not bound");
    return l.p[0];
}

// The equation of an entry name: 2 or 3 spatial dims, models, CPML profiles.
struct EqInfo {
    int ndim, nmodels, nprof;
    bool elastic;
};

EqInfo eq_of(const char *name)
{
    if (!std::strncmp(name, "acoustic2d", 10)) return {2, 1, 6, false};
    if (!std::strncmp(name, "acoustic3d", 10)) return {3, 1, 9, false};
    return {2, 3, 8, true};
}

// (M + 1) coefficients of a Buf into a host vector (device or host memory).
std::vector<float> host_coefs(const Buf &b, int M, const char *what)
{
    SWP_CHECK(b.defined_ && b.elem_size_ == 4 && b.numel_ >= M + 1, "sweep_call: ", what, " needs ", M + 1,
              " float32 values");
    std::vector<float> v(size_t(M) + 1);
    if (b.is_cuda_) SWP_HIP(hipMemcpy(v.data(), b.data_, v.size() * 4, hipMemcpyDeviceToHost));
    else std::memcpy(v.data(), b.data_, v.size() * 4);
    return v;
}

// The fields both directions share, onto p (which keeps pointers into lap/grad).
void common(const EqInfo &q, const BufList &models, const Buf &src, int M, int abcn, const Buf &sloc,
            const Buf &rloc, const IntSpan &sfi, const IntSpan &rfi, const BufList &pml, const IntSpan &plo,
            const IntSpan &phi, uint32_t nt, float dt, const FloatSpan &sp, std::vector<float> &lap,
            std::vector<float> &grad, const Buf &lapb, const Buf &gradb, swp_problem &p)
{
    std::memset(&p, 0, sizeof(p));
    SWP_CHECK(models.n == q.nmodels && models.p, "sweep_call: ", q.nmodels, " models expected, got ", models.n);
    const Buf &m0 = models.p[0];
    SWP_CHECK(m0.defined_ && m0.is_cuda_ && m0.elem_size_ == 4 && m0.ndim_ >= q.ndim, "sweep_call: models[0] must "
              "be a float32 device tensor of at least ", q.ndim, " dims");
    SWP_CHECK(src.defined_ && src.ndim_ == 3, "sweep_call: source must be (B, nsrc, nt)");
    p.ndim = q.ndim;
    p.nz = int32_t(m0.sizes_[m0.ndim_ - q.ndim]);
    p.ny = q.ndim == 3 ? int32_t(m0.sizes_[m0.ndim_ - 2]) : 1;
    p.nx = int32_t(m0.sizes_[m0.ndim_ - 1]);
    p.B = int32_t(src.sizes_[0]);
    p.M = M;
    p.nt = int32_t(nt);
    p.abcn = abcn;
    SWP_CHECK(int64_t(src.sizes_[2]) == int64_t(nt), "sweep_call: source has ", src.sizes_[2], " steps, nt is ", nt);
    const int64_t cells = int64_t(p.nz) * p.ny * p.nx;
    const int64_t per = m0.numel_ / cells;
    SWP_CHECK(per == 1 || per == p.B, "sweep_call: models[0] holds ", m0.numel_, " values, not 1 or B grids");
    p.model_batched = per == p.B && p.B > 1 ? 1 : 0;
    for (int k = 0; k < q.nmodels; ++k) {
        SWP_CHECK(models.p[k].numel_ == m0.numel_, "sweep_call: models differ in size");
        p.models[k] = static_cast<const float *>(models.p[k].data_);
    }
    // per-face PML pad: pad_lo / pad_hi in C axis order [z, (y,) x], a
    // missing or negative entry meaning abcn (csrc/cuda/common/context.h:125-146)
    for (int a = 0; a < 3 * 2; ++a) p.pad[a] = 0;
    const int faces[3][2] = {{0, 1}, {2, 3}, {4, 5}};   // z, y, x in swp_problem.pad
    for (int a = 0; a < q.ndim; ++a) {
        const int slot = q.ndim == 2 ? (a == 0 ? 0 : 2) : a;
        p.pad[faces[slot][0]] = (a < plo.n && plo.p[a] >= 0) ? plo.p[a] : abcn;
        p.pad[faces[slot][1]] = (a < phi.n && phi.p[a] >= 0) ? phi.p[a] : abcn;
    }
    p.dt = dt;
    // spacing in Cartesian order [dx, (dy,) dz] (propagator/_c.py:1129-1132)
    SWP_CHECK(sp.n == q.ndim && sp.p, "sweep_call: spacing needs ", q.ndim, " values");
    p.dx = sp.p[0];
    p.dy = q.ndim == 3 ? sp.p[1] : 0.f;
    p.dz = sp.p[q.ndim - 1];
    lap = host_coefs(lapb, M, "lap_coes");
    grad = host_coefs(gradb, M, "grad_coes");
    p.lap_coes = lap.data();
    p.grad_coes = grad.data();
    SWP_CHECK(pml.n == q.nprof && pml.p, "sweep_call: ", q.nprof, " CPML profiles expected, got ", pml.n);
    for (int k = 0; k < q.nprof; ++k) p.pml[k] = static_cast<const float *>(pml.p[k].data_);
    p.nsrc = int32_t(src.sizes_[1]);
    p.source = static_cast<const float *>(src.data_);
    SWP_CHECK(sloc.defined_ && sloc.elem_size_ == 4 && rloc.defined_ && rloc.elem_size_ == 4,
              "sweep_call: source and receiver locations must be int32 tensors");
    p.sources_loc = static_cast<const int32_t *>(sloc.data_);
    p.receivers_loc = static_cast<const int32_t *>(rloc.data_);
    p.nrec = rloc.ndim_ >= 2 ? int32_t(rloc.sizes_[1]) : 0;
    p.nsrc_fields = int32_t(sfi.n);
    p.nrec_fields = int32_t(rfi.n);
    SWP_CHECK(sfi.n >= 1 && sfi.n <= 5 && rfi.n >= 1 && rfi.n <= 5, "sweep_call: 1..5 source and receiver fields");
    for (int k = 0; k < sfi.n; ++k) p.src_fields[k] = sfi.p[k];
    for (int k = 0; k < rfi.n; ++k) p.rec_fields[k] = rfi.p[k];
}

// The forward->backward tape sits where the reference keeps the state of
// that memory mode, the same tensor on the forward and the backward input:
// u_allt_out / u_forward (full storage), boundary_gpu[0] (boundary saving),
// checkpoints[0] (chunk and binomial checkpointing).
const Buf &tape_of(int memory, const Buf &history, const BufList &boundary, const BufList &ckpts)
{
    if (memory == SWP_MEM_FULL) {
        SWP_CHECK(history.defined_, "sweep_call: full storage needs this core's tape bound as u_allt_out / u_forward");
        return history;
    }
    if (memory == SWP_MEM_BOUNDARY) return first(boundary, "boundary_gpu (this core's tape)");
    return first(ckpts, "checkpoints (this core's tape)");
}

int memory_of_forward(const ForwardInputCore &in)
{
    if (in.use_boundary_saving) return SWP_MEM_BOUNDARY;
    if (in.use_recursive_checkpoint) return SWP_MEM_BINOMIAL;
    if (in.use_checkpoint) return SWP_MEM_CKPT;
    if (in.save_all_wavefields) return SWP_MEM_FULL;
    return SWP_MEM_NONE;
}

int entry_memory(const char *name)
{
    const std::string n(name);
    if (n.size() >= 24 && n.compare(n.size() - 24, 24, "_backward_recursive_ckpt") == 0) return SWP_MEM_BINOMIAL;
    if (n.size() >= 14 && n.compare(n.size() - 14, 14, "_backward_ckpt") == 0) return SWP_MEM_CKPT;
    if (n.size() >= 12 && n.compare(n.size() - 12, 12, "_backward_bs") == 0) return SWP_MEM_BOUNDARY;
    return SWP_MEM_FULL;
}

void set_checkpointing(swp_problem &p, int32_t interval, int32_t count)
{
    if (p.memory == SWP_MEM_CKPT) p.ckpt_interval = interval > 0 ? interval : 1;
    if (p.memory == SWP_MEM_BINOMIAL) p.ckpt_interval = count > 0 ? count : 1;
}

// Translate a forward input; fills io with the record pointer.
void forward_problem(int entry, const ForwardInputCore &in, swp_problem &p, swp_forward_io &io,
                     std::vector<float> &lap, std::vector<float> &grad)
{
    const EqInfo q = eq_of(swp_entry_name(entry));
    common(q, in.models, in.source, in.M, in.abcn, in.sources_loc, in.receivers_loc, in.source_field_indices,
           in.receiver_field_indices, in.pml_vals, in.pad_lo, in.pad_hi, in.nt, in.dt, in.spacing, lap, grad,
           in.lap_coes, in.grad_coes, p);
    p.memory = memory_of_forward(in);
    set_checkpointing(p, in.checkpoint_interval, in.checkpoint_count);
    const Buf &ws = first(in.forward_workspace, "forward_workspace");
    p.workspace = ws.data_;
    p.workspace_bytes = buf_bytes(ws);
    if (p.memory != SWP_MEM_NONE) {
        const Buf &tp = tape_of(p.memory, in.u_allt_out, in.boundary_gpu, in.checkpoints);
        p.tape = tp.data_;
        p.tape_bytes = buf_bytes(tp);
    }
    io.record = in.record_out.defined_ ? static_cast<float *>(in.record_out.data_) : nullptr;
}

void backward_problem(int entry, const BackwardInputCore &in, swp_problem &p, swp_backward_io &io,
                      std::vector<float> &lap, std::vector<float> &grad)
{
    const char *name = swp_entry_name(entry);
    const EqInfo q = eq_of(name);
    common(q, in.models, in.forward_source, in.M, in.abcn, in.forward_sources_loc, in.adjoint_sources_loc,
           in.source_field_indices, in.receiver_field_indices, in.pml_vals, in.pad_lo, in.pad_hi, in.nt, in.dt,
           in.spacing, lap, grad, in.lap_coes, in.grad_coes, p);
    p.memory = entry_memory(name);
    set_checkpointing(p, in.checkpoint_interval, in.checkpoint_count);
    if (in.compute_illumination && in.illum_out.n >= 2) p.flags |= SWP_FLAG_ILLUMINATION;
    const Buf &ws = first(in.adjoint_workspace, "adjoint_workspace");
    p.workspace = ws.data_;
    p.workspace_bytes = buf_bytes(ws);
    const Buf &tp = tape_of(p.memory, in.u_forward, in.boundary_gpu, in.checkpoints);
    p.tape = tp.data_;
    p.tape_bytes = buf_bytes(tp);
    std::memset(&io, 0, sizeof(io));
    io.adjoint_source = static_cast<const float *>(in.adjoint_source.data_);
    // grads_out = [grad_wavelet?, one per model] (propagator/_c.py:238-239)
    const int64_t ng = in.grads_out.n;
    SWP_CHECK(ng == q.nmodels || ng == q.nmodels + 1, "sweep_call: grads_out holds ", ng, " buffers, expected ",
              q.nmodels, " (+1 with the wavelet gradient first)");
    int k0 = 0;
    if (ng == q.nmodels + 1) {
        io.grad_source = in.grads_out.p[0].defined_ ? static_cast<float *>(in.grads_out.p[0].data_) : nullptr;
        k0 = 1;
    }
    for (int k = 0; k < q.nmodels; ++k) {
        const Buf &g = in.grads_out.p[k0 + k];
        io.grad_models[k] = g.defined_ && g.numel_ > 0 ? static_cast<float *>(g.data_) : nullptr;
    }
    if (p.flags & SWP_FLAG_ILLUMINATION) {
        io.illum_source = static_cast<float *>(in.illum_out.p[0].data_);
        io.illum_receiver = static_cast<float *>(in.illum_out.p[1].data_);
    }
}

template <class F>
int guarded5(char *err, int cap, F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        put_err(e.what(), err, cap);
    } catch (...) {
        put_err("unknown error in libswp_core (sweep ABI 5)", err, cap);
    }
    return 1;
}

int entry_kind_checked(int entry)
{
    const int k = swp_entry_kind(entry);
    SWP_CHECK(k >= 0, "entry ", entry, " out of range [0, ", swp_entry_count(), ")");
    return k;
}

}  // namespace
}  // namespace abi5
}  // namespace swp

using namespace swp;
using namespace swp::abi5;

extern "C" {

SWP_CAPI int sweep_core_abi_version(void) { return 5; }
SWP_CAPI int64_t sweep_layout_count(void) { return kRefLayoutCount; }
SWP_CAPI int64_t sweep_layout_probe(int64_t i) { return (i >= 0 && i < kRefLayoutCount) ? kRefLayout[i] : -1; }
SWP_CAPI int sweep_entry_count(void) { return swp_entry_count(); }
SWP_CAPI const char *sweep_entry_name(int entry) { return swp_entry_name(entry); }
SWP_CAPI int sweep_entry_kind(int entry) { return swp_entry_kind(entry); }
SWP_CAPI int sweep_entry_available(int entry) { return swp_entry_available(entry); }
SWP_CAPI const char *sweep_core_device(void) { return swp_core_device(); }
SWP_CAPI void sweep_set_stream(void *stream) { swp_set_stream(stream); }
SWP_CAPI void *sweep_get_stream(void) { return swp_get_stream(); }

SWP_CAPI int sweep_call(int entry, const void *in, void *out, char *err, int err_cap)
{
    (void)out;   // the outputs are the buffers bound on the input
    swp_problem p;
    std::vector<float> lap, grad;
    int rc = guarded5(err, err_cap, [&] {
        SWP_CHECK(in != nullptr, "sweep_call: null input");
        if (entry_kind_checked(entry) == 0) {
            swp_forward_io io{};
            forward_problem(entry, *static_cast<const ForwardInputCore *>(in), p, io, lap, grad);
            char e2[2048] = {0};
            SWP_CHECK(swp_call(entry, &p, &io, e2, int(sizeof(e2))) == 0, e2);
        } else {
            swp_backward_io io{};
            backward_problem(entry, *static_cast<const BackwardInputCore *>(in), p, io, lap, grad);
            char e2[2048] = {0};
            SWP_CHECK(swp_call(entry, &p, &io, e2, int(sizeof(e2))) == 0, e2);
        }
    });
    return rc;
}

SWP_CAPI int sweep_query_bytes(int entry, const void *in, int64_t *workspace_bytes, int64_t *tape_bytes, char *err,
                               int err_cap)
{
    swp_problem p;
    std::vector<float> lap, grad;
    return guarded5(err, err_cap, [&] {
        SWP_CHECK(in != nullptr, "sweep_query_bytes: null input");
        // the buffers need not be bound yet: size with placeholders
        if (entry_kind_checked(entry) == 0) {
            ForwardInputCore f = *static_cast<const ForwardInputCore *>(in);
            Buf none{};
            none.defined_ = true;
            const BufList one{&none, 1};
            f.forward_workspace = one;
            f.boundary_gpu = one;
            f.checkpoints = one;
            f.u_allt_out = none;
            swp_forward_io io{};
            forward_problem(entry, f, p, io, lap, grad);
        } else {
            BackwardInputCore b = *static_cast<const BackwardInputCore *>(in);
            Buf none{};
            none.defined_ = true;
            const BufList one{&none, 1};
            b.adjoint_workspace = one;
            b.boundary_gpu = one;
            b.checkpoints = one;
            b.u_forward = none;
            swp_backward_io io{};
            backward_problem(entry, b, p, io, lap, grad);
        }
        char e2[2048] = {0};
        SWP_CHECK(swp_query(entry, &p, workspace_bytes, tape_bytes, e2, int(sizeof(e2))) == 0, e2);
    });
}

SWP_CAPI int sweep_sem_call(int entry, const void *in, int64_t arg, char *err, int err_cap)
{
    (void)entry; (void)in; (void)arg;
    put_err("sweep_sem_call: this core has no SEM entries", err, err_cap);
    return 1;
}

SWP_CAPI const char *sweep_entry_struct(int entry)
{
    (void)entry;
    return nullptr;
}

#define SWP_NO_RUNNER(side)                                                                                  \
    SWP_CAPI int sweep_##side##_runner_create(int entry, const void *in, void **handle, char *err, int cap)     \
    {                                                                                                        \
        (void)entry; (void)in;                                                                               \
        if (handle) *handle = nullptr;                                                                       \
        put_err("sweep_" #side "_runner_create: this core has no stepped runners (one call runs the whole "    \
                "time loop)", err, cap);                                                                     \
        return 1;                                                                                            \
    }                                                                                                        \
    SWP_CAPI int sweep_##side##_runner_run(void *h, int a, int b, int c, void *out, char *err, int cap)        \
    {                                                                                                        \
        (void)h; (void)a; (void)b; (void)c; (void)out;                                                       \
        put_err("sweep_" #side "_runner_run: no runner", err, cap);                                          \
        return 1;                                                                                            \
    }                                                                                                        \
    SWP_CAPI int sweep_##side##_runner_device_index(void *h)                                                 \
    {                                                                                                        \
        (void)h;                                                                                             \
        return -1;                                                                                           \
    }                                                                                                        \
    SWP_CAPI void sweep_##side##_runner_destroy(void *h) { (void)h; }
SWP_NO_RUNNER(forward)
SWP_NO_RUNNER(backward)
#undef SWP_NO_RUNNER

SWP_CAPI void *sweep_session_create(void) { return nullptr; }
SWP_CAPI int sweep_session_finish(void *s, char *err, int cap)
{
    (void)s; (void)err; (void)cap;
    return 0;
}
SWP_CAPI int sweep_session_used(void *s)
{
    (void)s;
    return 0;
}
SWP_CAPI void sweep_session_destroy(void *s) { (void)s; }

SWP_CAPI size_t sweep_visco_fft_workspace_bytes(int64_t B, int64_t nz, int64_t nx)
{
    (void)B; (void)nz; (void)nx;
    return 0;
}

SWP_CAPI int sweep_nvcomp_load(const char *path, char *err, int cap)
{
    (void)path;
    put_err("sweep_nvcomp_load: this core has no boundary compression", err, cap);
    return -1;
}
SWP_CAPI int sweep_nvcomp_loaded(void) { return 0; }
SWP_CAPI int64_t sweep_nvcomp_piece_bytes(void) { return 0; }
SWP_CAPI int64_t sweep_nvcomp_max_piece_out(int elem_size, char *err, int cap)
{
    (void)elem_size;
    put_err("sweep_nvcomp_max_piece_out: this core has no boundary compression", err, cap);
    return -1;
}
SWP_CAPI int64_t sweep_nvcomp_decompress_temp_bytes(int max_pieces, char *err, int cap)
{
    (void)max_pieces;
    put_err("sweep_nvcomp_decompress_temp_bytes: this core has no boundary compression", err, cap);
    return -1;
}

}  // extern "C"
