// The reference's C boundary, version 5, as libswp_core reads it: a mirror of
// the standard-layout structs its loader fills and hands to sweep_call
// (csrc/core/buf.h:66-83, buflist.h, outputs.h, input_core.h:10-132 of
// DeepWave-KAUST/sweep), field for field in declaration order.  Only the
// layout has to agree -- the loader compares sweep_layout_probe with its
// own ctypes mirror before any call -- and the static_asserts in
// sweep_abi5.hip check every offset here against that table.
#pragma once

#include <cstdint>

namespace swp {
namespace abi5 {

constexpr int kBufMaxDims = 8;   // BUF_MAX_DIMS
constexpr int kMaxGrads = 32;    // SWEEP_MAX_GRADS

// Storage-dtype tag of a Buf (BoundaryDtype).
enum : int32_t { kFP32 = 0, kFP16 = 1, kBF16 = 2, kINT8 = 3 };

// A non-owning strided view: element strides, sizes, byte width, where it lives.
struct Buf {
    void *data_;
    int64_t sizes_[kBufMaxDims];
    int64_t strides_[kBufMaxDims];
    int64_t numel_;
    int32_t ndim_;
    int32_t elem_size_;
    int32_t dtype_;
    bool defined_;
    bool is_cuda_;
    int32_t device_;
};

struct BufList {
    const Buf *p;
    int64_t n;
};
struct IntSpan {
    const int32_t *p;
    int64_t n;
};
struct FloatSpan {
    const float *p;
    int64_t n;
};
struct CStrList {
    const char *const *p;
    int64_t n;
};
struct GradList {
    Buf items[kMaxGrads];
    int64_t n;
};

struct ForwardOutputCore {
    Buf wavefield, last_two, record;
};
struct RTMOutputCore {
    Buf source_illumination, receiver_illumination, adcig;
};
struct BackwardOutputCore {
    GradList grads;
    Buf source_illumination, receiver_illumination, adcig;
};

struct ForwardInputCore {
    BufList models;
    Buf source;
    Buf lap_coes;
    Buf grad_coes;
    int32_t M;
    int32_t abcn;
    Buf sources_loc;
    Buf receivers_loc;
    IntSpan source_field_indices;
    IntSpan receiver_field_indices;
    BufList pml_vals;
    BufList wavefields;
    BufList forward_workspace;
    BufList derived_models;
    Buf last_two;
    BufList boundary_cpu;
    BufList boundary_gpu;
    BufList boundary_staging;
    BufList boundary_compress;
    CStrList boundary_disk_files;
    BufList checkpoints;
    Buf checkpoint_steps;
    bool save_all_wavefields;
    bool use_boundary_saving;
    bool use_checkpoint;
    bool use_recursive_checkpoint;
    bool checkpoint_on_cpu;
    bool boundary_on_cpu;
    bool boundary_on_disk;
    bool boundary_disk_async_read;
    bool use_pinned_memory;
    bool free_surface;
    int32_t fs_faces;
    IntSpan pad_lo;
    IntSpan pad_hi;
    BufList eq_aux;
    Buf topo_rows;
    bool has_topo;
    Buf topo_category;
    bool use_apm;
    uint32_t nt;
    float dt;
    FloatSpan spacing;
    int32_t transfer_interval;
    int32_t boundary_ring_buffers;
    int32_t boundary_tail_steps;
    void *boundary_session;
    int32_t checkpoint_interval;
    int32_t checkpoint_count;
    int32_t it_begin;
    int32_t it_end;
    Buf record_out;
    Buf u_allt_out;
    int32_t step_phase;
    int32_t cut_face_mask;
};

struct BackwardInputCore {
    Buf u_forward;
    BufList u_boundary;
    Buf u_last_two;
    BufList checkpoints;
    BufList checkpoint_replay;
    Buf checkpoint_steps;
    BufList adjoint_wavefields;
    BufList forward_wavefields;
    BufList adjoint_workspace;
    BufList derived_models;
    BufList boundary_cpu;
    BufList boundary_gpu;
    BufList boundary_staging;
    BufList boundary_compress;
    CStrList boundary_disk_files;
    BufList models;
    Buf adjoint_source;
    Buf forward_source;
    Buf lap_coes;
    Buf grad_coes;
    int32_t M;
    int32_t abcn;
    Buf adjoint_sources_loc;
    Buf forward_sources_loc;
    IntSpan source_field_indices;
    IntSpan receiver_field_indices;
    BufList pml_vals;
    BufList eq_aux;
    uint32_t nt;
    float dt;
    FloatSpan spacing;
    bool free_surface;
    int32_t fs_faces;
    IntSpan pad_lo;
    IntSpan pad_hi;
    Buf topo_rows;
    bool has_topo;
    Buf topo_category;
    bool use_apm;
    bool checkpoint_on_cpu;
    bool boundary_on_cpu;
    bool boundary_on_disk;
    bool boundary_disk_async_read;
    bool use_pinned_memory;
    int32_t transfer_interval;
    int32_t boundary_ring_buffers;
    int32_t boundary_tail_steps;
    void *boundary_session;
    int32_t checkpoint_interval;
    int32_t checkpoint_count;
    bool compute_illumination;
    int32_t bw_it_begin;
    int32_t bw_it_end;
    int32_t step_phase;
    BufList grads_out;
    BufList illum_out;
    Buf adcig_out;
    Buf grad_split_iii_out;
    int32_t cut_face_mask;
    bool compute_adcig;
    int32_t adcig_max_lag;
};

}  // namespace abi5
}  // namespace swp
