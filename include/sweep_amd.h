/*
 * sweep_amd.h -- the C boundary of libswp_core.so (MI355X / gfx950).
 *
 * The Python host (sweep_amd.propagator) calls these through ctypes; any
 * other host can bind them the same way.  Plain C: no torch types, plain
 * pointers and sizes.  Errors come back as a non-zero return with the message
 * copied into ``err``; nothing is thrown across.  The core never allocates:
 * the caller owns every buffer, including the per-call workspace and the
 * forward->backward tape, whose byte sizes it asks for with swp_query.
 *
 * Each entry point below replaces a function of the reference's C ABI
 * (src/sweep/csrc/core/capi.h of DeepWave-KAUST/sweep); the citation sits on
 * the declaration.  The input structs are this project's own: one
 * swp_problem (grid, models, coefficients, geometry, memory strategy, caller
 * buffers) in place of the reference's ForwardInputCore / BackwardInputCore
 * (csrc/core/input_core.h:10-132), whose wavefield lists the core here carves
 * out of the caller's workspace instead.
 */
#ifndef SWEEP_AMD_H
#define SWEEP_AMD_H

#include <stddef.h>
#include <stdint.h>

#define SWP_CAPI __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define SWP_ABI_VERSION 6

/* Memory strategy of a gradient run (reference: propagator/options.py
 * Full / BoundarySaving(storage='gpu') / Ckpt(mode='chunk'|'recursive')). */
enum {
    SWP_MEM_NONE = 0,      /* forward only: nothing kept for a backward */
    SWP_MEM_FULL = 1,      /* per-step imaging field (nt, B, grid) */
    SWP_MEM_BOUNDARY = 2,  /* boundary strips + the last two states */
    SWP_MEM_CKPT = 3,      /* state snapshots every ckpt_interval steps, chunk replay */
    SWP_MEM_BINOMIAL = 4,  /* ckpt_interval state snapshot slots, binomial reversal
                              (served by the *_backward_recursive_ckpt entries) */
};

/* swp_problem.flags */
enum {
    SWP_FLAG_ILLUMINATION = 1,   /* accumulate illumination in the backward */
    SWP_FLAG_BF16_STRIPS = 2,    /* SWP_MEM_BOUNDARY strips in bfloat16, round to nearest even
                                    (reference: BoundarySaving(storage_dtype='bf16'),
                                    propagator/options.py:204-212, boundarysaver.cu:29) */
    SWP_FLAG_GRAPH = 4,          /* the caller keeps this call's buffers at fixed addresses call after
                                    call: capture the call's launches into a graph and replay it
                                    (swp_set_graphs) */
};

/* One propagation problem on the runtime grid.  Arrays are row-major with x
 * fastest; a per-shot array is (B, nz[, ny], nx).  Coordinates are runtime
 * cell indices in (x[, y], z) order, like the reference's sources_loc /
 * receivers_loc.  Profiles follow the reference's pml_vals order:
 *   cpmlr (acoustic): az, bz, dbz, [ay, by, dby,] ax, bx, dbx
 *   cpmls (elastic):  az, bz, azh, bzh, ax, bx, axh, bxh
 * Every pointer is device memory except lap_coes / grad_coes, which the core
 * reads on the host (the kernels take the weights by value).                */
typedef struct swp_problem {
    int32_t ndim;            /* 2 or 3 */
    int32_t nz, ny, nx;      /* runtime grid (ny = 1 in 2-D) */
    int32_t B;               /* shots advanced together */
    int32_t M;               /* half stencil order (spatial_order / 2) */
    int32_t nt;              /* time steps */
    int32_t abcn;            /* absorbing width (reconstruction seed rim = abcn + M) */
    int32_t pad[6];          /* PML pad per face: z_lo, z_hi, y_lo, y_hi, x_lo, x_hi */
    float dt;
    float dz, dy, dx;
    const float *lap_coes;   /* HOST (M+1): [c0, c1..cM], zero-sum float32 second derivative */
    const float *grad_coes;  /* HOST (M+1): [0, c1..cM] centred first derivative */
    const float *pml[12];
    const float *models[3];  /* acoustic: vp; elastic: vp, vs, rho */
    int32_t model_batched;   /* 1: every model is (B, grid); 0: one shared grid */
    int32_t nsrc, nrec;
    const float *source;            /* (B, nsrc, nt) forward source wavelets */
    const int32_t *sources_loc;     /* (B, nsrc, ndim) */
    const int32_t *receivers_loc;   /* (B, nrec, ndim) */
    int32_t memory;          /* SWP_MEM_* */
    int32_t ckpt_interval;   /* SWP_MEM_CKPT: steps between snapshots */
    int32_t flags;           /* SWP_FLAG_ILLUMINATION, SWP_FLAG_BF16_STRIPS */
    int32_t reserved;
    /* Wavefield indices the sources inject into and the receivers record
     * (reference: source_field_indices / receiver_field_indices,
     * csrc/core/input_core.h:19-20).  Elastic: 0 vx, 1 vz, 2 sxx, 3 szz,
     * 4 sxz; the record and the adjoint source are (nrec_fields, B, nrec, nt).
     * Acoustic: one field (0). */
    int32_t nsrc_fields, nrec_fields;
    int32_t src_fields[5], rec_fields[5];
    void *tape;              /* forward -> backward state, swp_query's tape_bytes */
    int64_t tape_bytes;
    void *workspace;         /* per-call scratch, swp_query's workspace_bytes */
    int64_t workspace_bytes;
} swp_problem;

/* Outputs of a forward entry (kind 0). */
typedef struct swp_forward_io {
    float *record;           /* ([nrec_fields,] B, nrec, nt), time innermost (reference raw layout) */
} swp_forward_io;

/* Inputs/outputs of a backward entry (kind 1).  Gradients accumulate. */
typedef struct swp_backward_io {
    const float *adjoint_source;  /* ([nrec_fields,] B, nrec, nt) residual, raw record layout */
    float *grad_source;           /* (B, nsrc, nt) or NULL */
    float *grad_models[3];        /* (B, grid) per model, or NULL */
    float *illum_source;          /* (B, grid) sum_t u_tt^2, or NULL */
    float *illum_receiver;        /* (B, grid) sum_t lambda^2, or NULL */
} swp_backward_io;

/* Replaces sweep_core_abi_version (capi.h:22-23). */
SWP_CAPI int swp_core_abi_version(void);

/* Replaces sweep_layout_count / sweep_layout_probe (capi.h:25-30): sizeof and
 * offsetof of every struct field above, in declaration order, so the ctypes
 * mirror can refuse a drifted build.  -1 past the end. */
SWP_CAPI int64_t swp_layout_count(void);
SWP_CAPI int64_t swp_layout_probe(int64_t i);

/* Replace sweep_entry_count / _name / _kind / _available (capi.h:32-64).
 * Names are <C_NAME>_<op>: forward, backward, backward_bs, backward_ckpt,
 * backward_recursive_ckpt -- the callables equation._C() resolves. */
SWP_CAPI int swp_entry_count(void);
SWP_CAPI const char *swp_entry_name(int entry);
SWP_CAPI int swp_entry_kind(int entry);
SWP_CAPI int swp_entry_available(int entry);

/* Replaces sweep_core_device (capi.h:66-69): "cuda", the torch device type of
 * a ROCm GPU. */
SWP_CAPI const char *swp_core_device(void);

/* Workspace and tape bytes the entry needs for this problem (no reference
 * counterpart: the reference's Python sizes every buffer from cuda_layout). */
SWP_CAPI int swp_query(int entry, const swp_problem *in, int64_t *workspace_bytes,
                       int64_t *tape_bytes, char *err, int err_cap);

/* Replaces sweep_call (capi.h:71-72): io is swp_forward_io (kind 0) or
 * swp_backward_io (kind 1).  Launches on the thread's stream, asynchronously;
 * the caller keeps every buffer alive until that stream reaches the end. */
SWP_CAPI int swp_call(int entry, const swp_problem *in, void *io, char *err, int err_cap);

/* Replace sweep_set_stream / sweep_get_stream (capi.h:250-253): the
 * thread-local hipStream_t every launch of this thread goes on. */
SWP_CAPI void swp_set_stream(void *stream);
SWP_CAPI void *swp_get_stream(void);

/* Time-loop graphs (no reference counterpart: the reference launches every
 * step from its host loop, csrc/cuda/common/eq_driver.cuh:377-434, 842-861).
 * For a call whose problem carries SWP_FLAG_GRAPH, swp_call captures the
 * call's whole launch sequence -- every kernel of every time step -- into a
 * HIP graph on a side stream, and launches the graph on the thread's stream;
 * a later call with the same entry, device, problem (buffer addresses,
 * scalars, host coefficient values), io pointers, tuning and kernel-family
 * settings replays it instead of launching step by step.  Results are
 * bit-identical either way.  A shape whose buffers move between calls falls
 * back to eager launches after three captures without a replay.  Calls
 * without the flag, with kernel timing or the launch census on, or running
 * more than 4096 time steps (re-runs of a checkpointed backward included)
 * run eagerly.
 * swp_set_graphs: 1 on, 0 off, -1 the SWP_GRAPHS environment variable
 * (default on, "0" turns it off).  swp_graph_stats reports captures, replays,
 * eager fallbacks and cached graphs since the last swp_graph_clear, which
 * also frees every cached graph. */
SWP_CAPI void swp_set_graphs(int mode);
SWP_CAPI int swp_graph_stats(int64_t *captures, int64_t *replays, int64_t *eager, int64_t *cached);
SWP_CAPI void swp_graph_clear(void);

/* Kernel-class timing with HIP events on the launch stream.  stride > 0 turns
 * it on (and clears earlier samples), timing one launch in every `stride` of
 * each class; 0 turns it off.  swp_last_kernel_stats returns the summed
 * duration (ms) and count of the sampled launches of class `which` (0 forward
 * step, 1 adjoint step, 2 reconstruction, 3 the rest) since then.  No
 * reference counterpart: the roofline hook of bench.py. */
SWP_CAPI void swp_set_timing(int stride);
SWP_CAPI int swp_last_kernel_stats(int which, double *ms_total, int64_t *launches);

/* Launch census (no reference counterpart: a test hook).  on = 1 clears it
 * and turns it on, 0 turns it off.  While on, every launch of a 2-D z-march
 * kernel (sites a2d.fwd_int, a2d.fwd_band, a2d.adj_int, a2d.adj_band,
 * a2d.recon_zm), of a 3-D region kernel (a3d.fwd3.XYZ, a3d.adj3.XYZ with X,
 * Y, Z = 1 where the region lies in that axis' CPML margin, a3d.recon_zm)
 * and of the per-cell stencil kernels (a2d.fwd_cell, a2d.adj_cell,
 * a3d.fwd_cell, a3d.adj_cell) is counted with its workgroups.  swp_census writes "site launches workgroups" lines into
 * buf (NUL-terminated, truncated to cap) and returns the full text length,
 * or -1. */
SWP_CAPI void swp_census_enable(int on);
SWP_CAPI int64_t swp_census(char *buf, int64_t cap);

/* The binomial checkpoint schedule SWP_MEM_BINOMIAL runs, for inspection
 * and host-side tests: writes min(count, cap) (code, arg) int32 pairs --
 * code 0 step arg, 1 save to slot arg, 2 load slot arg, 3 re-run and reverse
 * step arg -- and returns count, or -1 on bad arguments.  swp_revolve_spine
 * returns the (slot, step) snapshots the forward pass takes.  Replaces the
 * recursive-bisection schedule of the reference's recursive checkpoint
 * backward (csrc/cuda/common/eq_driver.cuh:1055-1259, steps from
 * propagator/_c.py:2608-2615); no C function of the reference corresponds. */
SWP_CAPI int64_t swp_revolve_schedule(int32_t nt, int32_t slots, int32_t *ops, int64_t cap);
SWP_CAPI int64_t swp_revolve_spine(int32_t nt, int32_t slots, int32_t *pts, int64_t cap);

/* Launch-shape tuning (no reference counterpart: the reference fixes its
 * block shapes at compile time).  key: SWP_TUNE_*; value > 0 sets it, 0
 * restores the default (the SWP_TZ* environment variable if set, else the
 * built-in).  Applies to calls made after it, process-wide.  3-D records and
 * gradients are bit-identical under any setting.  2-D results move at the
 * float32 rounding level: gradients with the reconstruction segment height
 * (measured 4e-8 relative), records where the interior segment height moves
 * cells between the interior and band kernels (1.2e-6 at 96 x 128).  The
 * 2-D heights act on launches of at least 4M shot-cells, which run the
 * z-marching kernels (smaller ones run one thread per cell; the
 * SWP_KERNELS environment variable, percell or zmarch, forces a family).
 * swp_set_tuning returns 0, or 1 for an unknown key or a
 * negative value; swp_get_tuning returns the value set (0: default), -1 for
 * an unknown key. */
enum {
    SWP_TUNE_TZ2 = 0,        /* 2-D interior z-march segment rows (default 32) */
    SWP_TUNE_TZ2_BAND = 1,   /* 2-D CPML-band segment rows (default 8) */
    SWP_TUNE_TZ2_RECON = 2,  /* 2-D reconstruction segment rows (default 32) */
    SWP_TUNE_TZ3 = 3,        /* 3-D z-march segment planes (default 64) */
    SWP_TUNE_COUNT = 4
};
SWP_CAPI int swp_set_tuning(int32_t key, int32_t value);
SWP_CAPI int32_t swp_get_tuning(int32_t key);

/* Record wire format.  A forward entry writes its record raw, time innermost:
 * (nfield, B, nrec, nt); callers and losses use the canonical
 * (B, nt, nrec, nfield).  swp_record_to_canonical replaces the reference's
 * _cuda_record_to_canonical (propagator/_c.py:109-122), swp_record_to_raw its
 * _canonical_to_cuda_record (_c.py:158-186) for the adjoint source.  Device
 * pointers, not overlapping; 1 <= nfield <= 5; launched on the thread's
 * stream.  Returns 0, or 1 with a message in err. */
SWP_CAPI int swp_record_to_canonical(const float *raw, float *canonical, int32_t nfield, int32_t B,
                                     int32_t nrec, int32_t nt, char *err, int err_cap);
SWP_CAPI int swp_record_to_raw(const float *canonical, float *raw, int32_t nfield, int32_t B,
                               int32_t nrec, int32_t nt, char *err, int err_cap);

/* L2 misfit over canonical records: residual = (pred - obs) * mask[b, r]
 * (mask (B, nrec) or NULL for none, residual NULL to skip it) and
 * loss[0] = 0.5 * sum residual^2, accumulated in float64 in a fixed order
 * (the same value on every run).  partials: device scratch of
 * SWP_MISFIT_PARTIALS doubles.  residual is the adjoint source of the loss.
 * The reference leaves losses to a separate package (sweep_loss, not part of
 * it); this is the loss of its tests and FD checks, 0.5 * ||pred - obs||^2
 * with padded receivers masked out. */
#define SWP_MISFIT_PARTIALS 1024
SWP_CAPI int swp_l2_misfit(const float *pred, const float *obs, const float *mask, int32_t nfield, int32_t B,
                           int32_t nrec, int32_t nt, float *residual, double *loss, double *partials,
                           char *err, int err_cap);

/* swp_l2_misfit without the partials argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * partial sums go to a static device array of SWP_MISFIT_PARTIALS doubles
 * (8 KiB) that the library holds on each GPU it runs on until the process
 * ends; every block writes its own slot before the final sum reads it, so
 * nothing needs clearing between calls.  Calls that overlap on one GPU (two
 * streams, two threads) share that array and race: serialise them.  Same
 * result as swp_l2_misfit bit for bit; allocates nothing. */
SWP_CAPI int swp_l2_misfit_noalloc(const float *pred, const float *obs, const float *mask, int32_t nfield, int32_t B,
                                   int32_t nrec, int32_t nt, float *residual, double *loss, char *err, int err_cap);

/* ---- Learned preconditioner (2-D U-Net) kernels: no counterpart in the
 * reference, whose path has no learned component.  Tensors are NHWC views:
 * a device pointer to channel 0 of pixel 0 and a channel stride; bf16 tensors
 * are uint16 bit patterns (round to nearest even).  Convolutions run on the
 * bf16 matrix cores with float32 accumulation, same-size, ks 1 or 3; channel
 * counts, strides and pointers of bf16 views are multiples of 8 elements /
 * 16 bytes.  Weights are [cout][ks*ks][cin].  All launch on the thread's
 * stream and return 0, or 1 with a message. */
SWP_CAPI int swp_nn_conv_fwd(const uint16_t *in, int32_t N, int32_t H, int32_t W, int32_t cin, int32_t in_stride,
                             const uint16_t *w, int32_t cout, int32_t ks, const float *bias, int32_t relu,
                             void *out, int32_t out_stride, int32_t out_f32, char *err, int err_cap);
/* Float count of the scratch swp_nn_conv_wgrad needs for these extents. */
SWP_CAPI int64_t swp_nn_conv_wgrad_scratch(int32_t N, int32_t H, int32_t W, int32_t cin, int32_t cout, int32_t ks);
/* dw[co][tap][ci] = sum_p dout[p][co] in[p + off(tap)][ci] (float32,
 * overwritten, fixed summation order); db[co] = sum_p dout[p][co] unless NULL. */
SWP_CAPI int swp_nn_conv_wgrad(const uint16_t *in, int32_t N, int32_t H, int32_t W, int32_t cin, int32_t in_stride,
                               const uint16_t *dout, int32_t cout, int32_t dout_stride, int32_t ks, float *dw,
                               float *db, float *scratch, char *err, int err_cap);
/* wt[ci][tap][co] = w[co][ks*ks-1-tap][ci]: the data gradient of a
 * convolution is swp_nn_conv_fwd(dout, wt) with cin and cout swapped. */
SWP_CAPI int swp_nn_conv_flip(const uint16_t *w, int32_t cout, int32_t cin, int32_t ks, uint16_t *wt,
                              char *err, int err_cap);
SWP_CAPI int swp_nn_relu_bwd(uint16_t *d, int32_t d_stride, const uint16_t *act, int32_t act_stride, int64_t P,
                             int32_t C, char *err, int err_cap);
/* 2x2 average pool of an (N, H, W, C) view; the adjoint adds into din. */
SWP_CAPI int swp_nn_pool2_fwd(const uint16_t *in, int32_t in_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                              uint16_t *out, int32_t out_stride, char *err, int err_cap);
SWP_CAPI int swp_nn_pool2_bwd_acc(const uint16_t *dout, int32_t dout_stride, int32_t N, int32_t H, int32_t W,
                                  int32_t C, uint16_t *din, int32_t din_stride, char *err, int err_cap);
/* Nearest 2x upsample of an (N, H, W, C) view and its adjoint (overwrites din). */
SWP_CAPI int swp_nn_up2_fwd(const uint16_t *in, int32_t in_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                            uint16_t *out, int32_t out_stride, char *err, int err_cap);
SWP_CAPI int swp_nn_up2_bwd(const uint16_t *dout, int32_t dout_stride, int32_t N, int32_t H, int32_t W, int32_t C,
                            uint16_t *din, int32_t din_stride, char *err, int err_cap);
/* x[p][c] = bf16(scale[c] * planes[c][p]) for c < nsrc (<= 4), 0 up to stride. */
SWP_CAPI int swp_nn_pack_input(const float *const *planes, const float *scale, int32_t nsrc, int64_t P, uint16_t *x,
                               int32_t stride, char *err, int err_cap);
/* loss[0] = sum_p (out[p][0] - target[p])^2 mask[p] / count (float64, fixed
 * order; mask NULL: 1); dout[p][0] = bf16(2 (out - target) mask / count), the
 * other channels of dout zero.  partials: SWP_MISFIT_PARTIALS doubles. */
SWP_CAPI int swp_nn_mse(const float *out, int32_t out_stride, const float *target, const float *mask, int64_t P,
                        double count, uint16_t *dout, int32_t dout_stride, double *loss, double *partials,
                        char *err, int err_cap);
/* One Adam step (float32 master weights; wbf, if not NULL, gets their bf16). */
SWP_CAPI int swp_nn_adam(float *w, const float *g, float *m, float *v, uint16_t *wbf, int64_t n, float lr, float b1,
                         float b2, float eps, int32_t step, char *err, int err_cap);
SWP_CAPI int swp_nn_to_bf16(const float *src, uint16_t *dst, int64_t n, char *err, int err_cap);

#ifdef __cplusplus
}
#endif
#endif
