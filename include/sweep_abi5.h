/*
 * sweep_abi5.h -- libswp_core.so under the reference's own C ABI, version 5
 * (src/sweep/csrc/core/capi.h:21-268 of DeepWave-KAUST/sweep): the symbols
 * the reference's ctypes loader declares (backend/c/loader.py:52-79), so it
 * can load this library through SWEEP_CORE and run its ABI guard
 * (loader.py:91-118) against it.
 *
 * Inputs are the reference's ForwardInputCore / BackwardInputCore
 * (csrc/core/input_core.h:10-132), read through this library's own mirror of
 * their layout.  The entry table is this core's fifteen <C_NAME>_<op>
 * callables (acoustic2d, acoustic3d, elastic2d x forward, backward,
 * backward_bs, backward_ckpt, backward_recursive_ckpt), the names the
 * reference's equations resolve.  The core never allocates: where the
 * reference binds per-field wavefield and boundary lists, a caller binds one
 * workspace as forward_workspace[0] / adjoint_workspace[0] and one
 * forward->backward tape where the reference keeps that memory mode's state,
 * the same tensor on both calls: u_allt_out / u_forward (full storage),
 * boundary_gpu[0] (boundary saving), checkpoints[0] (checkpointing) -- of the
 * byte sizes sweep_query_bytes returns.  Functions of the reference's ABI this core has
 * no use for (SEM, stepped runners, boundary sessions, visco FFT, nvCOMP)
 * are exported and refuse or answer "none".
 */
#ifndef SWEEP_ABI5_H
#define SWEEP_ABI5_H

#include <stddef.h>
#include <stdint.h>

#include "sweep_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* capi.h:20-22: returns 5. */
SWP_CAPI int sweep_core_abi_version(void);
/* capi.h:24-29: the reference's layout probe, all 2333 entries (sizeof,
 * offsetof, BoundaryDtype values in its LAYOUT_PROBE order); -1 past the end. */
SWP_CAPI int64_t sweep_layout_count(void);
SWP_CAPI int64_t sweep_layout_probe(int64_t i);
/* capi.h:31-217: this core's entries (same ids as swp_entry_*). */
SWP_CAPI int sweep_entry_count(void);
SWP_CAPI const char *sweep_entry_name(int entry);
SWP_CAPI int sweep_entry_kind(int entry);
SWP_CAPI int sweep_entry_available(int entry);
/* capi.h:219-222: "cuda". */
SWP_CAPI const char *sweep_core_device(void);
/* capi.h:224-225: in = ForwardInputCore (kind 0) or BackwardInputCore (kind 1);
 * out (ForwardOutputCore / BackwardOutputCore) is not written: the outputs
 * are the buffers bound on the input (record_out; grads_out = [grad_wavelet
 * when the equation returns it, one per model]; illum_out).  Launches on the
 * thread's stream, asynchronously. */
SWP_CAPI int sweep_call(int entry, const void *in, void *out, char *err, int err_cap);
/* No reference counterpart: the workspace and tape bytes this core needs for
 * the call `in` describes (its workspace and tape Bufs may be unbound). */
SWP_CAPI int sweep_query_bytes(int entry, const void *in, int64_t *workspace_bytes, int64_t *tape_bytes,
                               char *err, int err_cap);
/* capi.h:227-230: no SEM entries here; refuses. */
SWP_CAPI int sweep_sem_call(int entry, const void *in, int64_t arg, char *err, int err_cap);
SWP_CAPI const char *sweep_entry_struct(int entry);
/* capi.h:232-243: no stepped runners here (one call runs the whole time loop). */
SWP_CAPI int sweep_forward_runner_create(int entry, const void *in, void **handle, char *err, int err_cap);
SWP_CAPI int sweep_forward_runner_run(void *handle, int it_begin, int it_end, int step_phase, void *out, char *err,
                                      int err_cap);
SWP_CAPI int sweep_forward_runner_device_index(void *handle);
SWP_CAPI void sweep_forward_runner_destroy(void *handle);
SWP_CAPI int sweep_backward_runner_create(int entry, const void *in, void **handle, char *err, int err_cap);
SWP_CAPI int sweep_backward_runner_run(void *handle, int bw_it_begin, int bw_it_end, int step_phase, void *out,
                                       char *err, int err_cap);
SWP_CAPI int sweep_backward_runner_device_index(void *handle);
SWP_CAPI void sweep_backward_runner_destroy(void *handle);
/* capi.h:245-250: no boundary-staging sessions; create returns NULL. */
SWP_CAPI void *sweep_session_create(void);
SWP_CAPI int sweep_session_finish(void *session, char *err, int err_cap);
SWP_CAPI int sweep_session_used(void *session);
SWP_CAPI void sweep_session_destroy(void *session);
/* capi.h:252-255: the thread-local launch stream (the same one swp_set_stream sets). */
SWP_CAPI void sweep_set_stream(void *stream);
SWP_CAPI void *sweep_get_stream(void);
/* capi.h:257-268: no visco FFT or nvCOMP here. */
SWP_CAPI size_t sweep_visco_fft_workspace_bytes(int64_t B, int64_t nz, int64_t nx);
SWP_CAPI int sweep_nvcomp_load(const char *path, char *err, int err_cap);
SWP_CAPI int sweep_nvcomp_loaded(void);
SWP_CAPI int64_t sweep_nvcomp_piece_bytes(void);
SWP_CAPI int64_t sweep_nvcomp_max_piece_out(int elem_size, char *err, int err_cap);
SWP_CAPI int64_t sweep_nvcomp_decompress_temp_bytes(int max_pieces, char *err, int err_cap);

#ifdef __cplusplus
}
#endif
#endif
