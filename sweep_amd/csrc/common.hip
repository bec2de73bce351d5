// Shared device helpers of libswp_core: zero fill, the point kernels that
// inject sources and gather receivers, and the kernel-class timer.
#include "common.h"

#include <algorithm>

namespace swp {

// ----------------------------------------------------------------- fill
__global__ void __launch_bounds__(256) fill_zero_kernel(float *__restrict__ p, int64_t n)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x * 4;
    for (int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4; i < n; i += stride) {
        if (i + 3 < n) {
            *reinterpret_cast<float4 *>(p + i) = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            for (int64_t k = i; k < n; ++k) p[k] = 0.f;
        }
    }
}

void fill_zero(float *p, int64_t n)
{
    if (n <= 0) return;
    SWP_CHECK((reinterpret_cast<uintptr_t>(p) & 15) == 0, "fill_zero: unaligned buffer");
    int64_t blocks = (n / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), p, n);
    SWP_LAUNCHED();
}

// ----------------------------------------------------------------- points
// One block per shot.  Phase 1 adds w[b, s, it] into u at each (in-grid)
// point with an atomic (several points may share a cell; the reference's
// add_source, common.cu:11-37); phase 2, after the block barrier, reads u at
// each gather point into out[b, r, it] (record_kernel, common.cu:119-144) or
// adds it there (accumulate_source_grad_2d, acoustic2d/kernels.cu:223-255).
// The phase-2 loads go around the vector L1 (agent-scope atomic loads) so
// they see the phase-1 atomics, which live in L2.
__global__ void __launch_bounds__(256) points_kernel(PointArgs a)
{
    const int b = blockIdx.x;
    float *u_add = a.u_add ? a.u_add + int64_t(b) * a.n : nullptr;
    if (u_add) {
        for (int s = threadIdx.x; s < a.nadd; s += blockDim.x) {
            const int32_t *l = a.add_loc + (int64_t(b) * a.nadd + s) * a.ndim;
            int ix = l[0], iy = a.ndim == 3 ? l[1] : 0, iz = l[a.ndim - 1];
            if (ix < 0 || ix >= a.nx || iy < 0 || iy >= a.ny || iz < 0 || iz >= a.nz) continue;
            const int64_t cell = (int64_t(iz) * a.ny + iy) * (a.px > 0 ? a.px : a.nx) + ix;
            atomicAdd(u_add + cell, a.sign * a.add_w[(int64_t(b) * a.nadd + s) * a.nt + a.it]);
        }
    }
    if (!a.u_get) return;
    __syncthreads();
    const float *u_get = a.u_get + int64_t(b) * a.n;
    for (int r = threadIdx.x; r < a.nget; r += blockDim.x) {
        const int32_t *l = a.get_loc + (int64_t(b) * a.nget + r) * a.ndim;
        int ix = l[0], iy = a.ndim == 3 ? l[1] : 0, iz = l[a.ndim - 1];
        if (ix < 0 || ix >= a.nx || iy < 0 || iy >= a.ny || iz < 0 || iz >= a.nz) continue;
        const int64_t cell = (int64_t(iz) * a.ny + iy) * (a.px > 0 ? a.px : a.nx) + ix;
        const float v = __hip_atomic_load(u_get + cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        float *o = a.get_out + (int64_t(b) * a.nget + r) * a.nt + a.it;
        if (a.get_accumulate) *o += v;
        else *o = v;
    }
}

void launch_points(const PointArgs &a, int B)
{
    if (B <= 0) return;
    if (!a.u_add && !a.u_get) return;
    hipLaunchKernelGGL(points_kernel, dim3(B), dim3(256), 0, stream(), a);
    SWP_LAUNCHED();
}

// ----------------------------------------------------------------- pitch
template <bool ADD>
__global__ void __launch_bounds__(256) pitch_kernel(float *__restrict__ dst, int64_t dp,
                                                    const float *__restrict__ src, int64_t sp, int64_t rows,
                                                    int width)
{
    const int64_t cols = ADD ? width : dp;
    const int64_t total = rows * cols;
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < total; k += int64_t(gridDim.x) * 256) {
        const int64_t r = k / cols, c = k % cols;
        if (ADD) dst[r * dp + c] += src[r * sp + c];
        else dst[r * dp + c] = c < width ? src[r * sp + c] : 0.f;
    }
}

template <bool ADD>
static void launch_pitch(float *dst, int64_t dp, const float *src, int64_t sp, int64_t rows, int width)
{
    const int64_t total = rows * (ADD ? width : dp);
    if (total <= 0) return;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(pitch_kernel<ADD>, dim3(unsigned(blocks)), dim3(256), 0, stream(), dst, dp, src, sp, rows,
                       width);
    SWP_LAUNCHED();
}

void pitch_copy(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width)
{
    launch_pitch<false>(dst, dpitch, src, spitch, rows, width);
}

void pitch_add(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width)
{
    launch_pitch<true>(dst, dpitch, src, spitch, rows, width);
}

// ----------------------------------------------------------------- bf16 pack
// Pairs (k, k + 1) from k = head go out as one 4-byte store (dst + head is
// 4-byte aligned); a leading or trailing odd element goes alone.
__global__ void __launch_bounds__(256) pack_bf16_kernel(uint16_t *__restrict__ dst, const float *__restrict__ src,
                                                        int64_t n, int head)
{
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t == 0 && head) dst[0] = bf16_rne(src[0]);
    for (int64_t k = head + t * 2; k < n; k += int64_t(gridDim.x) * 512) {
        if (k + 1 < n)
            *reinterpret_cast<uint32_t *>(dst + k) = uint32_t(bf16_rne(src[k])) | (uint32_t(bf16_rne(src[k + 1])) << 16);
        else
            dst[k] = bf16_rne(src[k]);
    }
}

void pack_bf16(float *dst, const float *src, int64_t n)
{
    if (n <= 0) return;
    const uintptr_t d = reinterpret_cast<uintptr_t>(dst);
    SWP_CHECK((d & 1) == 0 && (reinterpret_cast<uintptr_t>(src) & 3) == 0, "pack_bf16: unaligned buffers");
    const int head = (d & 3) ? 1 : 0;
    const int64_t blocks = std::min<int64_t>((n / 2 + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(pack_bf16_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(),
                       reinterpret_cast<uint16_t *>(dst), src, n, head);
    SWP_LAUNCHED();
}

// ----------------------------------------------------------------- timer
KTimer &timer()
{
    // One per process: torch runs a backward on its own worker thread, so a
    // thread-local timer would miss the backward's launches.  Calls that
    // time concurrently from several threads share it (samples interleave).
    static KTimer t;
    return t;
}

void KTimer::reset()
{
    for (auto &m : marks) m.clear();
    for (auto &s : seen) s = 0;
    next = 0;
}

int KTimer::begin(int cls)
{
    if (!on) return -1;
    if (seen[cls]++ % stride != 0) return -1;
    if (next + 2 > pool.size()) {
        size_t grow = pool.size() < 256 ? 256 : pool.size();
        for (size_t k = 0; k < grow; ++k) {
            hipEvent_t e;
            SWP_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
    }
    int tok = int(next);
    next += 2;
    SWP_HIP(hipEventRecord(pool[tok], stream()));
    marks[cls].push_back({tok, tok + 1});
    return tok;
}

void KTimer::end(int cls, int tok)
{
    (void)cls;
    if (tok < 0) return;
    SWP_HIP(hipEventRecord(pool[tok + 1], stream()));
}

void KTimer::stats(int cls, double *ms, int64_t *n)
{
    double total = 0.0;
    for (auto &pr : marks[cls]) {
        SWP_HIP(hipEventSynchronize(pool[pr.second]));
        float t = 0.f;
        SWP_HIP(hipEventElapsedTime(&t, pool[pr.first], pool[pr.second]));
        total += t;
    }
    *ms = total;
    *n = int64_t(marks[cls].size());
}

}  // namespace swp
