// Shared device-side declarations of libswp_core.
#pragma once

#include "core.h"

namespace swp {

// Arguments of the per-shot point kernel: an optional scatter-add of
// w[b, s, it] (times sign) into u_add at add_loc, then an optional gather of
// u_get at get_loc into out[b, r, it].  Coordinates are (x[, y], z).
struct PointArgs {
    float *u_add;
    const float *add_w;
    const int32_t *add_loc;
    int nadd;
    float sign;
    const float *u_get;
    float *get_out;
    const int32_t *get_loc;
    int nget;
    int get_accumulate;
    int it, nt, ndim;
    int nz, ny, nx;
    int px;      // row pitch of u_add / u_get (0: nx)
    int64_t n;   // cells per shot
};

void launch_points(const PointArgs &a, int B);

// Boundary strips hold fp32 or, with SWP_FLAG_BF16_STRIPS, bfloat16: round
// to nearest, ties to even, a NaN kept a quiet NaN (the rounding of the
// reference's __float2bfloat16).  A strip pointer is typed float* either way;
// strip_at / strip_put / strip_get count elements of the stored kind.
// XCD-local block numbering.  The dispatcher hands workgroups to the eight
// XCDs round-robin in launch order (workgroup i to XCD i mod 8) and each XCD
// has its own L2, so with the plain numbering a tile's row neighbours -- whose
// halo rows it re-reads -- sit in other L2s.  Renumbered, XCD k takes the k-th
// contiguous eighth of the grid (row-major), and neighbouring tiles share an
// L2.  The map is a bijection of [0, n): the first n mod 8 XCDs take one block
// more.
__device__ __forceinline__ uint32_t xcd_local_id(uint32_t pid, uint32_t n)
{
    const uint32_t xcd = pid % 8u, k = pid / 8u, q = n / 8u, r = n % 8u;
    return xcd < r ? xcd * (q + 1u) + k : r * (q + 1u) + (xcd - r) * q + k;
}

struct BlockXYZ {
    int x, y, z;
};

// This block's (x, y, z) in its grid under the XCD-local numbering.
__device__ __forceinline__ BlockXYZ xcd_block()
{
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t n = gx * gy * gridDim.z;
    const uint32_t l = xcd_local_id(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), n);
    return {int(l % gx), int((l / gx) % gy), int(l / (gx * gy))};
}

__host__ __device__ inline uint16_t bf16_rne(float f)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__host__ __device__ inline float bf16_float(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

__host__ __device__ inline float *strip_at(float *s, int64_t i, bool bf16)
{
    return bf16 ? reinterpret_cast<float *>(reinterpret_cast<uint16_t *>(s) + i) : s + i;
}

__host__ __device__ inline const float *strip_at(const float *s, int64_t i, bool bf16)
{
    return bf16 ? reinterpret_cast<const float *>(reinterpret_cast<const uint16_t *>(s) + i) : s + i;
}

__device__ __forceinline__ void strip_put(float *s, int64_t i, float v, bool bf16)
{
    if (bf16) reinterpret_cast<uint16_t *>(s)[i] = bf16_rne(v);
    else s[i] = v;
}

__device__ __forceinline__ float strip_get(const float *s, int64_t i, bool bf16)
{
    return bf16 ? bf16_float(reinterpret_cast<const uint16_t *>(s)[i]) : s[i];
}

// dst[i] = bf16_rne(src[i]), i < n: the forward saves a step's strips in
// float32 (scattered 2-byte stores from the stencil kernels measured 28%
// slower on a 3-D forward step) and packs them in one coalesced pass.
void pack_bf16(float *dst, const float *src, int64_t n);

// Row-pitched copies between the caller's arrays and the core's: rows rows
// of width floats, dst row r at dst + r * dpitch, src row r at src + r * spitch.
// pitch_copy writes dst (and zeroes dst's columns [width, dpitch)),
// pitch_add adds src into dst.
void pitch_copy(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width);
void pitch_add(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width);

// Whether equation 0 (acoustic2d), 1 (acoustic3d), 2 (elastic2d) is built.
bool eq_available(int eq);

}  // namespace swp
