// Kernels of the learned preconditioner: bfloat16 convolutions on the
// matrix cores (v_mfma_f32_16x16x32_bf16, float32 accumulation) and the
// element-wise passes of a 2-D U-Net's forward and backward.
//
// The convolution is an implicit GEMM over NHWC views: D[pixel, co] =
// sum_k A[pixel, k] B[k, co] with k = (tap, ci).  A 16 x 16 x 32 MFMA takes,
// in lane l, row l % 16 of A and column l % 16 of B, each with 8 of the 32
// k values (the same 8 in both, lane group l / 16 holding k = 8 (l / 16) ..
// + 7 here), and returns D[4 (l / 16) + e][l % 16] in element e.  With
// cin a multiple of 8 a lane's 8 k values are 8 consecutive channels of one
// tap: one 16-byte load from the NHWC input and one from the weights.
//
// The weight gradient reduces over pixels, which are not contiguous in
// either operand, so each 32-pixel chunk of dout and of the tap-shifted
// input is staged through LDS transposed to [channel][pixel]; the pixel
// range is split across workgroups and the float32 partial tiles are summed
// in a fixed order by a second kernel.
#include "unet.h"

#include <algorithm>
#include <cmath>

namespace swp {
namespace nn {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(p));
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }

__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- forward
// Block: 4 waves, 128 consecutive pixels of one image row (32 per wave, two
// 16-row A tiles) x 64 output channels (four 16-column B tiles).
// Grid: (ceil(W / 128), H, N * ceil(cout / 64)).
template <int KS, bool F32OUT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const uint16_t *__restrict__ in, int H, int W, int cin,
                                                       int is, const uint16_t *__restrict__ w, int cout,
                                                       const float *__restrict__ bias, int relu,
                                                       void *__restrict__ out, int os, int ncot)
{
    const int lane = int(threadIdx.x) & 63, wv = int(threadIdx.x) >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int y = int(blockIdx.y);
    const int n = int(blockIdx.z) / ncot, co0 = (int(blockIdx.z) % ncot) * 64;
    const int x0 = int(blockIdx.x) * 128 + wv * 32;
    constexpr int R = KS / 2;
    const int nchunk = (cin + 31) / 32;
    f32x4 acc[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < KS * KS; ++tap) {
        const int sy = y + tap / KS - R, dx = tap % KS - R;
        const bool rowok = sy >= 0 && sy < H;
        const int64_t rowbase = (int64_t(n) * H + sy) * W;
        for (int c = 0; c < nchunk; ++c) {
            const int ci = c * 32 + 8 * g;
            const bool ciok = ci < cin;
            bf16x8 a[2], b[4];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int sx = x0 + 16 * m + r16 + dx;
                a[m] = (rowok && ciok && sx >= 0 && sx < W) ? ld8(in + (rowbase + sx) * is + ci) : zero8();
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int co = co0 + 16 * t + r16;
                b[t] = (ciok && co < cout) ? ld8(w + (int64_t(co) * (KS * KS) + tap) * cin + ci) : zero8();
            }
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[m][t] = mma(a[m], b[t], acc[m][t]);
        }
    }
    const int64_t obase = (int64_t(n) * H + y) * W;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int co = co0 + 16 * t + r16;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int x = x0 + 16 * m + 4 * g + e;
                if (x >= W) continue;
                float v = acc[m][t][e] + bv;
                if (relu) v = fmaxf(v, 0.f);
                const int64_t i = (obase + x) * os + co;
                if constexpr (F32OUT) static_cast<float *>(out)[i] = v;
                else static_cast<uint16_t *>(out)[i] = f2bf(v);
            }
    }
}

// ---------------------------------------------------------------- weight gradient
// Block: 64 (co) x 64 (ci) for one tap over one pixel split; wave wv owns
// the 32 x 32 sub-tile (co (wv / 2) * 32, ci (wv % 2) * 32).
// Grid: (ceil(cin / 64), ceil(cout / 64), ks * ks * splits).
constexpr int kCP = 64;   // pixels per LDS chunk (two 32-pixel MFMA k-steps)
constexpr int kLP = 72;   // LDS row pitch (bf16): kCP pixels + 8, 16-byte aligned rows

template <int KS>
__global__ void __launch_bounds__(256) wgrad_kernel(const uint16_t *__restrict__ in, int H, int W, int cin, int is,
                                                    const uint16_t *__restrict__ dout, int cout, int ds,
                                                    float *__restrict__ partials, int splits, int64_t P,
                                                    int64_t per)
{
    __shared__ __attribute__((aligned(16))) uint16_t sD[64 * kLP];
    __shared__ __attribute__((aligned(16))) uint16_t sI[64 * kLP];
    const int tid = int(threadIdx.x), lane = tid & 63, wv = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int ci0 = int(blockIdx.x) * 64, co0 = int(blockIdx.y) * 64;
    const int tap = int(blockIdx.z) % (KS * KS), split = int(blockIdx.z) / (KS * KS);
    const int dy = tap / KS - KS / 2, dx = tap % KS - KS / 2;
    const int64_t pb = int64_t(split) * per, pe = std::min(P, pb + per);
    const int wco = (wv >> 1) * 32, wci = (wv & 1) * 32;
    const int lp = tid >> 3, lc = (tid & 7) * 8;
    const int64_t HW = int64_t(H) * W;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t p0 = pb; p0 < pe; p0 += kCP) {
#pragma unroll
        for (int h = 0; h < kCP / 32; ++h) {
            const int pl = lp + 32 * h;
            const int64_t p = p0 + pl;
            u32x4 vd = {0u, 0u, 0u, 0u}, vi = {0u, 0u, 0u, 0u};
            if (p < pe) {
                if (co0 + lc < cout) vd = *reinterpret_cast<const u32x4 *>(dout + p * ds + co0 + lc);
                if (ci0 + lc < cin) {
                    const int64_t nn = p / HW, rem = p % HW;
                    const int sy = int(rem / W) + dy, sx = int(rem % W) + dx;
                    if (sy >= 0 && sy < H && sx >= 0 && sx < W)
                        vi = *reinterpret_cast<const u32x4 *>(in + ((nn * H + sy) * W + sx) * is + ci0 + lc);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                sD[(lc + 2 * e) * kLP + pl] = uint16_t(vd[e] & 0xffffu);
                sD[(lc + 2 * e + 1) * kLP + pl] = uint16_t(vd[e] >> 16);
                sI[(lc + 2 * e) * kLP + pl] = uint16_t(vi[e] & 0xffffu);
                sI[(lc + 2 * e + 1) * kLP + pl] = uint16_t(vi[e] >> 16);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kCP / 32; ++k) {
            bf16x8 A[2], Bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) A[a] = ld8(sD + (wco + 16 * a + r16) * kLP + 32 * k + 8 * g);
#pragma unroll
            for (int b = 0; b < 2; ++b) Bv[b] = ld8(sI + (wci + 16 * b + r16) * kLP + 32 * k + 8 * g);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = mma(A[a], Bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    float *dst = partials + (int64_t(split) * (KS * KS) + tap) * cout * cin;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int ci = ci0 + wci + 16 * b + r16;
            if (ci >= cin) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int co = co0 + wco + 16 * a + 4 * g + e;
                if (co < cout) dst[int64_t(co) * cin + ci] = acc[a][b][e];
            }
        }
}

// dw[co, tap, ci] = sum over splits of partials[split, tap, co, ci].
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float *__restrict__ partials, int splits, int taps,
                                                           int cout, int cin, float *__restrict__ dw)
{
    const int64_t n = int64_t(taps) * cout * cin;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int ci = int(i % cin);
        const int64_t t2 = i / cin;
        const int tap = int(t2 % taps), co = int(t2 / taps);
        float s = 0.f;
        for (int sp = 0; sp < splits; ++sp) s += partials[((int64_t(sp) * taps + tap) * cout + co) * cin + ci];
        dw[i] = s;
    }
}

// partial[split, c] = sum over the split's pixels of x[p, c]; then out[c].
// Thread t takes channel c0 + t % CB and pixel sub-lane t / CB (J = 256 / CB
// sub-lanes stride the split's pixels), the J lane sums meet in LDS.
__global__ void __launch_bounds__(256) colsum_kernel(const uint16_t *__restrict__ x, int xs, int C, int64_t P,
                                                     int64_t per, float *__restrict__ partial)
{
    __shared__ float red[256];
    const int64_t pb = int64_t(blockIdx.x) * per, pe = std::min(P, pb + per);
    const int CB = C < 256 ? C : 256;
    const int J = 256 / CB;
    const int t = int(threadIdx.x), j = t / CB, cl = t % CB;
    for (int c0 = 0; c0 < C; c0 += CB) {
        const int c = c0 + cl;
        float s = 0.f;
        if (j < J && c < C)
            for (int64_t p = pb + j; p < pe; p += J) s += bf2f(x[p * xs + c]);
        red[t] = s;
        __syncthreads();
        if (j == 0 && c < C) {
            float tot = 0.f;
            for (int k = 0; k < J; ++k) tot += red[k * CB + cl];
            partial[int64_t(blockIdx.x) * C + c] = tot;
        }
        __syncthreads();
    }
}

// One block per channel: a fixed-order tree over the splits.
__global__ void __launch_bounds__(256) colsum_reduce_kernel(const float *__restrict__ partial, int splits, int C,
                                                            float *__restrict__ out)
{
    __shared__ float red[256];
    const int c = int(blockIdx.x);
    float s = 0.f;
    for (int sp = int(threadIdx.x); sp < splits; sp += 256) s += partial[int64_t(sp) * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void __launch_bounds__(256) flip_kernel(const uint16_t *__restrict__ w, int cout, int cin, int taps,
                                                   uint16_t *__restrict__ wt)
{
    const int64_t n = int64_t(cout) * taps * cin;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        // i indexes wt[ci, tap, co]
        const int co = int(i % cout);
        const int64_t t2 = i / cout;
        const int tap = int(t2 % taps), ci = int(t2 / taps);
        wt[i] = w[(int64_t(co) * taps + (taps - 1 - tap)) * cin + ci];
    }
}

// ---------------------------------------------------------------- element-wise
__global__ void __launch_bounds__(256) relu_bwd_kernel(uint16_t *__restrict__ d, int dsd,
                                                       const uint16_t *__restrict__ act, int as, int64_t P, int C)
{
    const int64_t n = P * C;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int64_t p = i / C;
        const int c = int(i % C);
        if (!(bf2f(act[p * as + c]) > 0.f)) d[p * dsd + c] = 0;
    }
}

// H, W: the input (fine) extents.
__global__ void __launch_bounds__(256) pool2_fwd_kernel(const uint16_t *__restrict__ in, int is, int N, int H, int W,
                                                        int C, uint16_t *__restrict__ out, int os)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = int64_t(N) * Ho * Wo * C;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int c = int(i % C);
        const int64_t q = i / C;
        const int xo = int(q % Wo), yo = int((q / Wo) % Ho);
        const int64_t nn = q / (int64_t(Wo) * Ho);
        const int64_t p00 = (nn * H + 2 * yo) * W + 2 * xo;
        const float s = (bf2f(in[p00 * is + c]) + bf2f(in[(p00 + 1) * is + c])) +
                        (bf2f(in[(p00 + W) * is + c]) + bf2f(in[(p00 + W + 1) * is + c]));
        out[q * os + c] = f2bf(0.25f * s);
    }
}

__global__ void __launch_bounds__(256) pool2_bwd_kernel(const uint16_t *__restrict__ dout, int dsd, int N, int H,
                                                        int W, int C, uint16_t *__restrict__ din, int dis)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = int64_t(N) * H * W * C;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int c = int(i % C);
        const int64_t p = i / C;
        const int x = int(p % W), y = int((p / W) % H);
        const int64_t nn = p / (int64_t(W) * H);
        const int64_t q = (nn * Ho + y / 2) * Wo + x / 2;
        din[p * dis + c] = f2bf(bf2f(din[p * dis + c]) + 0.25f * bf2f(dout[q * dsd + c]));
    }
}

// H, W: the coarse (input) extents.
__global__ void __launch_bounds__(256) up2_fwd_kernel(const uint16_t *__restrict__ in, int is, int N, int H, int W,
                                                      int C, uint16_t *__restrict__ out, int os)
{
    const int Hf = 2 * H, Wf = 2 * W;
    const int64_t n = int64_t(N) * Hf * Wf * C;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int c = int(i % C);
        const int64_t p = i / C;
        const int x = int(p % Wf), y = int((p / Wf) % Hf);
        const int64_t nn = p / (int64_t(Wf) * Hf);
        out[p * os + c] = in[((nn * H + y / 2) * W + x / 2) * is + c];
    }
}

__global__ void __launch_bounds__(256) up2_bwd_kernel(const uint16_t *__restrict__ dout, int dsd, int N, int H, int W,
                                                      int C, uint16_t *__restrict__ din, int dis)
{
    const int Wf = 2 * W;
    const int64_t n = int64_t(N) * H * W * C;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int c = int(i % C);
        const int64_t q = i / C;
        const int x = int(q % W), y = int((q / W) % H);
        const int64_t nn = q / (int64_t(W) * H);
        const int64_t f00 = (nn * 2 * H + 2 * y) * Wf + 2 * x;
        const float s = (bf2f(dout[f00 * dsd + c]) + bf2f(dout[(f00 + 1) * dsd + c])) +
                        (bf2f(dout[(f00 + Wf) * dsd + c]) + bf2f(dout[(f00 + Wf + 1) * dsd + c]));
        din[q * dis + c] = f2bf(s);
    }
}

struct Planes {
    const float *p[4];
    float s[4];
};

__global__ void __launch_bounds__(256) pack_input_kernel(Planes pl, int nsrc, int64_t P, uint16_t *__restrict__ x,
                                                         int xs)
{
    const int64_t n = P * xs;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const int c = int(i % xs);
        const int64_t p = i / xs;
        float v = 0.f;
        if (c < nsrc) v = pl.s[c] * pl.p[c][p];
        x[i] = f2bf(v);
    }
}

__global__ void __launch_bounds__(256) mse_kernel(const float *__restrict__ out, int os, const float *__restrict__ tgt,
                                                  const float *__restrict__ mask, int64_t P, double count,
                                                  uint16_t *__restrict__ dout, int dsd, double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    const float sc = float(2.0 / count);
    for (int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x; p < P; p += int64_t(gridDim.x) * 256) {
        const float r = out[p * os] - tgt[p];
        const float m = mask ? mask[p] : 1.f;
        acc += double(r) * double(r) * double(m);
        dout[p * dsd] = f2bf(sc * r * m);
        for (int c = 1; c < dsd; ++c) dout[p * dsd + c] = 0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_kernel(const double *__restrict__ partial, int np, double scale,
                                                  double *__restrict__ out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = int(threadIdx.x); i < np; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = scale * red[0];
}

__global__ void __launch_bounds__(256) adam_kernel(float *__restrict__ w, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v,
                                                   uint16_t *__restrict__ wbf, int64_t n, float lr, float b1,
                                                   float b2, float eps, float c1, float c2)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wbf) wbf[i] = f2bf(wi);
    }
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float *__restrict__ s, uint16_t *__restrict__ d, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) d[i] = f2bf(s[i]);
}

unsigned blocks_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384))); }

void check_view(const void *p, int stride, int channels, const char *what)
{
    SWP_CHECK(p != nullptr, what, ": null pointer");
    SWP_CHECK((reinterpret_cast<uintptr_t>(p) & 15) == 0, what, ": pointer not 16-byte aligned");
    SWP_CHECK(stride % 8 == 0 && channels % 8 == 0 && channels <= stride && channels > 0, what,
              ": channels (", channels, ") and stride (", stride, ") must be positive multiples of 8, channels <= stride");
}

int wgrad_splits(int64_t P, int cin, int cout, int ks)
{
    const int64_t tiles = int64_t((cin + 63) / 64) * ((cout + 63) / 64) * ks * ks;
    int64_t s = std::max<int64_t>(1, 2048 / tiles);
    s = std::min<int64_t>(s, std::max<int64_t>(1, P / 256));
    return int(std::min<int64_t>(s, 64));
}

constexpr int kColSplits = 512;

}  // namespace

void conv_fwd(const uint16_t *in, int N, int H, int W, int cin, int in_stride, const uint16_t *w, int cout, int ks,
              const float *bias, bool relu, void *out, int out_stride, bool out_f32)
{
    SWP_CHECK(ks == 1 || ks == 3, "conv: kernel size must be 1 or 3");
    SWP_CHECK(N >= 1 && H >= 1 && W >= 1 && N * int64_t((cout + 63) / 64) <= 65535 && H <= 65535,
              "conv: bad extents");
    check_view(in, in_stride, cin, "conv input");
    SWP_CHECK(w && (reinterpret_cast<uintptr_t>(w) & 15) == 0, "conv: weights null or not 16-byte aligned");
    SWP_CHECK(out && cout >= 1 && cout <= out_stride, "conv: bad output view");
    const int ncot = (cout + 63) / 64;
    const dim3 grid(unsigned((W + 127) / 128), unsigned(H), unsigned(N * ncot));
    const int r = relu ? 1 : 0;
#define SWP_CONV_LAUNCH(KS, F32)                                                                              \
    hipLaunchKernelGGL((conv_fwd_kernel<KS, F32>), grid, dim3(256), 0, stream(), in, H, W, cin, in_stride, w, \
                       cout, bias, r, out, out_stride, ncot)
    if (ks == 3) {
        if (out_f32) SWP_CONV_LAUNCH(3, true);
        else SWP_CONV_LAUNCH(3, false);
    } else {
        if (out_f32) SWP_CONV_LAUNCH(1, true);
        else SWP_CONV_LAUNCH(1, false);
    }
#undef SWP_CONV_LAUNCH
    SWP_LAUNCHED();
}

int64_t conv_wgrad_partials(int N, int H, int W, int cin, int cout, int ks)
{
    const int64_t P = int64_t(N) * H * W;
    return int64_t(wgrad_splits(P, cin, cout, ks)) * ks * ks * cout * cin + int64_t(kColSplits) * cout;
}

void conv_wgrad(const uint16_t *in, int N, int H, int W, int cin, int in_stride, const uint16_t *dout, int cout,
                int dout_stride, int ks, float *dw, float *db, float *partials)
{
    SWP_CHECK(ks == 1 || ks == 3, "wgrad: kernel size must be 1 or 3");
    SWP_CHECK(N >= 1 && H >= 1 && W >= 1, "wgrad: bad extents");
    check_view(in, in_stride, cin, "wgrad input");
    check_view(dout, dout_stride, cout, "wgrad dout");
    SWP_CHECK(dw && partials, "wgrad: null output or scratch");
    const int64_t P = int64_t(N) * H * W;
    const int splits = wgrad_splits(P, cin, cout, ks);
    int64_t per = (P + splits - 1) / splits;
    per = (per + kCP - 1) / kCP * kCP;
    const dim3 grid(unsigned((cin + 63) / 64), unsigned((cout + 63) / 64), unsigned(ks * ks * splits));
    if (ks == 3)
        hipLaunchKernelGGL(wgrad_kernel<3>, grid, dim3(256), 0, stream(), in, H, W, cin, in_stride, dout, cout,
                           dout_stride, partials, splits, P, per);
    else
        hipLaunchKernelGGL(wgrad_kernel<1>, grid, dim3(256), 0, stream(), in, H, W, cin, in_stride, dout, cout,
                           dout_stride, partials, splits, P, per);
    SWP_LAUNCHED();
    const int64_t n = int64_t(ks) * ks * cout * cin;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(blocks_for(n)), dim3(256), 0, stream(), partials, splits, ks * ks,
                       cout, cin, dw);
    SWP_LAUNCHED();
    if (db) {
        float *cp = partials + int64_t(splits) * ks * ks * cout * cin;
        const int64_t cper = (P + kColSplits - 1) / kColSplits;
        hipLaunchKernelGGL(colsum_kernel, dim3(kColSplits), dim3(256), 0, stream(), dout, dout_stride, cout, P, cper,
                           cp);
        SWP_LAUNCHED();
        hipLaunchKernelGGL(colsum_reduce_kernel, dim3(unsigned(cout)), dim3(256), 0, stream(), cp, kColSplits, cout,
                           db);
        SWP_LAUNCHED();
    }
}

void conv_flip(const uint16_t *w, int cout, int cin, int ks, uint16_t *wt)
{
    SWP_CHECK(w && wt && w != wt, "flip: bad buffers");
    const int64_t n = int64_t(cout) * ks * ks * cin;
    hipLaunchKernelGGL(flip_kernel, dim3(blocks_for(n)), dim3(256), 0, stream(), w, cout, cin, ks * ks, wt);
    SWP_LAUNCHED();
}

void relu_bwd(uint16_t *d, int d_stride, const uint16_t *act, int act_stride, int64_t P, int C)
{
    SWP_CHECK(d && act && C >= 1 && C <= d_stride && C <= act_stride, "relu_bwd: bad views");
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(blocks_for(P * C)), dim3(256), 0, stream(), d, d_stride, act,
                       act_stride, P, C);
    SWP_LAUNCHED();
}

void pool2_fwd(const uint16_t *in, int in_stride, int N, int H, int W, int C, uint16_t *out, int out_stride)
{
    SWP_CHECK(H % 2 == 0 && W % 2 == 0, "pool2: extents must be even");
    SWP_CHECK(in && out && C <= in_stride && C <= out_stride, "pool2: bad views");
    hipLaunchKernelGGL(pool2_fwd_kernel, dim3(blocks_for(int64_t(N) * H * W * C / 4)), dim3(256), 0, stream(), in,
                       in_stride, N, H, W, C, out, out_stride);
    SWP_LAUNCHED();
}

void pool2_bwd_acc(const uint16_t *dout, int dout_stride, int N, int H, int W, int C, uint16_t *din, int din_stride)
{
    SWP_CHECK(H % 2 == 0 && W % 2 == 0, "pool2: extents must be even");
    SWP_CHECK(dout && din && C <= dout_stride && C <= din_stride, "pool2: bad views");
    hipLaunchKernelGGL(pool2_bwd_kernel, dim3(blocks_for(int64_t(N) * H * W * C)), dim3(256), 0, stream(), dout,
                       dout_stride, N, H, W, C, din, din_stride);
    SWP_LAUNCHED();
}

void up2_fwd(const uint16_t *in, int in_stride, int N, int H, int W, int C, uint16_t *out, int out_stride)
{
    SWP_CHECK(in && out && C <= in_stride && C <= out_stride, "up2: bad views");
    hipLaunchKernelGGL(up2_fwd_kernel, dim3(blocks_for(int64_t(N) * H * W * C * 4)), dim3(256), 0, stream(), in,
                       in_stride, N, H, W, C, out, out_stride);
    SWP_LAUNCHED();
}

void up2_bwd(const uint16_t *dout, int dout_stride, int N, int H, int W, int C, uint16_t *din, int din_stride)
{
    SWP_CHECK(dout && din && C <= dout_stride && C <= din_stride, "up2: bad views");
    hipLaunchKernelGGL(up2_bwd_kernel, dim3(blocks_for(int64_t(N) * H * W * C)), dim3(256), 0, stream(), dout,
                       dout_stride, N, H, W, C, din, din_stride);
    SWP_LAUNCHED();
}

void pack_input(const float *const *src, const float *scale, int nsrc, int64_t P, uint16_t *x, int stride)
{
    SWP_CHECK(nsrc >= 1 && nsrc <= 4 && nsrc <= stride && x, "pack_input: bad arguments");
    Planes pl{};
    for (int c = 0; c < nsrc; ++c) {
        SWP_CHECK(src[c] != nullptr, "pack_input: null plane");
        pl.p[c] = src[c];
        pl.s[c] = scale[c];
    }
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(P * stride)), dim3(256), 0, stream(), pl, nsrc, P, x,
                       stride);
    SWP_LAUNCHED();
}

void mse(const float *out, int out_stride, const float *target, const float *mask, int64_t P, double count,
         uint16_t *dout, int dout_stride, double *loss, double *partials)
{
    SWP_CHECK(out && target && dout && loss && partials && count > 0, "mse: bad arguments");
    const unsigned nb = unsigned(std::max<int64_t>(1, std::min<int64_t>((P + 1023) / 1024, 1024)));
    hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, stream(), out, out_stride, target, mask, P, count, dout,
                       dout_stride, partials);
    SWP_LAUNCHED();
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream(), partials, int(nb), 1.0 / count, loss);
    SWP_LAUNCHED();
}

void adam(float *w, const float *g, float *m, float *v, uint16_t *wbf, int64_t n, float lr, float b1, float b2,
          float eps, int step)
{
    SWP_CHECK(w && g && m && v && step >= 1, "adam: bad arguments");
    const float c1 = float(1.0 - std::pow(double(b1), step)), c2 = float(1.0 - std::pow(double(b2), step));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, stream(), w, g, m, v, wbf, n, lr, b1, b2, eps,
                       c1, c2);
    SWP_LAUNCHED();
}

void to_bf16(const float *src, uint16_t *dst, int64_t n)
{
    SWP_CHECK(src && dst, "to_bf16: null buffer");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, stream(), src, dst, n);
    SWP_LAUNCHED();
}

}  // namespace nn
}  // namespace swp
