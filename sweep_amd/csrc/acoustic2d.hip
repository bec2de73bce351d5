// 2-D constant-density acoustic path on gfx950: forward step with the cpmlr
// CPML, exact discrete adjoint, boundary-saving reverse reconstruction with
// fused u_tt imaging, strip save/restore, full-storage and checkpoint
// imaging, illumination.  FP32 on the vector ALUs; wave64 rows along x.
//
// Arithmetic follows the reference's acoustic2d kernels term for term
// (csrc/cuda/equations/acoustic2d/forward_split.cu:39-168,
// adjoint_split.cu:27-190, kernels.cuh:472-574, kernels.cu:3-189); the launch
// structure, memory layout and time loops are this file's own.
#include "acoustic2d.h"
#include "acoustic2d_zm.h"
#include "revolve.h"

#include <algorithm>
#include <cstdlib>

namespace swp {
namespace a2d {

// ---------------------------------------------------------------- geometry
Geo make_geo(const swp_problem &p)
{
    Geo g{};
    g.nz = p.nz;
    g.nx = p.nx;
    g.B = p.B;
    g.M = p.M;
    g.nt = p.nt;
    g.pzl = p.pad[0];
    g.pzh = p.pad[1];
    g.pxl = p.pad[4];
    g.pxh = p.pad[5];
    g.X0 = g.pxl + g.M;
    g.X1 = g.nx - g.pxh - g.M;
    g.Z0 = g.pzl + g.M;
    g.Z1 = g.nz - g.pzh - g.M;
    g.abcn = p.abcn;
    g.dt = p.dt;
    g.dx = p.dx;
    g.dz = p.dz;
    // rows of every array the core allocates start on 128-byte lines
    g.px = (g.nx + 31) / 32 * 32;
    g.n = int64_t(g.nz) * g.px;
    g.nm = int64_t(g.nz) * g.nx;
    g.mstride = p.model_batched ? g.n : 0;
    g.W = g.M + 1;
    g.nxp = g.X1 - g.X0;
    g.nzp = g.Z1 - g.Z0;
    g.ns = 2LL * g.W * g.nxp + 2LL * g.W * g.nzp;
    g.sbf16 = (p.flags & SWP_FLAG_BF16_STRIPS) != 0;
    g.inv_dx2 = fd::scale_of(g.dx * g.dx);
    g.inv_dz2 = fd::scale_of(g.dz * g.dz);
    g.inv_dx = fd::scale_of(g.dx);
    g.inv_dz = fd::scale_of(g.dz);
    g.dt2 = g.dt * g.dt;
    g.inv_dt2 = fd::scale_of(g.dt * g.dt);
    return g;
}

Coef make_coef(const swp_problem &p, const float *lap_host, const float *grad_host)
{
    Coef c{};
    const int M = p.M;
    SWP_CHECK(M >= 1 && M <= kMaxM, "acoustic2d: half order M=", M, " outside [1, ", kMaxM, "]");
    for (int k = 0; k <= M; ++k) {
        c.lap[k] = lap_host[k];
        c.grad[k] = grad_host[k];
    }
    // The fixed-order forward stencils of orders 2..8 hold their weights as
    // float32 constants (operators/laplace.cuh:157-341, gradient.cuh:56-78):
    // c_M closes the zero sum in float32 arithmetic.
    const float f2 = 2.0f;
    switch (M) {
    case 1:
        c.slap[0] = 1.0f;
        c.sgrad[0] = 0.5f;
        break;
    case 2: {
        const float c0 = 5.0f / f2, c1 = 4.0f / 3.0f;
        c.slap_c0 = c0;
        c.slap[0] = c1;
        c.slap[1] = 0.5f * c0 - c1;
        c.sgrad[0] = 8.0f / 12.0f;
        c.sgrad[1] = -1.0f / 12.0f;
        break;
    }
    case 3: {
        const float c0 = 49.0f / 18.0f, c1 = 3.0f / 2.0f, c2 = -3.0f / 20.0f;
        c.slap[0] = c1;
        c.slap[1] = c2;
        c.slap[2] = (0.5f * c0 - c1) - c2;
        c.sgrad[0] = 3.0f / 4.0f;
        c.sgrad[1] = -3.0f / 20.0f;
        c.sgrad[2] = 1.0f / 60.0f;
        break;
    }
    case 4: {
        const float c0 = 205.0f / 72.0f, c1 = 8.0f / 5.0f, c2 = -1.0f / 5.0f, c3 = 8.0f / 315.0f;
        c.slap[0] = c1;
        c.slap[1] = c2;
        c.slap[2] = c3;
        c.slap[3] = ((0.5f * c0 - c1) - c2) - c3;
        c.sgrad[0] = 4.0f / 5.0f;
        c.sgrad[1] = -1.0f / 5.0f;
        c.sgrad[2] = 4.0f / 105.0f;
        c.sgrad[3] = -1.0f / 280.0f;
        break;
    }
    default:
        break;
    }
    return c;
}

// ---------------------------------------------------------------- stencils
// Forward second derivative along stride s: order 4 in summed form, the other
// fixed orders in difference form, runtime order from the table.
template <int MS>
__device__ __forceinline__ float lap_fwd(const Coef &c, const float *__restrict__ u, int64_t i,
                                         int64_t s, fd::Scale inv_h2, int M)
{
    const float u0 = u[i];
    if constexpr (MS == 1) {
        return fd::scaled((u[i - s] - u0) + (u[i + s] - u0), inv_h2);
    } else if constexpr (MS == 2) {
        return fd::scaled(c.slap[0] * (u[i - s] + u[i + s]) + c.slap[1] * (u[i - 2 * s] + u[i + 2 * s]) -
                          c.slap_c0 * u0, inv_h2);
    } else if constexpr (MS == 3 || MS == 4) {
        float acc = 0.f;
#pragma unroll
        for (int k = 1; k <= MS; ++k)
            acc += c.slap[k - 1] * ((u[i - k * s] - u0) + (u[i + k * s] - u0));
        return fd::scaled(acc, inv_h2);
    } else {
        float acc = 0.f;
        for (int k = 1; k <= M; ++k) acc += c.lap[k] * ((u[i + k * s] - u0) + (u[i - k * s] - u0));
        return fd::scaled(acc, inv_h2);
    }
}

// Centred first derivative of f(k), k = -M..M (gradient.cuh:47-88).
template <int MS, class F>
__device__ __forceinline__ float d1(const Coef &c, F f, fd::Scale inv_h, int M)
{
    if constexpr (MS == 1) {
        return fd::scaled_half(f(1) - f(-1), inv_h);
    } else if constexpr (MS >= 2 && MS <= 4) {
        float acc = 0.f;
#pragma unroll
        for (int k = 1; k <= MS; ++k) acc += c.sgrad[k - 1] * (f(k) - f(-k));
        return fd::scaled(acc, inv_h);
    } else {
        float acc = 0.f;
        for (int k = 1; k <= M; ++k) acc += c.grad[k] * (f(k) - f(-k));
        return fd::scaled(acc, inv_h);
    }
}

// ---------------------------------------------------------------- forward
// One thread per computed cell [M, n-M) of one shot; blocks of 64 x 4 cells
// (one wave64 per row segment, coalesced along x).  Interior cells take the
// bare leapfrog; cells within the CPML compute band of an axis run that axis'
// cpmlr update.  Optionally saves u_now into the shot's boundary strips and
// writes the full-mode image v^2 * w_sum.
template <int MS>
__global__ void __launch_bounds__(256) fwd_kernel(Geo g, Coef c, Prof pr, StepFields f,
                                                  float *__restrict__ u_this,
                                                  float *__restrict__ strip)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = M + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = M + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = int64_t(iz) * g.px + ix;
    const float *__restrict__ un = f.u_now + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const int64_t nx = g.px;   // row pitch

    if (strip) {   // boundary save of u_now (driver_traits.cuh:248-265, strip.cuh)
        float *s = strip_at(strip, int64_t(b) * g.ns, g.sbf16);
        const float v = un[i];
        const bool inx = ix >= g.X0 && ix < g.X1, inz = iz >= g.Z0 && iz < g.Z1;
        if (inx && iz >= g.Z0 && iz < g.Z0 + g.W)
            strip_put(s, int64_t(iz - g.Z0) * g.nxp + (ix - g.X0), v, g.sbf16);
        if (inx && iz >= g.Z1 - g.W && iz < g.Z1)
            strip_put(s, int64_t(g.W) * g.nxp + int64_t(iz - (g.Z1 - g.W)) * g.nxp + (ix - g.X0), v, g.sbf16);
        if (inz && ix >= g.X0 && ix < g.X0 + g.W)
            strip_put(s, 2LL * g.W * g.nxp + int64_t(iz - g.Z0) * g.W + (ix - g.X0), v, g.sbf16);
        if (inz && ix >= g.X1 - g.W && ix < g.X1)
            strip_put(s, 2LL * g.W * g.nxp + int64_t(g.nzp) * g.W + int64_t(iz - g.Z0) * g.W + (ix - (g.X1 - g.W)),
                      v, g.sbf16);
    }

    const float lap_x = lap_fwd<MS>(c, un, i, 1, g.inv_dx2, M);
    const float lap_z = lap_fwd<MS>(c, un, i, nx, g.inv_dz2, M);
    const float v = vp[i];
    const float v2_dt2 = (v * v) * g.dt * g.dt;
    const bool cx = ix < g.pxl + 2 * M || ix >= g.nx - g.pxh - 2 * M;
    const bool cz = iz < g.pzl + 2 * M || iz >= g.nz - g.pzh - 2 * M;
    float w_sum;
    if (!cx && !cz) {
        w_sum = lap_x + lap_z;
    } else {
        w_sum = 0.0f;
        if (cx) {
            const float *__restrict__ psix = f.psix + off;
            const float ax_ = pr.ax[ix], bx_ = pr.bx[ix], dbx_ = pr.dbx[ix];
            const float dudx = d1<MS>(c, [&](int k) { return un[i + k]; }, g.inv_dx, M);
            const float dap = d1<MS>(c, [&](int k) { return pr.ax[ix + k] * psix[i + k]; }, g.inv_dx, M);
            const float tmpx = ((1.0f + bx_) * lap_x + dbx_ * dudx) + dap;
            const float zx = f.zetax[off + i];
            w_sum += (1.0f + bx_) * tmpx + ax_ * zx;
            f.psixn[off + i] = bx_ * dudx + ax_ * psix[i];
            f.zetax[off + i] = bx_ * tmpx + ax_ * zx;
        } else {
            w_sum += lap_x;
        }
        if (cz) {
            const float *__restrict__ psiz = f.psiz + off;
            const float az_ = pr.az[iz], bz_ = pr.bz[iz], dbz_ = pr.dbz[iz];
            const float dudz = d1<MS>(c, [&](int k) { return un[i + k * nx]; }, g.inv_dz, M);
            const float dap = d1<MS>(c, [&](int k) { return pr.az[iz + k] * psiz[i + k * nx]; }, g.inv_dz, M);
            const float tmpz = ((1.0f + bz_) * lap_z + dbz_ * dudz) + dap;
            const float zz = f.zetaz[off + i];
            w_sum += (1.0f + bz_) * tmpz + az_ * zz;
            f.psizn[off + i] = bz_ * dudz + az_ * psiz[i];
            f.zetaz[off + i] = bz_ * tmpz + az_ * zz;
        } else {
            w_sum += lap_z;
        }
    }
    f.u_next[off + i] = 2.0f * un[i] - f.u_prev[off + i] + v2_dt2 * w_sum;
    if (u_this) u_this[off + i] = (v * v) * w_sum;
}

// ---------------------------------------------------------------- adjoint
// Exact transpose of fwd_kernel (+ source/record), one thread per computed
// cell: lambda_next = 2 lambda - lambda_prev + L^T(lambda), with the CPML
// memory variables double-buffered (read old, write new).  img != nullptr
// fuses the lagged full-mode imaging grad += 2 dt^2 img lambda / v.
template <int MS>
__global__ void __launch_bounds__(256) adj_kernel(Geo g, Coef c, Prof pr, AdjFields f,
                                                  const float *__restrict__ img,
                                                  float *__restrict__ grad)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = M + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = M + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = int64_t(iz) * g.px + ix;
    const int64_t nx = g.px;   // row pitch
    const float *__restrict__ ln = f.l_now + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const float dt2 = g.dt2;
    const float gun = ln[i];

    if (grad && ix >= g.X0 && ix < g.X1 && iz >= g.Z0 && iz < g.Z1)
        grad[off + i] += 2.f * g.dt * g.dt * img[off + i] * gun / vp[i];

    const float v0 = vp[i];
    const float gw = v0 * v0 * dt2 * gun;
    const bool pure = ix >= g.X0 + M + 1 && ix < g.X1 - M - 1 && iz >= g.Z0 + M + 1 && iz < g.Z1 - M - 1;
    if (pure) {
        float lapx = 0.0f, lapz = 0.0f;
        auto GW = [&](int64_t n) { return vp[n] * vp[n] * dt2 * ln[n]; };
        if constexpr (MS > 0) {
#pragma unroll
            for (int k = 1; k <= MS; ++k) {
                lapx += c.lap[k] * ((GW(i + k) - gw) + (GW(i - k) - gw));
                lapz += c.lap[k] * ((GW(i + k * nx) - gw) + (GW(i - k * nx) - gw));
            }
        } else {
            for (int k = 1; k <= M; ++k) {
                lapx += c.lap[k] * ((GW(i + k) - gw) + (GW(i - k) - gw));
                lapz += c.lap[k] * ((GW(i + k * nx) - gw) + (GW(i - k * nx) - gw));
            }
        }
        f.l_next[off + i] = 2.0f * gun - f.l_prev[off + i] + lapx * g.inv_dx2.inv + lapz * g.inv_dz2.inv;
        return;
    }
    const float *__restrict__ psix = f.psix + off;
    const float *__restrict__ psiz = f.psiz + off;
    const float *__restrict__ zetax = f.zetax + off;
    const float *__restrict__ zetaz = f.zetaz + off;
    const float ax_ = pr.ax[ix], az_ = pr.az[iz], bx_ = pr.bx[ix], bz_ = pr.bz[iz];
    const float lx0 = (1.0f + bx_) * (bx_ * zetax[i] + (1.0f + bx_) * gw);
    const float lz0 = (1.0f + bz_) * (bz_ * zetaz[i] + (1.0f + bz_) * gw);
    float Lx = 0.f, Dgx = 0.f, Dqx = 0.f, Lz = 0.f, Dgz = 0.f, Dqz = 0.f;
    for (int k = 1; k <= M; ++k) {
        {
            const int64_t np = i + k, nm = i - k;
            const int cp = ix + k, cm = ix - k;
            const float bp = pr.bx[cp], bm = pr.bx[cm];
            const float tp = bp * zetax[np] + (1.0f + bp) * (vp[np] * vp[np] * dt2) * ln[np];
            const float tm = bm * zetax[nm] + (1.0f + bm) * (vp[nm] * vp[nm] * dt2) * ln[nm];
            Lx += c.lap[k] * (((1.0f + bp) * tp - lx0) + ((1.0f + bm) * tm - lx0));
            Dgx += c.grad[k] * ((bp * psix[np] + pr.dbx[cp] * tp) - (bm * psix[nm] + pr.dbx[cm] * tm));
            Dqx += c.grad[k] * (tp - tm);
        }
        {
            const int64_t np = i + k * nx, nm = i - k * nx;
            const int cp = iz + k, cm = iz - k;
            const float bp = pr.bz[cp], bm = pr.bz[cm];
            const float tp = bp * zetaz[np] + (1.0f + bp) * (vp[np] * vp[np] * dt2) * ln[np];
            const float tm = bm * zetaz[nm] + (1.0f + bm) * (vp[nm] * vp[nm] * dt2) * ln[nm];
            Lz += c.lap[k] * (((1.0f + bp) * tp - lz0) + ((1.0f + bm) * tm - lz0));
            Dgz += c.grad[k] * ((bp * psiz[np] + pr.dbz[cp] * tp) - (bm * psiz[nm] + pr.dbz[cm] * tm));
            Dqz += c.grad[k] * (tp - tm);
        }
    }
    Lx *= g.inv_dx2.inv;
    Dgx *= g.inv_dx.inv;
    Dqx *= g.inv_dx.inv;
    Lz *= g.inv_dz2.inv;
    Dgz *= g.inv_dz.inv;
    Dqz *= g.inv_dz.inv;
    f.l_next[off + i] = 2.0f * gun - f.l_prev[off + i] + (Lx + Lz - Dgx - Dgz);
    f.psixo[off + i] = ax_ * (psix[i] - Dqx);
    f.psizo[off + i] = az_ * (psiz[i] - Dqz);
    f.zetaxo[off + i] = ax_ * (zetax[i] + gw);
    f.zetazo[off + i] = az_ * (zetaz[i] + gw);
}

// ---------------------------------------------------------------- BS recon
// Reverse leapfrog without PML inside the restore band, fused u_tt imaging on
// the cells the strip restore will not overwrite (kernels.cuh:472-574).
template <int MS>
__global__ void __launch_bounds__(256) recon_kernel(Geo g, Coef c, const float *__restrict__ vpb,
                                                    const float *__restrict__ u_prev,
                                                    const float *__restrict__ u_now,
                                                    float *__restrict__ u_next,
                                                    const float *__restrict__ adj,
                                                    float *__restrict__ grad, int x_lo, int x_hi,
                                                    int z_lo, int z_hi)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = x_lo + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = z_lo + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= x_hi || iz >= z_hi) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = int64_t(iz) * g.px + ix;
    const float *__restrict__ un = u_now + off;
    const float *__restrict__ vp = vpb + int64_t(b) * g.mstride;
    const float w_sum = lap_fwd<MS>(c, un, i, 1, g.inv_dx2, M) + lap_fwd<MS>(c, un, i, g.px, g.inv_dz2, M);
    const float v = vp[i];
    const float u0 = un[i];
    const float up = u_prev[off + i];
    const float unx = 2.0f * u0 - up + (v * v) * g.dt * g.dt * w_sum;
    u_next[off + i] = unx;
    if (grad && ix >= g.X0 + g.W && ix < g.X1 - g.W && iz >= g.Z0 + g.W && iz < g.Z1 - g.W) {
        const float u_tt = fd::scaled(unx - 2 * u0 + up, g.inv_dt2);
        grad[off + i] += 2.f * g.dt * g.dt * u_tt * adj[off + i] / v;
    }
}

// Strip restore of w^{it-1} into u_next, the forward source of step it taken
// back out where it falls in a strip (common.cu:252-282), and the u_tt imaging
// of the strip cells (kernels.cu:83-131).  One thread per strip cell, each
// cell once: top and bottom rows over the box width, then the side columns
// over the inner rows (the band kernel's enumeration).
__global__ void __launch_bounds__(256) strip_restore_kernel(
    Geo g, const float *__restrict__ slot, float *__restrict__ u_next,
    const float *__restrict__ u_now, const float *__restrict__ u_prev,
    const float *__restrict__ adj, const float *__restrict__ vpb, float *__restrict__ grad,
    const float *__restrict__ src, const int32_t *__restrict__ src_loc, int nsrc, int it,
    int wzl, int wzh, int wxl, int wxh)
{
    const int b = blockIdx.y;
    int t = int(blockIdx.x) * 256 + int(threadIdx.x);
    const int bw = g.X1 - g.X0;
    const int zi0 = g.Z0 + wzl, zi1 = g.Z1 - wzh;
    const int bh = zi1 - zi0;
    const int n_top = wzl * bw, n_bot = wzh * bw, n_left = wxl * bh, n_right = wxh * bh;
    int ix, iz;
    if (t < n_top) { iz = g.Z0 + t / bw; ix = g.X0 + t % bw; }
    else if ((t -= n_top) < n_bot) { iz = zi1 + t / bw; ix = g.X0 + t % bw; }
    else if ((t -= n_bot) < n_left) { ix = g.X0 + t % wxl; iz = zi0 + t / wxl; }
    else if ((t -= n_left) < n_right) { ix = (g.X1 - wxh) + t % wxh; iz = zi0 + t / wxh; }
    else return;
    // saved value: from the top/bottom band when the cell is in one, else a side band
    const float *s = strip_at(slot, int64_t(b) * g.ns, g.sbf16);
    int64_t k;
    if (iz < g.Z0 + g.W) k = int64_t(iz - g.Z0) * g.nxp + (ix - g.X0);
    else if (iz >= g.Z1 - g.W) k = int64_t(g.W) * g.nxp + int64_t(iz - (g.Z1 - g.W)) * g.nxp + (ix - g.X0);
    else if (ix < g.X0 + g.W) k = 2LL * g.W * g.nxp + int64_t(iz - g.Z0) * g.W + (ix - g.X0);
    else k = 2LL * g.W * g.nxp + int64_t(g.nzp) * g.W + int64_t(iz - g.Z0) * g.W + (ix - (g.X1 - g.W));
    float val = strip_get(s, k, g.sbf16);
    for (int k = 0; k < nsrc; ++k) {
        const int32_t *l = src_loc + (int64_t(b) * nsrc + k) * 2;
        if (l[0] == ix && l[1] == iz) val += -src[(int64_t(b) * nsrc + k) * g.nt + it];
    }
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = int64_t(iz) * g.px + ix;
    u_next[off + i] = val;
    if (grad) {
        const float u_tt = fd::scaled(val - 2 * u_now[off + i] + u_prev[off + i], g.inv_dt2);
        grad[off + i] += 2.f * g.dt * g.dt * u_tt * adj[off + i] / vpb[int64_t(b) * g.mstride + i];
    }
}

// grad += 2 dt^2 img lambda / v over the physical box (kernels.cu:3-36);
// optional illumination S += u_tt^2 (img is u_tt, or formed from three
// levels when fnow != nullptr), R += lambda^2 (kernels.cu:133-189).
__global__ void __launch_bounds__(256) image_kernel(Geo g, const float *__restrict__ img,
                                                    const float *__restrict__ fnow,
                                                    const float *__restrict__ fprev,
                                                    const float *__restrict__ lam,
                                                    const float *__restrict__ vpb,
                                                    float *__restrict__ grad, float *__restrict__ S,
                                                    float *__restrict__ R)
{
    const int ix = g.X0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = g.Z0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.X1 || iz >= g.Z1) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = int64_t(iz) * g.px + ix;
    const float l = lam[off + i];
    if (grad) grad[off + i] += 2.f * g.dt * g.dt * img[off + i] * l / vpb[int64_t(b) * g.mstride + i];
    if (R) R[off + i] += l * l;
    if (S) {
        float u_tt;
        if (!fnow) u_tt = img[off + i];
        else u_tt = fd::scaled(fnow[off + i] - 2 * fprev[off + i] + img[off + i], g.inv_dt2);
        S[off + i] += u_tt * u_tt;
    }
}

// Zero the absorbing rim (width abcn + M) of a reconstruction seed
// (common.cu:362-400).
__global__ void __launch_bounds__(256) zero_rim_kernel(Geo g, float *__restrict__ u)
{
    const int ix = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx || iz >= g.nz) return;
    const int w = g.abcn + g.M;
    if (ix < w || ix >= g.nx - w || iz < w || iz >= g.nz - w)
        u[int64_t(b) * g.n + int64_t(iz) * g.px + ix] = 0.f;
}

__global__ void __launch_bounds__(256) copy_kernel(float *__restrict__ dst,
                                                   const float *__restrict__ src, int64_t n)
{
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n; k += int64_t(gridDim.x) * 256)
        dst[k] = src[k];
}

// ---------------------------------------------------------------- launches
static void copy(float *dst, const float *src, int64_t n)
{
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(copy_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), dst, src, n);
    SWP_LAUNCHED();
}

static dim3 grid_for(int nxc, int nzc, int B)
{
    return dim3(unsigned((nxc + 63) / 64), unsigned((nzc + 3) / 4), unsigned(B));
}

template <class K>
static void by_order(int M, K &&k)
{
    switch (M) {
    case 1: k(std::integral_constant<int, 1>{}); break;
    case 2: k(std::integral_constant<int, 2>{}); break;
    case 3: k(std::integral_constant<int, 3>{}); break;
    case 4: k(std::integral_constant<int, 4>{}); break;
    default: k(std::integral_constant<int, 0>{}); break;
    }
}

void launch_fwd(const Ctx &x, const StepFields &f, float *u_this, float *strip)
{
    TimedScope ts(K_FWD);
    if (x.zm) {
        launch_fwd_zm(x, x.plan, f, u_this, strip);
        return;
    }
    const Geo &g = x.g;
    const dim3 grid = grid_for(g.nx - 2 * g.M, g.nz - 2 * g.M, g.B);
    census_add("a2d.fwd_cell", int64_t(grid.x) * grid.y * grid.z);
    by_order(g.M, [&](auto ms) {
        hipLaunchKernelGGL(fwd_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c,
                           x.pr, f, u_this, strip);
    });
    SWP_LAUNCHED();
}

void launch_adj(const Ctx &x, const AdjFields &f, const float *img, float *grad)
{
    TimedScope ts(K_ADJ);
    if (x.zm) {
        launch_adj_zm(x, x.plan, f, img, grad);
        return;
    }
    const Geo &g = x.g;
    const dim3 grid = grid_for(g.nx - 2 * g.M, g.nz - 2 * g.M, g.B);
    census_add("a2d.adj_cell", int64_t(grid.x) * grid.y * grid.z);
    by_order(g.M, [&](auto ms) {
        hipLaunchKernelGGL(adj_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c,
                           x.pr, f, img, grad);
    });
    SWP_LAUNCHED();
}

// ---------------------------------------------------------------- drivers
struct FwdState {
    float *u[3];
    float *psix, *psixn, *psiz, *psizn, *zetax, *zetaz;
};

static void carve_fwd(Carver &cv, const Geo &g, FwdState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto &p : s.u) p = cv.take<float>(N);
    s.psix = cv.take<float>(N);
    s.psixn = cv.take<float>(N);
    s.psiz = cv.take<float>(N);
    s.psizn = cv.take<float>(N);
    s.zetax = cv.take<float>(N);
    s.zetaz = cv.take<float>(N);
}

static void zero_fwd(const Geo &g, FwdState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto p : s.u) fill_zero(p, N);
    fill_zero(s.psix, N); fill_zero(s.psixn, N); fill_zero(s.psiz, N); fill_zero(s.psizn, N);
    fill_zero(s.zetax, N); fill_zero(s.zetaz, N);
}

struct AdjState {
    float *l[3];
    float *psix, *psixn, *psiz, *psizn, *zetax, *zetaxn, *zetaz, *zetazn;
};

static void carve_adj(Carver &cv, const Geo &g, AdjState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto &p : s.l) p = cv.take<float>(N);
    s.psix = cv.take<float>(N); s.psixn = cv.take<float>(N);
    s.psiz = cv.take<float>(N); s.psizn = cv.take<float>(N);
    s.zetax = cv.take<float>(N); s.zetaxn = cv.take<float>(N);
    s.zetaz = cv.take<float>(N); s.zetazn = cv.take<float>(N);
}

static void zero_adj(const Geo &g, AdjState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto p : s.l) fill_zero(p, N);
    fill_zero(s.psix, N); fill_zero(s.psixn, N); fill_zero(s.psiz, N); fill_zero(s.psizn, N);
    fill_zero(s.zetax, N); fill_zero(s.zetaxn, N); fill_zero(s.zetaz, N); fill_zero(s.zetazn, N);
}

// Tape layout per memory mode.
struct Tape {
    float *u_this = nullptr;    // FULL: (nt, B, n)
    float *strips = nullptr;    // BOUNDARY: (nt, B, ns)
    float *last_two = nullptr;  // BOUNDARY: (2, B, n) = (u_prev, u_now) after the last step
    float *snaps = nullptr;     // CKPT: (nchunk, 6, B, n); BINOMIAL: (slots, 6, B, n)
    int nchunk = 0;
};

static Tape carve_tape(Carver &cv, const swp_problem &p, const Geo &g)
{
    Tape t;
    const int64_t N = int64_t(g.B) * g.n;
    if (p.memory == SWP_MEM_FULL) {
        t.u_this = cv.take<float>(int64_t(g.nt) * N);
    } else if (p.memory == SWP_MEM_BOUNDARY) {
        const int64_t cells = int64_t(g.nt) * g.B * g.ns;
        t.strips = g.sbf16 ? reinterpret_cast<float *>(cv.take<uint16_t>(cells)) : cv.take<float>(cells);
        t.last_two = cv.take<float>(2 * N);
    } else if (p.memory == SWP_MEM_CKPT) {
        SWP_CHECK(p.ckpt_interval >= 1, "ckpt_interval must be >= 1");
        t.nchunk = (g.nt + p.ckpt_interval - 1) / p.ckpt_interval;
        t.snaps = cv.take<float>(int64_t(t.nchunk) * 6 * N);
    } else if (p.memory == SWP_MEM_BINOMIAL) {
        SWP_CHECK(p.ckpt_interval >= 1, "binomial checkpointing needs >= 1 snapshot slot");
        t.nchunk = p.ckpt_interval;
        t.snaps = cv.take<float>(int64_t(t.nchunk) * 6 * N);
    }
    return t;
}

// The core's row-pitched copies of the caller's model and of the gradient
// and illumination outputs (the caller's arrays are (B|1, nz, nx) dense).
struct IoBufs {
    float *vp = nullptr;                           // (B or 1, n)
    float *grad = nullptr, *S = nullptr, *R = nullptr;   // (B, n), backward only
};

static IoBufs carve_io(Carver &cv, const swp_problem &p, const Geo &g, int kind)
{
    IoBufs b;
    b.vp = cv.take<float>((p.model_batched ? int64_t(g.B) : 1) * g.n);
    if (kind == 1) {
        const int64_t N = int64_t(g.B) * g.n;
        b.grad = cv.take<float>(N);
        b.S = cv.take<float>(N);
        b.R = cv.take<float>(N);
    }
    return b;
}

// q = p with the model pointer on the core's pitched copy (filled here).
static swp_problem with_pitched_model(const swp_problem &p, const Geo &g, const IoBufs &b)
{
    swp_problem q = p;
    pitch_copy(b.vp, g.px, p.models[0], g.nx, (p.model_batched ? int64_t(g.B) : 1) * g.nz, g.nx);
    q.models[0] = b.vp;
    return q;
}

static PointArgs point_args(const Geo &g)
{
    PointArgs a{};
    a.nt = g.nt;
    a.ndim = 2;
    a.nz = g.nz;
    a.ny = 1;
    a.nx = g.nx;
    a.px = g.px;
    a.n = g.n;
    a.sign = 1.0f;
    return a;
}

static void snapshot(const Geo &g, const FwdState &s, float *dst)
{
    const int64_t N = int64_t(g.B) * g.n;
    const float *src[6] = {s.u[0], s.u[1], s.psix, s.psiz, s.zetax, s.zetaz};
    for (int k = 0; k < 6; ++k) copy(dst + k * N, src[k], N);
}

static void restore_snapshot(const Geo &g, FwdState &s, const float *src)
{
    const int64_t N = int64_t(g.B) * g.n;
    float *dst[6] = {s.u[0], s.u[1], s.psix, s.psiz, s.zetax, s.zetaz};
    for (int k = 0; k < 6; ++k) copy(dst[k], src + k * N, N);
}

// One forward step on the state s (u triple rotated, psi swapped afterwards).
static void forward_step(const Ctx &x, const swp_problem &p, FwdState &s, int it, float *u_this,
                         float *strip, float *record)
{
    const Geo &g = x.g;
    StepFields f{p.models[0], s.u[0], s.u[1], s.u[2], s.psix, s.psiz, s.psixn, s.psizn, s.zetax, s.zetaz};
    launch_fwd(x, f, u_this, strip);
    PointArgs a = point_args(g);
    a.it = it;
    a.u_add = s.u[2];
    a.add_w = p.source;
    a.add_loc = p.sources_loc;
    a.nadd = p.nsrc;
    if (record) {
        a.u_get = s.u[2];
        a.get_out = record;
        a.get_loc = p.receivers_loc;
        a.nget = p.nrec;
    }
    {
        TimedScope ts(K_OTHER);
        launch_points(a, g.B);
    }
    float *t = s.u[0];
    s.u[0] = s.u[1];
    s.u[1] = s.u[2];
    s.u[2] = t;
    std::swap(s.psix, s.psixn);
    std::swap(s.psiz, s.psizn);
}

int64_t workspace_bytes(const swp_problem &p, int kind)
{
    Geo g = make_geo(p);
    Carver cv(nullptr, 0);
    FwdState fs;
    carve_fwd(cv, g, fs);
    if (kind == 1) {
        AdjState as;
        carve_adj(cv, g, as);
    }
    carve_io(cv, p, g, kind);
    if (kind == 1 && p.memory == SWP_MEM_CKPT) cv.take<float>(int64_t(p.ckpt_interval) * g.B * g.n);
    if (kind == 1 && p.memory == SWP_MEM_BINOMIAL) cv.take<float>(int64_t(g.B) * g.n);
    return cv.used + 256;
}

int64_t tape_bytes(const swp_problem &p)
{
    Geo g = make_geo(p);
    Carver cv(nullptr, 0);
    carve_tape(cv, p, g);
    return cv.used + 256;
}

void forward(const Ctx &x, const swp_problem &p_in, float *record)
{
    const Geo &g = x.g;
    Carver ws(p_in.workspace, p_in.workspace_bytes);
    FwdState s;
    carve_fwd(ws, g, s);
    zero_fwd(g, s);
    const swp_problem p = with_pitched_model(p_in, g, carve_io(ws, p_in, g, 0));
    Carver tc(p.tape, p.tape_bytes);
    Tape t = carve_tape(tc, p, g);
    const int64_t N = int64_t(g.B) * g.n;
    std::vector<SpinePt> spine;
    if (p.memory == SWP_MEM_BINOMIAL) spine = revolve_spine(g.nt, t.nchunk);
    size_t sp = 0;
    for (int it = 0; it < g.nt; ++it) {
        if (p.memory == SWP_MEM_CKPT && it % p.ckpt_interval == 0)
            snapshot(g, s, t.snaps + int64_t(it / p.ckpt_interval) * 6 * N);
        for (; sp < spine.size() && spine[sp].step == it; ++sp) snapshot(g, s, t.snaps + int64_t(spine[sp].slot) * 6 * N);
        forward_step(x, p, s, it, t.u_this ? t.u_this + int64_t(it) * N : nullptr,
                     t.strips ? strip_at(t.strips, int64_t(it) * g.B * g.ns, g.sbf16) : nullptr, record);
    }
    if (t.last_two) {
        copy(t.last_two, s.u[0], N);
        copy(t.last_two + N, s.u[1], N);
    }
}

// One adjoint step: stencil, residual injection at the receivers, rotation,
// then the wavelet-gradient gather at the sources (eq_driver.cuh:689-698).
static void adjoint_step(const Ctx &x, const swp_problem &p, const swp_backward_io &io,
                         AdjState &a, int it, const float *img, float *grad)
{
    const Geo &g = x.g;
    AdjFields f{p.models[0], a.l[0], a.l[1], a.l[2], a.psix, a.psiz, a.zetax, a.zetaz,
                a.psixn, a.psizn, a.zetaxn, a.zetazn};
    launch_adj(x, f, img, grad);
    PointArgs pa = point_args(g);
    pa.it = it;
    pa.u_add = a.l[2];
    pa.add_w = io.adjoint_source;
    pa.add_loc = p.receivers_loc;
    pa.nadd = p.nrec;
    if (io.grad_source) {
        pa.u_get = a.l[2];
        pa.get_out = io.grad_source;
        pa.get_loc = p.sources_loc;
        pa.nget = p.nsrc;
        pa.get_accumulate = 1;
    }
    {
        TimedScope ts(K_OTHER);
        launch_points(pa, g.B);
    }
    float *t = a.l[0];
    a.l[0] = a.l[1];
    a.l[1] = a.l[2];
    a.l[2] = t;
    std::swap(a.psix, a.psixn);
    std::swap(a.psiz, a.psizn);
    std::swap(a.zetax, a.zetaxn);
    std::swap(a.zetaz, a.zetazn);
}

static void image(const Ctx &x, const swp_problem &p, const float *img, const float *fnow,
                  const float *fprev, const float *lam, float *grad, float *S, float *R)
{
    const Geo &g = x.g;
    if (!grad && !S && !R) return;
    TimedScope ts(K_OTHER);
    hipLaunchKernelGGL(image_kernel, grid_for(g.X1 - g.X0, g.Z1 - g.Z0, g.B), dim3(64, 4), 0,
                       stream(), g, img, fnow, fprev, lam, p.models[0], grad, S, R);
    SWP_LAUNCHED();
}

static void backward_run(const Ctx &x, const swp_problem &p, const swp_backward_io &io, Carver &ws,
                         FwdState &fs, AdjState &a);

void backward(const Ctx &x, const swp_problem &p_in, const swp_backward_io &io_in)
{
    const Geo &g = x.g;
    Carver ws(p_in.workspace, p_in.workspace_bytes);
    FwdState fs;
    carve_fwd(ws, g, fs);
    AdjState a;
    carve_adj(ws, g, a);
    zero_adj(g, a);
    const IoBufs b = carve_io(ws, p_in, g, 1);
    const swp_problem p = with_pitched_model(p_in, g, b);
    // gradients and illumination accumulate in the core's pitched buffers,
    // then add into the caller's
    swp_backward_io io = io_in;
    const int64_t N = int64_t(g.B) * g.n, rows = int64_t(g.B) * g.nz;
    const bool illum = (p.flags & 1) != 0;
    if (io_in.grad_models[0]) { fill_zero(b.grad, N); io.grad_models[0] = b.grad; }
    if (illum && io_in.illum_source) { fill_zero(b.S, N); io.illum_source = b.S; }
    if (illum && io_in.illum_receiver) { fill_zero(b.R, N); io.illum_receiver = b.R; }
    backward_run(x, p, io, ws, fs, a);
    if (io_in.grad_models[0]) pitch_add(io_in.grad_models[0], g.nx, b.grad, g.px, rows, g.nx);
    if (illum && io_in.illum_source) pitch_add(io_in.illum_source, g.nx, b.S, g.px, rows, g.nx);
    if (illum && io_in.illum_receiver) pitch_add(io_in.illum_receiver, g.nx, b.R, g.px, rows, g.nx);
}

static void backward_run(const Ctx &x, const swp_problem &p, const swp_backward_io &io, Carver &ws,
                         FwdState &fs, AdjState &a)
{
    const Geo &g = x.g;
    const int64_t N = int64_t(g.B) * g.n;
    Carver tc(p.tape, p.tape_bytes);
    Tape t = carve_tape(tc, p, g);
    float *grad = io.grad_models[0];
    const bool illum = (p.flags & 1) != 0;
    float *S = illum ? io.illum_source : nullptr;
    float *R = illum ? io.illum_receiver : nullptr;

    if (p.memory == SWP_MEM_FULL) {
        for (int it = g.nt - 1; it >= 0; --it) {
            const float *img = (grad && it + 1 < g.nt) ? t.u_this + int64_t(it + 1) * N : nullptr;
            adjoint_step(x, p, io, a, it, img, img ? grad : nullptr);
            if (S || R) image(x, p, t.u_this + int64_t(it) * N, nullptr, nullptr, a.l[1], nullptr, S, R);
        }
        if (grad && g.nt > 0) image(x, p, t.u_this, nullptr, nullptr, a.l[1], grad, nullptr, nullptr);
        return;
    }

    if (p.memory == SWP_MEM_BINOMIAL) {
        // the forward left the spine snapshots; the schedule re-runs steps
        // from them and reverses each step right after re-running it
        float *cap = ws.take<float>(N);
        zero_fwd(g, fs);   // the cells no step writes (rim, psi outside the bands) stay zero
        for (const RevOp &o : revolve_schedule(g.nt, t.nchunk)) {
            switch (o.code) {
            case REV_STEP: forward_step(x, p, fs, o.arg, nullptr, nullptr, nullptr); break;
            case REV_SAVE: snapshot(g, fs, t.snaps + int64_t(o.arg) * 6 * N); break;
            case REV_LOAD: restore_snapshot(g, fs, t.snaps + int64_t(o.arg) * 6 * N); break;
            default:
                forward_step(x, p, fs, o.arg, cap, nullptr, nullptr);
                adjoint_step(x, p, io, a, o.arg, nullptr, nullptr);
                image(x, p, cap, nullptr, nullptr, a.l[1], grad, S, R);
                break;
            }
        }
        return;
    }

    if (p.memory == SWP_MEM_CKPT) {
        const int I = p.ckpt_interval;
        float *chunk = ws.take<float>(int64_t(I) * N);
        for (int c = t.nchunk - 1; c >= 0; --c) {
            const int it0 = c * I, it1 = std::min(g.nt, it0 + I);
            zero_fwd(g, fs);
            restore_snapshot(g, fs, t.snaps + int64_t(c) * 6 * N);
            for (int it = it0; it < it1; ++it)
                forward_step(x, p, fs, it, chunk + int64_t(it - it0) * N, nullptr, nullptr);
            for (int it = it1 - 1; it >= it0; --it) {
                adjoint_step(x, p, io, a, it, nullptr, nullptr);
                image(x, p, chunk + int64_t(it - it0) * N, nullptr, nullptr, a.l[1], grad, S, R);
            }
        }
        return;
    }

    SWP_CHECK(p.memory == SWP_MEM_BOUNDARY, "acoustic2d backward: unknown memory mode ", p.memory);
    // Reconstruction seed: (prev, now) = (w^nt, w^(nt-1)) with the absorbing rim zeroed.
    float *r[3] = {fs.u[0], fs.u[1], fs.u[2]};
    fill_zero(r[2], N);
    copy(r[0], t.last_two + N, N);
    copy(r[1], t.last_two, N);
    {
        const dim3 grid = grid_for(g.nx, g.nz, g.B);
        hipLaunchKernelGGL(zero_rim_kernel, grid, dim3(64, 4), 0, stream(), g, r[0]);
        hipLaunchKernelGGL(zero_rim_kernel, grid, dim3(64, 4), 0, stream(), g, r[1]);
        SWP_LAUNCHED();
    }
    const int bxl = g.pxl > 0 ? g.pxl + 2 * g.M + 1 : 2 * g.M;
    const int bxh = g.pxh > 0 ? g.pxh + 2 * g.M + 1 : 2 * g.M;
    const int bzl = g.pzl > 0 ? g.pzl + 2 * g.M + 1 : 2 * g.M;
    const int bzh = g.pzh > 0 ? g.pzh + 2 * g.M + 1 : 2 * g.M;
    int wzl = g.W, wzh = g.W, wxl = g.W, wxh = g.W;
    {   // bs_strip_widths (strip.cuh:64-68): a thin box images no cell twice
        const int ez = std::max(g.Z1 - g.Z0, 0), ex = std::max(g.X1 - g.X0, 0);
        wzl = std::min(wzl, ez); wzh = std::min(wzh, std::max(ez - wzl, 0));
        wxl = std::min(wxl, ex); wxh = std::min(wxh, std::max(ex - wxl, 0));
    }
    const int bw = g.X1 - g.X0, bh = g.Z1 - g.Z0 - wzl - wzh;
    const int n_strip = (wzl + wzh) * bw + (wxl + wxh) * std::max(bh, 0);
    for (int it = g.nt - 1; it >= 1; --it) {
        adjoint_step(x, p, io, a, it, nullptr, nullptr);
        const float *lam = a.l[1];
        if (bxl < g.nx - bxh && bzl < g.nz - bzh) {
            TimedScope ts(K_RECON);
            if (x.zm) {
                launch_recon_zm(x, p.models[0], r[0], r[1], r[2], lam, grad, bxl, g.nx - bxh, bzl,
                                g.nz - bzh, x.tz_recon);
            } else {
                by_order(g.M, [&](auto ms) {
                    hipLaunchKernelGGL(recon_kernel<decltype(ms)::value>,
                                       grid_for(g.nx - bxh - bxl, g.nz - bzh - bzl, g.B), dim3(64, 4), 0,
                                       stream(), g, x.c, p.models[0], r[0], r[1], r[2], lam, grad, bxl,
                                       g.nx - bxh, bzl, g.nz - bzh);
                });
                SWP_LAUNCHED();
            }
        }
        if (n_strip > 0) {
            TimedScope ts(K_OTHER);
            hipLaunchKernelGGL(strip_restore_kernel, dim3(unsigned((n_strip + 255) / 256), unsigned(g.B)),
                               dim3(256), 0, stream(), g, strip_at(t.strips, int64_t(it - 1) * g.B * g.ns, g.sbf16), r[2],
                               r[1], r[0], lam, p.models[0], grad, p.source, p.sources_loc, p.nsrc, it,
                               wzl, wzh, wxl, wxh);
            SWP_LAUNCHED();
        }
        if (S || R) image(x, p, r[0], r[2], r[1], lam, nullptr, S, R);
        PointArgs pa = point_args(g);
        pa.it = it;
        pa.u_add = r[2];
        pa.add_w = p.source;
        pa.add_loc = p.sources_loc;
        pa.nadd = p.nsrc;
        {
            TimedScope ts(K_OTHER);
            launch_points(pa, g.B);
        }
        float *tp = r[0];
        r[0] = r[1];
        r[1] = r[2];
        r[2] = tp;
    }
    if (g.nt > 0) {
        adjoint_step(x, p, io, a, 0, nullptr, nullptr);
        if (R) image(x, p, nullptr, nullptr, nullptr, a.l[1], nullptr, nullptr, R);
    }
}

Ctx make_ctx(const swp_problem &p)
{
    SWP_CHECK(p.ndim == 2, "acoustic2d: ndim must be 2, got ", p.ndim);
    SWP_CHECK(p.nz > 2 * p.M && p.nx > 2 * p.M, "acoustic2d: grid smaller than the stencil");
    SWP_CHECK(p.B >= 1, "acoustic2d: B must be >= 1");
    SWP_CHECK(p.models[0] != nullptr, "acoustic2d: vp is required");
    for (int k = 0; k < 6; ++k) SWP_CHECK(p.pml[k] != nullptr, "acoustic2d: CPML profile ", k, " missing");
    SWP_CHECK(p.lap_coes && p.grad_coes, "acoustic2d: FD coefficient tables missing");
    Ctx x;
    x.g = make_geo(p);
    SWP_CHECK(x.g.X1 > x.g.X0 && x.g.Z1 > x.g.Z0, "acoustic2d: empty physical box");
    x.c = make_coef(p, p.lap_coes, p.grad_coes);
    x.pr = Prof{p.pml[0], p.pml[1], p.pml[2], p.pml[3], p.pml[4], p.pml[5]};
    // Kernel family: the z-marching kernels from kZmMinCells cells per launch
    // (B x nz x nx on the runtime grid) up, the one-thread-per-cell kernels
    // (any order) below.  A z-march block walks its segment row by row, one
    // load latency per row, so on a small launch the few blocks it makes
    // leave the chip idle: at 96 x 128 and one shot it takes 84 us per step of
    // forward + backward against 24 for the per-cell kernels; the two meet
    // near 4.7M cells (147 against 160 us) and at 384 x 1152 x 32 shots the
    // z-march is 2.8x faster.  SWP_KERNELS=percell / zmarch forces a family.
    // The segment heights of the z-marching kernels come from the tuning
    // table (swp_set_tuning), else SWP_TZ / SWP_TZ_BAND / SWP_TZ_RECON.
    const char *kern = std::getenv("SWP_KERNELS");
    const std::string ks = kern ? kern : "";
    const int64_t cells = int64_t(x.g.B) * x.g.nz * x.g.nx;
    x.zm = zm_supported(x.g) && ks != "percell" && (ks == "zmarch" || cells >= kZmMinCells);
    x.tz_recon = tuning(SWP_TUNE_TZ2_RECON, "SWP_TZ_RECON", 32);
    x.plan = make_plan(x.g, tuning(SWP_TUNE_TZ2, "SWP_TZ", 32), tuning(SWP_TUNE_TZ2_BAND, "SWP_TZ_BAND", 8));
    return x;
}

}  // namespace a2d
}  // namespace swp
