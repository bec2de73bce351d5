// 3-D constant-density acoustic path on gfx950: forward step with the
// product-rule cpmlr CPML, exact discrete adjoint, boundary-saving reverse
// reconstruction with fused u_tt imaging, six-face strip save/restore,
// full-storage and checkpoint imaging, illumination.  FP32 on the vector
// ALUs; wave64 rows along x.
//
// Arithmetic follows the reference's acoustic3d kernels term for term
// (csrc/cuda/equations/acoustic3d/forward_split.cu:81-328,
// adjoint_split.cu:56-318, kernels.cu:3-285); the launch structure, memory
// layout and time loops are this file's own.  Time loop order as the
// reference's driver (acoustic3d/driver_traits.cuh, common/eq_driver.cuh):
// forward step -> strip save -> source -> record -> rotate; adjoint step ->
// residual -> rotate -> wavelet-gradient gather.
#include "acoustic3d.h"
#include "acoustic3d_zm.h"
#include "revolve.h"
#include "fd.h"

#include <algorithm>
#include <cstdlib>
#include <string>

namespace swp {
namespace a3d {

bool available() { return true; }

Geo make_geo(const swp_problem &p)
{
    Geo g{};
    g.nz = p.nz;
    g.ny = p.ny;
    g.nx = p.nx;
    g.B = p.B;
    g.M = p.M;
    g.nt = p.nt;
    g.pzl = p.pad[0];
    g.pzh = p.pad[1];
    g.pyl = p.pad[2];
    g.pyh = p.pad[3];
    g.pxl = p.pad[4];
    g.pxh = p.pad[5];
    g.X0 = g.pxl + g.M;
    g.X1 = g.nx - g.pxh - g.M;
    g.Y0 = g.pyl + g.M;
    g.Y1 = g.ny - g.pyh - g.M;
    g.Z0 = g.pzl + g.M;
    g.Z1 = g.nz - g.pzh - g.M;
    g.abcn = p.abcn;
    g.W = g.M + 1;
    g.nxp = g.X1 - g.X0;
    g.nyp = g.Y1 - g.Y0;
    g.nzp = g.Z1 - g.Z0;
    // rows of every array the core allocates start on 128-byte lines
    g.px = (g.nx + 31) / 32 * 32;
    g.sxy = int64_t(g.px) * g.ny;
    g.n = g.sxy * g.nz;
    g.nm = int64_t(g.nz) * g.ny * g.nx;
    g.mstride = p.model_batched ? g.n : 0;
    const int64_t W = g.W;
    g.ns = 2 * W * g.nyp * g.nxp + 2 * W * g.nzp * g.nxp + 2 * W * g.nzp * g.nyp;
    g.sbf16 = (p.flags & SWP_FLAG_BF16_STRIPS) != 0;
    g.dt = p.dt;
    g.dx = p.dx;
    g.dy = p.dy;
    g.dz = p.dz;
    g.inv_dx2 = fd::scale_of(g.dx * g.dx);
    g.inv_dy2 = fd::scale_of(g.dy * g.dy);
    g.inv_dz2 = fd::scale_of(g.dz * g.dz);
    g.inv_dx = fd::scale_of(g.dx);
    g.inv_dy = fd::scale_of(g.dy);
    g.inv_dz = fd::scale_of(g.dz);
    g.dt2 = g.dt * g.dt;
    g.inv_dt2 = fd::scale_of(g.dt * g.dt);
    return g;
}

// ---------------------------------------------------------------- forward
// One thread per cell of a box inside the computed box [M, n-M)^3 of one
// shot; blocks of 64 x 4 cells in (x, y), one z plane.  Cells inside phys_box(1) take the bare
// leapfrog; the shell applies, on each axis whose CPML band holds the cell
// (pos < lo + 1 or pos >= hi - 1), the product-rule update
//   tmp = ((1 + b) L + db du) + a dpsi + (da) psi,  w += (1 + b) tmp + a zeta,
//   psi' = b du + a psi,  zeta = b tmp + a zeta
// (forward_split.cu:132-249), the plain second derivative on the others.
template <int MS>
__global__ void __launch_bounds__(256) fwd_kernel(Geo g, Coef c, Prof pr, StepFields f,
                                                  float *__restrict__ u_this, float *__restrict__ strip, Box bx)
{
    const int M = MS > 0 ? MS : g.M;
    const int nzb = bx.z1 - bx.z0;
    const int ix = bx.x0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iy = bx.y0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int iz = bx.z0 + int(blockIdx.z) % nzb;
    const int b = int(blockIdx.z) / nzb;
    if (ix >= bx.x1 || iy >= bx.y1) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = (int64_t(iz) * g.ny + iy) * g.px + ix;
    const float *__restrict__ un = f.u_now + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    if (strip) save_strips(g, strip + int64_t(b) * g.ns, ix, iy, iz, un[i]);

    const int64_t S[3] = {1, g.px, g.sxy};
    const int pos[3] = {ix, iy, iz};
    const int lo[3] = {g.X0, g.Y0, g.Z0}, hi[3] = {g.X1, g.Y1, g.Z1};
    const fd::Scale ih2[3] = {g.inv_dx2, g.inv_dy2, g.inv_dz2}, ih[3] = {g.inv_dx, g.inv_dy, g.inv_dz};
    float L[3];
    bool inb[3];
    bool any = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t s = S[a];
        L[a] = fd::lap<MS>(c, [&](int k) { return un[i + k * s]; }, ih2[a], M);
        inb[a] = pos[a] < lo[a] + 1 || pos[a] >= hi[a] - 1;
        any = any || inb[a];
    }
    float w_sum;
    if (!any) {
        w_sum = L[0] + L[1] + L[2];
    } else {
        w_sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!inb[a]) {
                w_sum += L[a];
                continue;
            }
            const int64_t s = S[a];
            const int p = pos[a];
            const float *__restrict__ psi = f.psi[a] + off;
            const float *__restrict__ A = pr.a[a];
            const float du = fd::d1<MS>(c, [&](int k) { return un[i + k * s]; }, ih[a], M);
            const float dpsi = fd::d1<MS>(c, [&](int k) { return psi[i + k * s]; }, ih[a], M);
            const float da = fd::d1<MS>(c, [&](int k) { return A[p + k]; }, ih[a], M);
            const float a_ = A[p], b_ = pr.b[a][p], db_ = pr.db[a][p];
            const float psi0 = psi[i];
            const float tmp = ((1.0f + b_) * L[a] + db_ * du) + a_ * dpsi + da * psi0;
            float *__restrict__ zeta = f.zeta[a] + off;
            const float z0 = zeta[i];
            w_sum += (1.0f + b_) * tmp + a_ * z0;
            f.psin[a][off + i] = b_ * du + a_ * psi0;
            zeta[i] = b_ * tmp + a_ * z0;
        }
    }
    const float v = vp[i];
    f.u_next[off + i] = 2.0f * un[i] - f.u_prev[off + i] + (v * v) * g.dt * g.dt * w_sum;
    if (u_this) u_this[off + i] = (v * v) * w_sum;
}

// ---------------------------------------------------------------- adjoint
// Exact transpose of fwd_kernel (+ source/record).  Cells outside
// phys_box(M + 1) on an axis run that axis' transposed CPML update
// (adjoint_split.cu:56-318) with psi/zeta double-buffered; img != nullptr
// fuses the lagged full-mode imaging grad += 2 dt^2 img lambda / v.
template <int MS>
__global__ void __launch_bounds__(256) adj_kernel(Geo g, Coef c, Prof pr, AdjFields f,
                                                  const float *__restrict__ img, float *__restrict__ grad, Box bx)
{
    const int M = MS > 0 ? MS : g.M;
    const int nzb = bx.z1 - bx.z0;
    const int ix = bx.x0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iy = bx.y0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int iz = bx.z0 + int(blockIdx.z) % nzb;
    const int b = int(blockIdx.z) / nzb;
    if (ix >= bx.x1 || iy >= bx.y1) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = (int64_t(iz) * g.ny + iy) * g.px + ix;
    const float *__restrict__ ln = f.l_now + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const float dt2 = g.dt2;
    const float gun = ln[i];
    if (grad && ix >= g.X0 && ix < g.X1 && iy >= g.Y0 && iy < g.Y1 && iz >= g.Z0 && iz < g.Z1)
        grad[off + i] += 2.f * g.dt * g.dt * img[off + i] * gun / vp[i];
    const float v0 = vp[i];
    const float gw = v0 * v0 * dt2 * gun;
    auto GW = [&](int64_t nn) { return vp[nn] * vp[nn] * dt2 * ln[nn]; };
    const int64_t S[3] = {1, g.px, g.sxy};
    const int pos[3] = {ix, iy, iz};
    const int lo[3] = {g.X0, g.Y0, g.Z0}, hi[3] = {g.X1, g.Y1, g.Z1};
    const fd::Scale ih2[3] = {g.inv_dx2, g.inv_dy2, g.inv_dz2}, ih[3] = {g.inv_dx, g.inv_dy, g.inv_dz};
    bool in[3];
    bool any = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        in[a] = pos[a] < lo[a] + M + 1 || pos[a] >= hi[a] - M - 1;
        any = any || in[a];
    }
    if (!any) {
        float lp[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int64_t s = S[a];
            float acc = 0.f;
            if constexpr (MS > 0) {
#pragma unroll
                for (int k = 1; k <= MS; ++k) acc += c.lap[k] * ((GW(i + k * s) - gw) + (GW(i - k * s) - gw));
            } else {
                for (int k = 1; k <= M; ++k) acc += c.lap[k] * ((GW(i + k * s) - gw) + (GW(i - k * s) - gw));
            }
            lp[a] = acc;
        }
        f.l_next[off + i] = 2.0f * gun - f.l_prev[off + i] + lp[0] * g.inv_dx2.inv + lp[1] * g.inv_dy2.inv + lp[2] * g.inv_dz2.inv;
        return;
    }
    float Lt[3], Dg[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t s = S[a];
        if (!in[a]) {
            float L = 0.f;
            for (int k = 1; k <= M; ++k) L += c.lap[k] * ((GW(i + k * s) - gw) + (GW(i - k * s) - gw));
            Lt[a] = L * ih2[a].inv;
            Dg[a] = 0.f;
            continue;
        }
        const float *__restrict__ Aa = pr.a[a];
        const float *__restrict__ Ba = pr.b[a];
        const float *__restrict__ Da = pr.db[a];
        const int p = pos[a];
        const float a_ = Aa[p], b_ = Ba[p];
        const float da = fd::d1<MS>(c, [&](int k) { return Aa[p + k]; }, ih[a], M);
        const float *__restrict__ zeta = f.zeta[a] + off;
        const float *__restrict__ psi = f.psi[a] + off;
        const float gt0 = b_ * zeta[i] + (1.0f + b_) * gw;
        const float l0 = (1.0f + b_) * gt0;
        float L = 0.f, Dgg = 0.f, Dq = 0.f;
        for (int k = 1; k <= M; ++k) {
            const int64_t np = i + k * s, nm = i - k * s;
            const int cp = p + k, cm = p - k;
            const float bp = Ba[cp], bm = Ba[cm];
            const float tp = bp * zeta[np] + (1.0f + bp) * (vp[np] * vp[np] * dt2) * ln[np];
            const float tm = bm * zeta[nm] + (1.0f + bm) * (vp[nm] * vp[nm] * dt2) * ln[nm];
            L += c.lap[k] * (((1.0f + bp) * tp - l0) + ((1.0f + bm) * tm - l0));
            Dgg += c.grad[k] * ((bp * psi[np] + Da[cp] * tp) - (bm * psi[nm] + Da[cm] * tm));
            Dq += c.grad[k] * (Aa[cp] * tp - Aa[cm] * tm);
        }
        Lt[a] = L * ih2[a].inv;
        Dg[a] = Dgg * ih[a].inv;
        Dq *= ih[a].inv;
        f.psio[a][off + i] = a_ * psi[i] + da * gt0 - Dq;
        f.zetao[a][off + i] = a_ * (zeta[i] + gw);
    }
    f.l_next[off + i] = 2.0f * gun - f.l_prev[off + i] + (Lt[0] + Lt[1] + Lt[2] - Dg[0] - Dg[1] - Dg[2]);
}

// ---------------------------------------------------------------- BS recon
// Reverse leapfrog without PML on [lo, hi) per axis, fused u_tt imaging on
// the cells the strip restore will not overwrite (forward_split.cu:256-328).
template <int MS>
__global__ void __launch_bounds__(256) recon_kernel(Geo g, Coef c, const float *__restrict__ vpb,
                                                    const float *__restrict__ u_prev,
                                                    const float *__restrict__ u_now, float *__restrict__ u_next,
                                                    const float *__restrict__ adj, float *__restrict__ grad, Box bx)
{
    const int M = MS > 0 ? MS : g.M;
    const int nzb = bx.z1 - bx.z0;
    const int ix = bx.x0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iy = bx.y0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int iz = bx.z0 + int(blockIdx.z) % nzb;
    const int b = int(blockIdx.z) / nzb;
    if (ix >= bx.x1 || iy >= bx.y1) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = (int64_t(iz) * g.ny + iy) * g.px + ix;
    const float *__restrict__ un = u_now + off;
    const float *__restrict__ vp = vpb + int64_t(b) * g.mstride;
    const float w_sum = fd::lap<MS>(c, [&](int k) { return un[i + k]; }, g.inv_dx2, M) +
                        fd::lap<MS>(c, [&](int k) { return un[i + k * int64_t(g.px)]; }, g.inv_dy2, M) +
                        fd::lap<MS>(c, [&](int k) { return un[i + k * g.sxy]; }, g.inv_dz2, M);
    const float v = vp[i];
    const float u0 = un[i];
    const float up = u_prev[off + i];
    const float unx = 2.0f * u0 - up + (v * v) * g.dt * g.dt * w_sum;
    u_next[off + i] = unx;
    if (grad && ix >= g.X0 + g.W && ix < g.X1 - g.W && iy >= g.Y0 + g.W && iy < g.Y1 - g.W &&
        iz >= g.Z0 + g.W && iz < g.Z1 - g.W) {
        const float u_tt = fd::scaled(unx - 2 * u0 + up, g.inv_dt2);
        grad[off + i] += 2.f * g.dt * g.dt * u_tt * adj[off + i] / v;
    }
}

// Strip restore of w^{it-1} into u_next, the forward source of step it taken
// back out where it falls in a strip (common.cu:252-282), and the u_tt imaging
// of the strip cells (kernels.cu:94-148).  One thread per strip cell, each
// cell once: the z slabs over the whole box section, then the y slabs over
// the inner z range, then the x slabs over the inner z and y ranges.
__global__ void __launch_bounds__(256) strip_restore_kernel(
    Geo g, const float *__restrict__ slot, float *__restrict__ u_next, const float *__restrict__ u_now,
    const float *__restrict__ u_prev, const float *__restrict__ adj, const float *__restrict__ vpb,
    float *__restrict__ grad, const float *__restrict__ src, const int32_t *__restrict__ src_loc, int nsrc,
    int it, StripW w)
{
    const int b = blockIdx.y;
    int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t nxp = g.nxp, nyp = g.nyp;
    const int zi0 = g.Z0 + w.zl, zi1 = g.Z1 - w.zh, yi0 = g.Y0 + w.yl, yi1 = g.Y1 - w.yh;
    const int64_t mz = zi1 - zi0, my = yi1 - yi0;
    const int64_t nA = int64_t(w.zl + w.zh) * nyp * nxp;
    const int64_t nC = mz * (w.yl + w.yh) * nxp;
    const int64_t nE = mz * my * (w.xl + w.xh);
    int ix, iy, iz;
    if (t < nA) {
        const int64_t pl = t / (nyp * nxp), r = t % (nyp * nxp);
        iz = pl < w.zl ? g.Z0 + int(pl) : g.Z1 - w.zh + int(pl - w.zl);
        iy = g.Y0 + int(r / nxp);
        ix = g.X0 + int(r % nxp);
    } else if ((t -= nA) < nC) {
        const int64_t rows = w.yl + w.yh;
        const int64_t pl = t / (rows * nxp), r = t % (rows * nxp);
        iz = zi0 + int(pl);
        const int row = int(r / nxp);
        iy = row < w.yl ? g.Y0 + row : g.Y1 - w.yh + (row - w.yl);
        ix = g.X0 + int(r % nxp);
    } else if ((t -= nC) < nE) {
        const int64_t cols = w.xl + w.xh;
        const int64_t pl = t / (my * cols), r = t % (my * cols);
        iz = zi0 + int(pl);
        iy = yi0 + int(r / cols);
        const int col = int(r % cols);
        ix = col < w.xl ? g.X0 + col : g.X1 - w.xh + (col - w.xl);
    } else {
        return;
    }
    float val = load_strip(g, strip_at(slot, int64_t(b) * g.ns, g.sbf16), ix, iy, iz);
    for (int k = 0; k < nsrc; ++k) {
        const int32_t *l = src_loc + (int64_t(b) * nsrc + k) * 3;
        if (l[0] == ix && l[1] == iy && l[2] == iz) val += -src[(int64_t(b) * nsrc + k) * g.nt + it];
    }
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = (int64_t(iz) * g.ny + iy) * g.px + ix;
    u_next[off + i] = val;
    if (grad) {
        const float u_tt = fd::scaled(val - 2 * u_now[off + i] + u_prev[off + i], g.inv_dt2);
        grad[off + i] += 2.f * g.dt * g.dt * u_tt * adj[off + i] / vpb[int64_t(b) * g.mstride + i];
    }
}

// grad += 2 dt^2 img lambda / v over the physical box (kernels.cu:3-42);
// optional illumination S += u_tt^2 (img is u_tt, or formed from three
// levels when fnow != nullptr), R += lambda^2 (kernels.cu:150-209).
__global__ void __launch_bounds__(256) image_kernel(Geo g, const float *__restrict__ img,
                                                    const float *__restrict__ fnow, const float *__restrict__ fprev,
                                                    const float *__restrict__ lam, const float *__restrict__ vpb,
                                                    float *__restrict__ grad, float *__restrict__ S,
                                                    float *__restrict__ R)
{
    const int ix = g.X0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iy = g.Y0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int iz = g.Z0 + int(blockIdx.z) % g.nzp;
    const int b = int(blockIdx.z) / g.nzp;
    if (ix >= g.X1 || iy >= g.Y1) return;
    const int64_t off = int64_t(b) * g.n;
    const int64_t i = (int64_t(iz) * g.ny + iy) * g.px + ix;
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

__global__ void __launch_bounds__(256) copy_kernel(float *__restrict__ dst, const float *__restrict__ src, int64_t n)
{
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n; k += int64_t(gridDim.x) * 256) dst[k] = src[k];
}

// ---------------------------------------------------------------- launches
static void copy(float *dst, const float *src, int64_t n)
{
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(copy_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), dst, src, n);
    SWP_LAUNCHED();
}

// Per-cell launch over an (nxc, nyc, nzc) box of each of B shots.
static dim3 grid_for(int nxc, int nyc, int nzc, int B)
{
    const int64_t gz = int64_t(nzc) * B;
    SWP_CHECK(gz <= 65535, "acoustic3d: nz * B = ", gz, " exceeds the launch limit 65535");
    return dim3(unsigned((nxc + 63) / 64), unsigned((nyc + 3) / 4), unsigned(gz));
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

static Box computed_box(const Geo &g) { return {g.M, g.nx - g.M, g.M, g.ny - g.M, g.M, g.nz - g.M}; }

template <class K>
static void per_box(const Geo &g, bool shell_only, K &&k)
{
    Box boxes[6];
    int n = 1;
    if (shell_only) n = shell_boxes(g, boxes);
    else boxes[0] = computed_box(g);
    for (int j = 0; j < n; ++j) {
        const Box &b = boxes[j];
        k(b, grid_for(b.x1 - b.x0, b.y1 - b.y0, b.z1 - b.z0, g.B));
    }
}

static void launch_fwd(const Ctx &x, const StepFields &f, float *u_this, float *strip)
{
    TimedScope ts(K_FWD);
    const Geo &g = x.g;
    if (x.zm) launch_fwd_zm(x, f, u_this, strip, !x.shell_percell);
    if (x.zm && !x.shell_percell) return;
    by_order(g.M, [&](auto ms) {
        per_box(g, x.zm, [&](const Box &bx, dim3 grid) {
            census_add("a3d.fwd_cell", int64_t(grid.x) * grid.y * grid.z);
            hipLaunchKernelGGL(fwd_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c, x.pr, f,
                               u_this, strip, bx);
        });
    });
    SWP_LAUNCHED();
}

static void launch_adj(const Ctx &x, const AdjFields &f, const float *img, float *grad)
{
    TimedScope ts(K_ADJ);
    const Geo &g = x.g;
    if (x.zm) launch_adj_zm(x, f, img, grad, !x.shell_percell);
    if (x.zm && !x.shell_percell) return;
    by_order(g.M, [&](auto ms) {
        per_box(g, x.zm, [&](const Box &bx, dim3 grid) {
            census_add("a3d.adj_cell", int64_t(grid.x) * grid.y * grid.z);
            hipLaunchKernelGGL(adj_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c, x.pr, f,
                               img, grad, bx);
        });
    });
    SWP_LAUNCHED();
}

static void launch_recon(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                         const float *adj, float *grad, const Box &bx)
{
    TimedScope ts(K_RECON);
    const Geo &g = x.g;
    if (x.zm) {
        launch_recon_zm(x, vp, u_prev, u_now, u_next, adj, grad, bx.x0, bx.x1, bx.y0, bx.y1, bx.z0, bx.z1);
        return;
    }
    const dim3 grid = grid_for(bx.x1 - bx.x0, bx.y1 - bx.y0, bx.z1 - bx.z0, g.B);
    by_order(g.M, [&](auto ms) {
        hipLaunchKernelGGL(recon_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c, vp, u_prev,
                           u_now, u_next, adj, grad, bx);
    });
    SWP_LAUNCHED();
}

// ---------------------------------------------------------------- drivers
struct FwdState {
    float *u[3];
    float *psi[3], *psin[3], *zeta[3];
};

static void carve_fwd(Carver &cv, const Geo &g, FwdState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto &p : s.u) p = cv.take<float>(N);
    for (int a = 0; a < 3; ++a) {
        s.psi[a] = cv.take<float>(N);
        s.psin[a] = cv.take<float>(N);
        s.zeta[a] = cv.take<float>(N);
    }
}

static void zero_fwd(const Geo &g, FwdState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto p : s.u) fill_zero(p, N);
    for (int a = 0; a < 3; ++a) {
        fill_zero(s.psi[a], N);
        fill_zero(s.psin[a], N);
        fill_zero(s.zeta[a], N);
    }
}

struct AdjState {
    float *l[3];
    float *psi[3], *psin[3], *zeta[3], *zetan[3];
};

static void carve_adj(Carver &cv, const Geo &g, AdjState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto &p : s.l) p = cv.take<float>(N);
    for (int a = 0; a < 3; ++a) {
        s.psi[a] = cv.take<float>(N);
        s.psin[a] = cv.take<float>(N);
        s.zeta[a] = cv.take<float>(N);
        s.zetan[a] = cv.take<float>(N);
    }
}

static void zero_adj(const Geo &g, AdjState &s)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (auto p : s.l) fill_zero(p, N);
    for (int a = 0; a < 3; ++a) {
        fill_zero(s.psi[a], N);
        fill_zero(s.psin[a], N);
        fill_zero(s.zeta[a], N);
        fill_zero(s.zetan[a], N);
    }
}

constexpr int kSnapFields = 8;   // u_prev, u_now, psi x/y/z, zeta x/y/z

struct Tape {
    float *u_this = nullptr;    // FULL: (nt, B, n)
    float *strips = nullptr;    // BOUNDARY: (nt, B, ns)
    float *last_two = nullptr;  // BOUNDARY: (2, B, n) = (u_prev, u_now) after the last step
    float *snaps = nullptr;     // CKPT: (nchunk, 8, B, n)
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
        t.snaps = cv.take<float>(int64_t(t.nchunk) * kSnapFields * N);
    } else if (p.memory == SWP_MEM_BINOMIAL) {
        SWP_CHECK(p.ckpt_interval >= 1, "binomial checkpointing needs >= 1 snapshot slot");
        t.nchunk = p.ckpt_interval;
        t.snaps = cv.take<float>(int64_t(t.nchunk) * kSnapFields * N);
    }
    return t;
}

// The core's row-pitched copies of the caller's model and of the gradient
// and illumination outputs (the caller's arrays are (B|1, nz, ny, nx) dense).
struct IoBufs {
    float *vp = nullptr;                                 // (B or 1, n)
    float *grad = nullptr, *S = nullptr, *R = nullptr;   // (B, n), backward only (S, R with illumination)
    float *stage = nullptr;                              // (B, ns) float32 strips of one step, bf16 tape only
};

static IoBufs carve_io(Carver &cv, const swp_problem &p, const Geo &g, int kind)
{
    IoBufs b;
    b.vp = cv.take<float>((p.model_batched ? int64_t(g.B) : 1) * g.n);
    if (kind == 0 && p.memory == SWP_MEM_BOUNDARY && g.sbf16) b.stage = cv.take<float>(int64_t(g.B) * g.ns);
    if (kind == 1) {
        const int64_t N = int64_t(g.B) * g.n;
        b.grad = cv.take<float>(N);
        if (p.flags & 1) {
            b.S = cv.take<float>(N);
            b.R = cv.take<float>(N);
        }
    }
    return b;
}

// q = p with the model pointer on the core's pitched copy (filled here).
static swp_problem with_pitched_model(const swp_problem &p, const Geo &g, const IoBufs &b)
{
    swp_problem q = p;
    pitch_copy(b.vp, g.px, p.models[0], g.nx, (p.model_batched ? int64_t(g.B) : 1) * g.nz * g.ny, g.nx);
    q.models[0] = b.vp;
    return q;
}

static PointArgs point_args(const Geo &g)
{
    PointArgs a{};
    a.nt = g.nt;
    a.ndim = 3;
    a.nz = g.nz;
    a.ny = g.ny;
    a.nx = g.nx;
    a.px = g.px;
    a.n = g.n;
    a.sign = 1.0f;
    return a;
}

static void snap_ptrs(FwdState &s, float *out[kSnapFields])
{
    out[0] = s.u[0];
    out[1] = s.u[1];
    for (int a = 0; a < 3; ++a) {
        out[2 + a] = s.psi[a];
        out[5 + a] = s.zeta[a];
    }
}

static void snapshot(const Geo &g, FwdState &s, float *dst)
{
    const int64_t N = int64_t(g.B) * g.n;
    float *src[kSnapFields];
    snap_ptrs(s, src);
    for (int k = 0; k < kSnapFields; ++k) copy(dst + k * N, src[k], N);
}

static void restore_snapshot(const Geo &g, FwdState &s, const float *src)
{
    const int64_t N = int64_t(g.B) * g.n;
    float *dst[kSnapFields];
    snap_ptrs(s, dst);
    for (int k = 0; k < kSnapFields; ++k) copy(dst[k], src + k * N, N);
}

static void forward_step(const Ctx &x, const swp_problem &p, FwdState &s, int it, float *u_this, float *strip,
                         float *record)
{
    const Geo &g = x.g;
    StepFields f{};
    f.vp = p.models[0];
    f.u_prev = s.u[0];
    f.u_now = s.u[1];
    f.u_next = s.u[2];
    for (int a = 0; a < 3; ++a) {
        f.psi[a] = s.psi[a];
        f.psin[a] = s.psin[a];
        f.zeta[a] = s.zeta[a];
    }
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
    for (int k = 0; k < 3; ++k) std::swap(s.psi[k], s.psin[k]);
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

void forward(const swp_problem &p_in, float *record)
{
    const Ctx x = make_ctx(p_in);
    const Geo &g = x.g;
    Carver ws(p_in.workspace, p_in.workspace_bytes);
    FwdState s;
    carve_fwd(ws, g, s);
    zero_fwd(g, s);
    const IoBufs io = carve_io(ws, p_in, g, 0);
    const swp_problem p = with_pitched_model(p_in, g, io);
    Carver tc(p.tape, p.tape_bytes);
    Tape t = carve_tape(tc, p, g);
    const int64_t N = int64_t(g.B) * g.n;
    std::vector<SpinePt> spine;
    if (p.memory == SWP_MEM_BINOMIAL) spine = revolve_spine(g.nt, t.nchunk);
    size_t sp = 0;
    for (int it = 0; it < g.nt; ++it) {
        if (p.memory == SWP_MEM_CKPT && it % p.ckpt_interval == 0)
            snapshot(g, s, t.snaps + int64_t(it / p.ckpt_interval) * kSnapFields * N);
        for (; sp < spine.size() && spine[sp].step == it; ++sp)
            snapshot(g, s, t.snaps + int64_t(spine[sp].slot) * kSnapFields * N);
        float *tape_strips = t.strips ? strip_at(t.strips, int64_t(it) * g.B * g.ns, g.sbf16) : nullptr;
        forward_step(x, p, s, it, t.u_this ? t.u_this + int64_t(it) * N : nullptr,
                     io.stage ? io.stage : tape_strips, record);
        if (io.stage) pack_bf16(tape_strips, io.stage, int64_t(g.B) * g.ns);
    }
    if (t.last_two) {
        copy(t.last_two, s.u[0], N);
        copy(t.last_two + N, s.u[1], N);
    }
}

static void adjoint_step(const Ctx &x, const swp_problem &p, const swp_backward_io &io, AdjState &a, int it,
                         const float *img, float *grad)
{
    const Geo &g = x.g;
    AdjFields f{};
    f.vp = p.models[0];
    f.l_prev = a.l[0];
    f.l_now = a.l[1];
    f.l_next = a.l[2];
    for (int k = 0; k < 3; ++k) {
        f.psi[k] = a.psi[k];
        f.zeta[k] = a.zeta[k];
        f.psio[k] = a.psin[k];
        f.zetao[k] = a.zetan[k];
    }
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
    for (int k = 0; k < 3; ++k) {
        std::swap(a.psi[k], a.psin[k]);
        std::swap(a.zeta[k], a.zetan[k]);
    }
}

static void image(const Ctx &x, const swp_problem &p, const float *img, const float *fnow, const float *fprev,
                  const float *lam, float *grad, float *S, float *R)
{
    const Geo &g = x.g;
    if (!grad && !S && !R) return;
    TimedScope ts(K_OTHER);
    hipLaunchKernelGGL(image_kernel, grid_for(g.nxp, g.nyp, g.nzp, g.B), dim3(64, 4), 0, stream(), g, img, fnow,
                       fprev, lam, p.models[0], grad, S, R);
    SWP_LAUNCHED();
}

static StripW strip_widths(const Geo &g)
{
    StripW w;
    const int ez = std::max(g.nzp, 0), ey = std::max(g.nyp, 0), ex = std::max(g.nxp, 0);
    w.zl = std::min(g.W, ez);
    w.zh = std::min(g.W, std::max(ez - w.zl, 0));
    w.yl = std::min(g.W, ey);
    w.yh = std::min(g.W, std::max(ey - w.yl, 0));
    w.xl = std::min(g.W, ex);
    w.xh = std::min(g.W, std::max(ex - w.xl, 0));
    return w;
}

static void backward_run(const Ctx &x, const swp_problem &p, const swp_backward_io &io, Carver &ws,
                         FwdState &fs, AdjState &a);

void backward(const swp_problem &p_in, const swp_backward_io &io_in)
{
    const Ctx x = make_ctx(p_in);
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
    const int64_t N = int64_t(g.B) * g.n, rows = int64_t(g.B) * g.nz * g.ny;
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
            case REV_SAVE: snapshot(g, fs, t.snaps + int64_t(o.arg) * kSnapFields * N); break;
            case REV_LOAD: restore_snapshot(g, fs, t.snaps + int64_t(o.arg) * kSnapFields * N); break;
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
            restore_snapshot(g, fs, t.snaps + int64_t(c) * kSnapFields * N);
            for (int it = it0; it < it1; ++it) forward_step(x, p, fs, it, chunk + int64_t(it - it0) * N, nullptr, nullptr);
            for (int it = it1 - 1; it >= it0; --it) {
                adjoint_step(x, p, io, a, it, nullptr, nullptr);
                image(x, p, chunk + int64_t(it - it0) * N, nullptr, nullptr, a.l[1], grad, S, R);
            }
        }
        return;
    }

    SWP_CHECK(p.memory == SWP_MEM_BOUNDARY, "acoustic3d backward: unknown memory mode ", p.memory);
    // Reconstruction seed (prev, now) = (w^nt, w^(nt-1)); the 3-D driver keeps
    // the absorbing rim (no zeroing, acoustic3d/driver_traits.cuh).
    float *r[3] = {fs.u[0], fs.u[1], fs.u[2]};
    fill_zero(r[2], N);
    copy(r[0], t.last_two + N, N);
    copy(r[1], t.last_two, N);
    const int M = g.M;
    Box bx;
    bx.x0 = g.pxl > 0 ? g.pxl + 2 * M + 1 : 2 * M;
    bx.x1 = g.nx - (g.pxh > 0 ? g.pxh + 2 * M + 1 : 2 * M);
    bx.y0 = g.pyl > 0 ? g.pyl + 2 * M + 1 : 2 * M;
    bx.y1 = g.ny - (g.pyh > 0 ? g.pyh + 2 * M + 1 : 2 * M);
    bx.z0 = g.pzl > 0 ? g.pzl + 2 * M + 1 : 2 * M;
    bx.z1 = g.nz - (g.pzh > 0 ? g.pzh + 2 * M + 1 : 2 * M);
    const bool recon = bx.x0 < bx.x1 && bx.y0 < bx.y1 && bx.z0 < bx.z1;
    const StripW sw = strip_widths(g);
    const int64_t mz = std::max(g.nzp - sw.zl - sw.zh, 0), my = std::max(g.nyp - sw.yl - sw.yh, 0);
    const int64_t n_strip = int64_t(sw.zl + sw.zh) * g.nyp * g.nxp + mz * (sw.yl + sw.yh) * g.nxp +
                            mz * my * (sw.xl + sw.xh);
    for (int it = g.nt - 1; it >= 1; --it) {
        adjoint_step(x, p, io, a, it, nullptr, nullptr);
        const float *lam = a.l[1];
        if (recon) launch_recon(x, p.models[0], r[0], r[1], r[2], lam, grad, bx);
        if (n_strip > 0) {
            TimedScope ts(K_OTHER);
            hipLaunchKernelGGL(strip_restore_kernel, dim3(unsigned((n_strip + 255) / 256), unsigned(g.B)), dim3(256),
                               0, stream(), g, strip_at(t.strips, int64_t(it - 1) * g.B * g.ns, g.sbf16), r[2], r[1], r[0], lam,
                               p.models[0], grad, p.source, p.sources_loc, p.nsrc, it, sw);
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
    SWP_CHECK(p.ndim == 3, "acoustic3d: ndim must be 3, got ", p.ndim);
    SWP_CHECK(p.nz > 2 * p.M && p.ny > 2 * p.M && p.nx > 2 * p.M, "acoustic3d: grid smaller than the stencil");
    SWP_CHECK(p.B >= 1, "acoustic3d: B must be >= 1");
    SWP_CHECK(p.models[0] != nullptr, "acoustic3d: vp is required");
    for (int k = 0; k < 9; ++k) SWP_CHECK(p.pml[k] != nullptr, "acoustic3d: CPML profile ", k, " missing");
    SWP_CHECK(p.lap_coes && p.grad_coes, "acoustic3d: FD coefficient tables missing");
    Ctx x;
    x.g = make_geo(p);
    SWP_CHECK(x.g.X1 > x.g.X0 && x.g.Y1 > x.g.Y0 && x.g.Z1 > x.g.Z0, "acoustic3d: empty physical box");
    x.c = a2d::make_coef(p, p.lap_coes, p.grad_coes);
    // profiles arrive in shape order (z, y, x): [a, b, db] per axis
    for (int a = 0; a < 3; ++a) {
        const int src = 2 - a;
        x.pr.a[a] = p.pml[3 * src + 0];
        x.pr.b[a] = p.pml[3 * src + 1];
        x.pr.db[a] = p.pml[3 * src + 2];
    }
    x.tz = tuning(SWP_TUNE_TZ3, "SWP_TZ3", 64);
    // Kernel family, as in 2-D: the z-marching kernels from kZmMinCells3
    // shot-cells (B x nz x ny x nx on the runtime grid) up, one thread per
    // cell below, where the few long z-march segments of a small launch are
    // latency-bound (40 x 36 x 44, one shot: 672 against 37 us per step of
    // forward + backward; the two meet near 41M shot-cells).
    // SWP_KERNELS=percell / zmarch forces a family.
    const char *kern = std::getenv("SWP_KERNELS");
    const std::string ks = kern ? kern : "";
    const int64_t cells = int64_t(x.g.B) * x.g.nz * x.g.ny * x.g.nx;
    x.zm = zm_supported(x.g, x.tz) && ks != "percell" && (ks == "zmarch" || cells >= kZmMinCells3);
    // SWP_SHELL3=percell runs the CPML shell with the per-cell kernels over
    // six slabs instead of the region-specialised z-marching kernels.
    const char *sh = std::getenv("SWP_SHELL3");
    x.shell_percell = sh && std::string(sh) == "percell";
    return x;
}

}  // namespace a3d
}  // namespace swp
