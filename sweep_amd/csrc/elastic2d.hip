// 2-D isotropic elastic path on gfx950: velocity-stress on the Virieux
// staggered grid with the cpmls CPML (memory variables m = a m + b d), the
// exact adjoint (stress / velocity prepare + apply), vp / vs / rho imaging
// with the body-force rho corrections at velocity sources and receivers,
// full-storage, checkpoint (chunk replay, binomial) and boundary-saving
// backwards.  FP32 on the vector ALUs; wave64 rows along x.
//
// Arithmetic follows the reference's elastic2d kernels term for term
// (csrc/cuda/equations/elastic2d/kernels.cuh:48-330 forward, 547-1055
// adjoint, 1183-1260 imaging; operators/staggered.cuh; common/common.cu:40-116,
// 182-210; common/derived_models.cu:27-37) without the free-surface
// branches; the time loops follow common/sg_driver.cuh:23-110, 285-340.
#include "elastic2d.h"
#include "revolve.h"

#include <algorithm>
#include <cstddef>
#include <type_traits>

namespace swp {
namespace e2d {

bool available() { return true; }

namespace {

struct Fields {
    float *f[NSTATE];
};

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
    g.n = int64_t(g.nz) * g.nx;
    g.mstride = p.model_batched ? g.n : 0;
    g.dt = p.dt;
    g.dx = p.dx;
    g.dz = p.dz;
    g.inv_dx = 1.f / g.dx;
    g.inv_dz = 1.f / g.dz;
    return g;
}

// Staggered first derivative along stride s at cell i: FWD the forward type
// (u[i+1..] - u[i..]), else the backward type (operators/staggered.cuh).
template <int MS, bool FWD>
__device__ __forceinline__ float sgd(const SCoef &c, const float *__restrict__ u, int64_t i, int64_t s,
                                     float inv, int M)
{
    auto U = [&](int k) { return u[i + k * s]; };
    if constexpr (MS == 1) {
        return FWD ? (U(1) - U(0)) * inv : (U(0) - U(-1)) * inv;
    } else if constexpr (MS == 2) {
        constexpr float c1 = 9.f / 8.f, c2 = -1.f / 24.f;
        return FWD ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2))) * inv;
    } else if constexpr (MS == 3) {
        constexpr float c1 = 75.f / 64.f, c2 = -25.f / 384.f, c3 = 3.f / 640.f;
        return FWD ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1)) + c3 * (U(3) - U(-2))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2)) + c3 * (U(2) - U(-3))) * inv;
    } else if constexpr (MS == 4) {
        constexpr float c1 = 1225.f / 1024.f, c2 = -245.f / 3072.f, c3 = 49.f / 5120.f, c4 = -5.f / 7168.f;
        return FWD ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1)) + c3 * (U(3) - U(-2)) + c4 * (U(4) - U(-3))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2)) + c3 * (U(2) - U(-3)) + c4 * (U(3) - U(-4))) * inv;
    } else {
        float acc = 0.f;
        for (int m = 0; m < M; ++m)
            acc += FWD ? c.c[m] * (U(m + 1) - U(-m)) : c.c[m] * (U(m) - U(-m - 1));
        return acc * inv;
    }
}

__device__ __forceinline__ bool in_pml(const Geo &g, int ix, int iz)
{
    return ix < g.X0 || ix >= g.X1 || iz < g.Z0 || iz >= g.Z1;
}

// ---------------------------------------------------------------- boundary strips
// Boundary saving keeps, per face of the physical box, a band W = M + 1
// cells wide starting M cells outside the face (boundary/strip.cuh:81-114
// with offset -M, no tangential pad): top and bottom bands span x in
// [X0, X1), left and right bands z in [Z0, Z1).  Strip cell s of a shot:
// top (W, nxp), bottom (W, nxp), left (nzp, W), right (nzp, W).
__host__ __device__ inline int64_t strip_cells(const Geo &g)
{
    return int64_t(2) * (g.M + 1) * ((g.X1 - g.X0) + (g.Z1 - g.Z0));
}

__device__ __forceinline__ void strip_cell(const Geo &g, int64_t s, int &ix, int &iz)
{
    const int W = g.M + 1, nxp = g.X1 - g.X0, nzp = g.Z1 - g.Z0;
    const int64_t a = int64_t(W) * nxp, c = int64_t(W) * nzp;
    if (s < a) {
        iz = g.Z0 - g.M + int(s / nxp);
        ix = g.X0 + int(s % nxp);
    } else if (s < 2 * a) {
        s -= a;
        iz = g.Z1 - 1 + int(s / nxp);
        ix = g.X0 + int(s % nxp);
    } else if (s < 2 * a + c) {
        s -= 2 * a;
        iz = g.Z0 + int(s / W);
        ix = g.X0 - g.M + int(s % W);
    } else {
        s -= 2 * a + c;
        iz = g.Z0 + int(s / W);
        ix = g.X1 - 1 + int(s % W);
    }
}

enum { STRIP_SAVE = 0, STRIP_RESTORE = 1, STRIP_COPY = 2 };

// Fields [f0, f1) at the strip cells: SAVE strip[f, b, s] = src_f, RESTORE
// dst_f = strip[f, b, s], COPY dst_f = src_f.  strip: one step's (5, B, ns).
// Grid (ceil(ns / 256), B).
__global__ void __launch_bounds__(256) strip_kernel(Geo g, Fields src, Fields dst, float *__restrict__ strip, int f0,
                                                    int f1, int mode)
{
    const int64_t ns = strip_cells(g);
    const int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (s >= ns) return;
    const int b = blockIdx.y;
    int ix, iz;
    strip_cell(g, s, ix, iz);
    const int64_t i = int64_t(b) * g.n + int64_t(iz) * g.nx + ix;
    for (int f = f0; f < f1; ++f) {
        const int64_t k = (int64_t(f) * g.B + b) * ns + s;
        if (mode == STRIP_SAVE) strip[k] = src.f[f][i];
        else if (mode == STRIP_RESTORE) dst.f[f][i] = strip[k];
        else dst.f[f][i] = src.f[f][i];
    }
}

// Reverse steps without CPML on [X0 + 1, X1 - 1) x [Z0 + 1, Z1 - 1), the
// physical box less one cell (abcn + M + 1 from the grid edge), the inverse
// of the forward updates there (kernels.cuh:365-545): stress first
// (s -= dt C grad v), then velocity (v -= dt / rho div s), which first copies
// v into the carriers (cvx, cvz) for the next reverse step's imaging.
template <int MS>
__global__ void __launch_bounds__(256) stress_nopml_kernel(Geo g, SCoef c, Fields s, const float *__restrict__ lamb,
                                                           const float *__restrict__ mub)
{
    const int M = MS > 0 ? MS : g.M;
    const BlockXYZ blk = xcd_block();
    const int ix = g.X0 + 1 + blk.x * 64 + int(threadIdx.x);
    const int iz = g.Z0 + 1 + blk.y * 4 + int(threadIdx.y);
    const int b = blk.z;
    if (ix >= g.X1 - 1 || iz >= g.Z1 - 1) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float *vx = s.f[VX] + off, *vz = s.f[VZ] + off;
    const float dvx_dx = sgd<MS, false>(c, vx, i, 1, g.inv_dx, M);
    const float dvz_dz = sgd<MS, false>(c, vz, i, nx, g.inv_dz, M);
    const float dvx_dz = sgd<MS, true>(c, vx, i, nx, g.inv_dz, M);
    const float dvz_dx = sgd<MS, true>(c, vz, i, 1, g.inv_dx, M);
    const int64_t mi = int64_t(b) * g.mstride + i;
    const float la = lamb[mi], mu = mub[mi];
    s.f[SXX][off + i] -= g.dt * ((la + 2.f * mu) * dvx_dx + la * dvz_dz);
    s.f[SZZ][off + i] -= g.dt * ((la + 2.f * mu) * dvz_dz + la * dvx_dx);
    s.f[SXZ][off + i] -= g.dt * mu * (dvx_dz + dvz_dx);
}

template <int MS>
__global__ void __launch_bounds__(256) vel_nopml_kernel(Geo g, SCoef c, Fields s, const float *__restrict__ rhob,
                                                        float *__restrict__ cvx, float *__restrict__ cvz)
{
    const int M = MS > 0 ? MS : g.M;
    const BlockXYZ blk = xcd_block();
    const int ix = g.X0 + 1 + blk.x * 64 + int(threadIdx.x);
    const int iz = g.Z0 + 1 + blk.y * 4 + int(threadIdx.y);
    const int b = blk.z;
    if (ix >= g.X1 - 1 || iz >= g.Z1 - 1) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float *sxx = s.f[SXX] + off, *szz = s.f[SZZ] + off, *sxz = s.f[SXZ] + off;
    const float dsxx_dx = sgd<MS, true>(c, sxx, i, 1, g.inv_dx, M);
    const float dsxz_dz = sgd<MS, false>(c, sxz, i, nx, g.inv_dz, M);
    const float dsxz_dx = sgd<MS, false>(c, sxz, i, 1, g.inv_dx, M);
    const float dszz_dz = sgd<MS, true>(c, szz, i, nx, g.inv_dz, M);
    const float inv_rho = 1.f / rhob[int64_t(b) * g.mstride + i];
    cvx[off + i] = s.f[VX][off + i];
    cvz[off + i] = s.f[VZ][off + i];
    s.f[VX][off + i] -= g.dt * inv_rho * (dsxx_dx + dsxz_dz);
    s.f[VZ][off + i] -= g.dt * inv_rho * (dsxz_dx + dszz_dz);
}

// ---------------------------------------------------------------- models
// mu = (rho vs) vs, lambda = rho (vp vp - (2 vs) vs) (derived_models.cu:27-37)
__global__ void __launch_bounds__(256) lame_kernel(const float *__restrict__ vp, const float *__restrict__ vs,
                                                   const float *__restrict__ rho, float *__restrict__ mu,
                                                   float *__restrict__ lam, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const float v = vp[i], s = vs[i], r = rho[i];
        mu[i] = (r * s) * s;
        lam[i] = r * ((v * v) - (2.0f * s) * s);
    }
}

// ---------------------------------------------------------------- forward
// velocity: v += dt / rho (div s), CPML memory on the band (kernels.cuh:48-146)
template <int MS>
__global__ void __launch_bounds__(256) vel_kernel(Geo g, SCoef c, Prof pr, Fields s, const float *__restrict__ rhob)
{
    const int M = MS > 0 ? MS : g.M;
    const BlockXYZ blk = xcd_block();
    const int ix = M + blk.x * 64 + int(threadIdx.x);
    const int iz = M + blk.y * 4 + int(threadIdx.y);
    const int b = blk.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float *sxx = s.f[SXX] + off, *szz = s.f[SZZ] + off, *sxz = s.f[SXZ] + off;
    float dsxx_dx = sgd<MS, true>(c, sxx, i, 1, g.inv_dx, M);
    float dsxz_dz = sgd<MS, false>(c, sxz, i, nx, g.inv_dz, M);
    float dsxz_dx = sgd<MS, false>(c, sxz, i, 1, g.inv_dx, M);
    float dszz_dz = sgd<MS, true>(c, szz, i, nx, g.inv_dz, M);
    const float inv_rho = 1.f / rhob[int64_t(b) * g.mstride + i];
    if (in_pml(g, ix, iz)) {
        float m;
        m = pr.azh[iz] * s.f[M_SZZZ][off + i] + pr.bzh[iz] * dszz_dz;
        s.f[M_SZZZ][off + i] = m;
        dszz_dz += m;
        m = pr.ax[ix] * s.f[M_SXZX][off + i] + pr.bx[ix] * dsxz_dx;
        s.f[M_SXZX][off + i] = m;
        dsxz_dx += m;
        m = pr.az[iz] * s.f[M_SXZZ][off + i] + pr.bz[iz] * dsxz_dz;
        s.f[M_SXZZ][off + i] = m;
        dsxz_dz += m;
        m = pr.axh[ix] * s.f[M_SXXX][off + i] + pr.bxh[ix] * dsxx_dx;
        s.f[M_SXXX][off + i] = m;
        dsxx_dx += m;
    }
    s.f[VX][off + i] += g.dt * inv_rho * (dsxx_dx + dsxz_dz);
    s.f[VZ][off + i] += g.dt * inv_rho * (dsxz_dx + dszz_dz);
}

// stress: s += dt C grad v, CPML memory on the band; optional capture of the
// velocities for the imaging (kernels.cuh:182-330)
template <int MS>
__global__ void __launch_bounds__(256) stress_kernel(Geo g, SCoef c, Prof pr, Fields s, const float *__restrict__ lamb,
                                                     const float *__restrict__ mub, float *__restrict__ vcap)
{
    const int M = MS > 0 ? MS : g.M;
    const BlockXYZ blk = xcd_block();
    const int ix = M + blk.x * 64 + int(threadIdx.x);
    const int iz = M + blk.y * 4 + int(threadIdx.y);
    const int b = blk.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float *vx = s.f[VX] + off, *vz = s.f[VZ] + off;
    float dvx_dx = sgd<MS, false>(c, vx, i, 1, g.inv_dx, M);
    float dvz_dz = sgd<MS, false>(c, vz, i, nx, g.inv_dz, M);
    float dvx_dz = sgd<MS, true>(c, vx, i, nx, g.inv_dz, M);
    float dvz_dx = sgd<MS, true>(c, vz, i, 1, g.inv_dx, M);
    const int64_t mi = int64_t(b) * g.mstride + i;
    const float la = lamb[mi], mu = mub[mi];
    const float dt = g.dt;
    if (!in_pml(g, ix, iz)) {
        s.f[SXX][off + i] += dt * ((la + 2.f * mu) * dvx_dx + la * dvz_dz);
        s.f[SZZ][off + i] += dt * ((la + 2.f * mu) * dvz_dz + la * dvx_dx);
        s.f[SXZ][off + i] += dt * mu * (dvx_dz + dvz_dx);
    } else {
        float m;
        m = pr.az[iz] * s.f[M_VZZ][off + i] + pr.bz[iz] * dvz_dz;
        s.f[M_VZZ][off + i] = m;
        dvz_dz += m;
        m = pr.ax[ix] * s.f[M_VXX][off + i] + pr.bx[ix] * dvx_dx;
        s.f[M_VXX][off + i] = m;
        dvx_dx += m;
        s.f[SXX][off + i] += dt * ((la + 2.f * mu) * dvx_dx + la * dvz_dz);
        s.f[SZZ][off + i] += dt * ((la + 2.f * mu) * dvz_dz + la * dvx_dx);
        m = pr.azh[iz] * s.f[M_VXZ][off + i] + pr.bzh[iz] * dvx_dz;
        s.f[M_VXZ][off + i] = m;
        dvx_dz += m;
        m = pr.axh[ix] * s.f[M_VZX][off + i] + pr.bxh[ix] * dvz_dx;
        s.f[M_VZX][off + i] = m;
        dvz_dx += m;
        s.f[SXZ][off + i] += dt * mu * (dvx_dz + dvz_dx);
    }
    if (vcap) {   // (2, B, n): vx then vz
        vcap[off + i] = vx[i];
        vcap[int64_t(g.B) * g.n + off + i] = vz[i];
    }
}

// ---------------------------------------------------------------- adjoint
// stress prepare (every cell): q = dt C^T s_bar, CPML-filtered on the band
// (kernels.cuh:547-742)
__global__ void __launch_bounds__(256) stress_prep_kernel(Geo g, Prof pr, Fields a, const float *__restrict__ lamb,
                                                          const float *__restrict__ mub, Fields q)
{
    const int ix = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx || iz >= g.nz) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix;
    const int64_t mi = int64_t(b) * g.mstride + i;
    const float la = lamb[mi], mu = mub[mi], dt = g.dt;
    const float bar_sxx = a.f[SXX][off + i], bar_szz = a.f[SZZ][off + i], bar_sxz = a.f[SXZ][off + i];
    const float bar_dvx_dx = dt * ((la + 2.f * mu) * bar_sxx + la * bar_szz);
    const float bar_dvz_dz = dt * ((la + 2.f * mu) * bar_szz + la * bar_sxx);
    const float bar_dvx_dz = dt * mu * bar_sxz;
    const float bar_dvz_dx = dt * mu * bar_sxz;
    // q.f[0..3] = qxx, qzz, qxz, qzx
    if (!in_pml(g, ix, iz)) {
        q.f[0][off + i] = bar_dvx_dx;
        q.f[1][off + i] = bar_dvz_dz;
        q.f[2][off + i] = bar_dvx_dz;
        q.f[3][off + i] = bar_dvz_dx;
        return;
    }
    const float t_vxx = a.f[M_VXX][off + i] + bar_dvx_dx, t_vzz = a.f[M_VZZ][off + i] + bar_dvz_dz;
    const float t_vxz = a.f[M_VXZ][off + i] + bar_dvx_dz, t_vzx = a.f[M_VZX][off + i] + bar_dvz_dx;
    q.f[0][off + i] = bar_dvx_dx + pr.bx[ix] * t_vxx;
    q.f[1][off + i] = bar_dvz_dz + pr.bz[iz] * t_vzz;
    q.f[2][off + i] = bar_dvx_dz + pr.bzh[iz] * t_vxz;
    q.f[3][off + i] = bar_dvz_dx + pr.bxh[ix] * t_vzx;
    a.f[M_VXX][off + i] = pr.ax[ix] * t_vxx;
    a.f[M_VZX][off + i] = pr.axh[ix] * t_vzx;
    a.f[M_VZZ][off + i] = pr.az[iz] * t_vzz;
    a.f[M_VXZ][off + i] = pr.azh[iz] * t_vxz;
}

// stress apply: v_bar += transposed derivatives of q (kernels.cuh:744-832)
template <int MS>
__global__ void __launch_bounds__(256) stress_apply_kernel(Geo g, SCoef c, Fields a, Fields q)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = M + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = M + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float dqxx_dx = sgd<MS, true>(c, q.f[0] + off, i, 1, g.inv_dx, M);
    const float dqxz_dz = sgd<MS, false>(c, q.f[2] + off, i, nx, g.inv_dz, M);
    const float dqzx_dx = sgd<MS, false>(c, q.f[3] + off, i, 1, g.inv_dx, M);
    const float dqzz_dz = sgd<MS, true>(c, q.f[1] + off, i, nx, g.inv_dz, M);
    a.f[VX][off + i] += dqxx_dx + dqxz_dz;
    a.f[VZ][off + i] += dqzx_dx + dqzz_dz;
}

// velocity prepare (every cell): p = dt / rho v_bar, CPML-filtered (kernels.cuh:870-960)
__global__ void __launch_bounds__(256) vel_prep_kernel(Geo g, Prof pr, Fields a, const float *__restrict__ rhob, Fields p)
{
    const int ix = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx || iz >= g.nz) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix;
    const float inv_rho = 1.f / rhob[int64_t(b) * g.mstride + i];
    const float vx = a.f[VX][off + i], vz = a.f[VZ][off + i];
    const float bar_dsxx_dx = g.dt * inv_rho * vx, bar_dsxz_dz = g.dt * inv_rho * vx;
    const float bar_dsxz_dx = g.dt * inv_rho * vz, bar_dszz_dz = g.dt * inv_rho * vz;
    // p.f[4..7] = pxx, pzz, pxz, pzx
    if (!in_pml(g, ix, iz)) {
        p.f[4][off + i] = bar_dsxx_dx;
        p.f[6][off + i] = bar_dsxz_dz;
        p.f[7][off + i] = bar_dsxz_dx;
        p.f[5][off + i] = bar_dszz_dz;
        return;
    }
    const float t_sxxx = a.f[M_SXXX][off + i] + bar_dsxx_dx, t_sxzz = a.f[M_SXZZ][off + i] + bar_dsxz_dz;
    const float t_sxzx = a.f[M_SXZX][off + i] + bar_dsxz_dx, t_szzz = a.f[M_SZZZ][off + i] + bar_dszz_dz;
    p.f[4][off + i] = bar_dsxx_dx + pr.bxh[ix] * t_sxxx;
    p.f[6][off + i] = bar_dsxz_dz + pr.bz[iz] * t_sxzz;
    p.f[7][off + i] = bar_dsxz_dx + pr.bx[ix] * t_sxzx;
    p.f[5][off + i] = bar_dszz_dz + pr.bzh[iz] * t_szzz;
    a.f[M_SXXX][off + i] = pr.axh[ix] * t_sxxx;
    a.f[M_SXZX][off + i] = pr.ax[ix] * t_sxzx;
    a.f[M_SXZZ][off + i] = pr.az[iz] * t_sxzz;
    a.f[M_SZZZ][off + i] = pr.azh[iz] * t_szzz;
}

// velocity apply: s_bar += transposed derivatives of p (kernels.cuh:985-1055)
template <int MS>
__global__ void __launch_bounds__(256) vel_apply_kernel(Geo g, SCoef c, Fields a, Fields p)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = M + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = M + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.nx - M || iz >= g.nz - M) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const float dpxx_dx = sgd<MS, false>(c, p.f[4] + off, i, 1, g.inv_dx, M);
    const float dpzz_dz = sgd<MS, false>(c, p.f[5] + off, i, nx, g.inv_dz, M);
    const float dpxz_dz = sgd<MS, true>(c, p.f[6] + off, i, nx, g.inv_dz, M);
    const float dpzx_dx = sgd<MS, true>(c, p.f[7] + off, i, 1, g.inv_dx, M);
    a.f[SXX][off + i] += dpxx_dx;
    a.f[SZZ][off + i] += dpzz_dz;
    a.f[SXZ][off + i] += dpxz_dz + dpzx_dx;
}

// vp / vs / rho imaging over the physical box (kernels.cuh:1183-1260)
struct ImgArgs {
    const float *vxn, *vzn, *vxp, *vzp;   // v(it), v(it+1), (B, n) each
    const float *vp, *vs, *rho;
    float *gvp, *gvs, *grho;
};

template <int MS>
__global__ void __launch_bounds__(256) image_kernel(Geo g, SCoef c, Fields a, ImgArgs m)
{
    const int M = MS > 0 ? MS : g.M;
    const int ix = g.X0 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int iz = g.Z0 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int b = blockIdx.z;
    if (ix >= g.X1 || iz >= g.Z1) return;
    const int64_t off = int64_t(b) * g.n, i = int64_t(iz) * g.nx + ix, nx = g.nx;
    const int64_t mi = int64_t(b) * g.mstride + i;
    const float *fvx = m.vxn + off, *fvz = m.vzn + off;
    const float fvx_x = sgd<MS, false>(c, fvx, i, 1, g.inv_dx, M);
    const float fvz_z = sgd<MS, false>(c, fvz, i, nx, g.inv_dz, M);
    const float fvx_z = sgd<MS, true>(c, fvx, i, nx, g.inv_dz, M);
    const float fvz_x = sgd<MS, true>(c, fvz, i, 1, g.inv_dx, M);
    const float bar_sxx = a.f[SXX][off + i], bar_szz = a.f[SZZ][off + i], bar_sxz = a.f[SXZ][off + i];
    const float grad_lambda = (bar_sxx + bar_szz) * (fvx_x + fvz_z);
    float grad_mu = 2.f * (bar_sxx * fvx_x + bar_szz * fvz_z);
    grad_mu += bar_sxz * (fvx_z + fvz_x);
    const float vp = m.vp[mi], vs = m.vs[mi], rho = m.rho[mi], dt = g.dt;
    m.gvp[off + i] += -2.f * rho * vp * grad_lambda * dt;
    m.gvs[off + i] += -(-4.f * rho * vs * grad_lambda + 2.f * rho * vs * grad_mu) * dt;
    float gr = m.grho[off + i];
    gr += (a.f[VX][off + i] * (fvx[i] - m.vxp[off + i]) + a.f[VZ][off + i] * (fvz[i] - m.vzp[off + i])) / rho;
    gr -= grad_lambda * (vp * vp - 2.f * vs * vs) * dt + grad_mu * (vs * vs) * dt;
    m.grho[off + i] = gr;
}

// Body-force rho corrections: at the velocity sources before the residual
// (common.cu:40-75) and at the velocity receivers after the step
// (common.cu:77-116).  One thread per (shot, point).
__global__ void __launch_bounds__(256) rho_src_kernel(Geo g, float *__restrict__ grho, const float *__restrict__ adj,
                                                      const float *__restrict__ rhob, const float *__restrict__ w,
                                                      const int32_t *__restrict__ loc, int nsrc, int it)
{
    const int b = blockIdx.y, s = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (s >= nsrc) return;
    const int32_t *l = loc + (int64_t(b) * nsrc + s) * 2;
    const int ix = l[0], iz = l[1];
    if (ix < 0 || ix >= g.nx || iz < 0 || iz >= g.nz) return;
    const int64_t i = int64_t(iz) * g.nx + ix, off = int64_t(b) * g.n;
    atomicAdd(grho + off + i, adj[off + i] * w[(int64_t(b) * nsrc + s) * g.nt + it] / rhob[int64_t(b) * g.mstride + i]);
}

__global__ void __launch_bounds__(256) rho_rec_kernel(Geo g, float *__restrict__ grho, const float *__restrict__ fn,
                                                      const float *__restrict__ fp, const float *__restrict__ rhob,
                                                      const float *__restrict__ res, const int32_t *__restrict__ loc,
                                                      int nrec, int it)
{
    const int b = blockIdx.y, r = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (r >= nrec) return;
    const int32_t *l = loc + (int64_t(b) * nrec + r) * 2;
    const int ix = l[0], iz = l[1], M = g.M;
    if (ix < M || ix >= g.nx - M || iz < M || iz >= g.nz - M) return;
    const int64_t i = int64_t(iz) * g.nx + ix, off = int64_t(b) * g.n;
    atomicAdd(grho + off + i,
              -res[(int64_t(b) * nrec + r) * g.nt + it] * (fn[off + i] - fp[off + i]) / rhob[int64_t(b) * g.mstride + i]);
}

__global__ void __launch_bounds__(256) copy_kernel(float *__restrict__ dst, const float *__restrict__ src, int64_t n)
{
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n; k += int64_t(gridDim.x) * 256) dst[k] = src[k];
}

// ---------------------------------------------------------------- host
struct Ctx {
    Geo g;
    SCoef c;
    Prof pr;
    int nsf, nrf;
    int sf[5], rf[5];
};

Ctx make_ctx(const swp_problem &p)
{
    SWP_CHECK(p.ndim == 2, "elastic2d: ndim must be 2, got ", p.ndim);
    SWP_CHECK(p.M >= 1 && p.M <= kMaxM, "elastic2d: half order M=", p.M, " outside [1, ", kMaxM, "]");
    SWP_CHECK(p.nz > 2 * p.M && p.nx > 2 * p.M, "elastic2d: grid smaller than the stencil");
    SWP_CHECK(p.models[0] && p.models[1] && p.models[2], "elastic2d: vp, vs and rho are required");
    for (int k = 0; k < 8; ++k) SWP_CHECK(p.pml[k] != nullptr, "elastic2d: CPML profile ", k, " missing");
    SWP_CHECK(p.grad_coes, "elastic2d: staggered coefficient table missing");
    SWP_CHECK(p.nsrc_fields >= 1 && p.nsrc_fields <= 5 && p.nrec_fields >= 1 && p.nrec_fields <= 5,
              "elastic2d: need 1..5 source and receiver fields");
    Ctx x{};
    x.g = make_geo(p);
    SWP_CHECK(x.g.X1 > x.g.X0 && x.g.Z1 > x.g.Z0, "elastic2d: empty physical box");
    for (int m = 0; m < p.M; ++m) x.c.c[m] = p.grad_coes[m];
    x.pr = Prof{p.pml[0], p.pml[1], p.pml[2], p.pml[3], p.pml[4], p.pml[5], p.pml[6], p.pml[7]};
    x.nsf = p.nsrc_fields;
    x.nrf = p.nrec_fields;
    for (int k = 0; k < x.nsf; ++k) {
        x.sf[k] = p.src_fields[k];
        SWP_CHECK(x.sf[k] >= 0 && x.sf[k] <= 4, "elastic2d: source field ", x.sf[k], " out of range");
    }
    for (int k = 0; k < x.nrf; ++k) {
        x.rf[k] = p.rec_fields[k];
        SWP_CHECK(x.rf[k] >= 0 && x.rf[k] <= 3, "elastic2d: receiver field ", x.rf[k], " out of range");
    }
    return x;
}

template <class K>
void by_order(int M, K &&k)
{
    switch (M) {
    case 1: k(std::integral_constant<int, 1>{}); break;
    case 2: k(std::integral_constant<int, 2>{}); break;
    case 3: k(std::integral_constant<int, 3>{}); break;
    case 4: k(std::integral_constant<int, 4>{}); break;
    default: k(std::integral_constant<int, 0>{}); break;
    }
}

dim3 grid_for(int nxc, int nzc, int B) { return dim3(unsigned((nxc + 63) / 64), unsigned((nzc + 3) / 4), unsigned(B)); }

void copy(float *dst, const float *src, int64_t n)
{
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(copy_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), dst, src, n);
    SWP_LAUNCHED();
}

struct Models {
    float *mu, *lam;   // model-sized (shared or per shot)
};

struct Work {
    Models md;
    Fields s;      // forward state
    Fields a;      // adjoint state
    Fields qp;     // q (0..3), p (4..7)
    float *zero;   // (2, B, n) zero velocities: v(nt)
    float *carry;  // (2, B, n) v(chunk end) for the checkpoint reverse; v(it + 1) for the binomial and
                   // boundary-saving ones
    float *chunk;  // (I, 2, B, n) replayed velocities; (2, B, n) v(it) for the binomial reverse
};

int64_t model_cells(const Geo &g) { return g.mstride ? int64_t(g.B) * g.n : g.n; }

void carve(Carver &cv, const swp_problem &p, const Geo &g, int kind, Work &w)
{
    const int64_t N = int64_t(g.B) * g.n;
    w.md.mu = cv.take<float>(model_cells(g));
    w.md.lam = cv.take<float>(model_cells(g));
    for (auto &f : w.s.f) f = cv.take<float>(N);
    if (kind == 0) return;
    for (auto &f : w.a.f) f = cv.take<float>(N);
    for (int k = 0; k < 8; ++k) w.qp.f[k] = cv.take<float>(N);
    w.zero = cv.take<float>(2 * N);
    if (p.memory == SWP_MEM_CKPT) {
        w.carry = cv.take<float>(2 * N);
        w.chunk = cv.take<float>(int64_t(p.ckpt_interval) * 2 * N);
    } else if (p.memory == SWP_MEM_BINOMIAL) {
        w.carry = cv.take<float>(2 * N);
        w.chunk = cv.take<float>(2 * N);
    } else if (p.memory == SWP_MEM_BOUNDARY) {
        w.carry = cv.take<float>(2 * N);
    }
}

constexpr int kSnap = NSTATE;   // every field of the state

struct Tape {
    float *vel = nullptr;    // FULL: (nt, 2, B, n) velocities after each step's velocity update
    float *snaps = nullptr;  // CKPT: (nchunk, 13, B, n)
    float *strips = nullptr; // BOUNDARY: (nt, 5, B, ns) the five fields' strips after each step
    float *last = nullptr;   // BOUNDARY: (5, B, n) the five fields after the last step
    int nchunk = 0;
};

Tape carve_tape(Carver &cv, const swp_problem &p, const Geo &g)
{
    Tape t;
    const int64_t N = int64_t(g.B) * g.n;
    if (p.memory == SWP_MEM_FULL) {
        t.vel = cv.take<float>(int64_t(g.nt) * 2 * N);
    } else if (p.memory == SWP_MEM_CKPT) {
        SWP_CHECK(p.ckpt_interval >= 1, "ckpt_interval must be >= 1");
        t.nchunk = (g.nt + p.ckpt_interval - 1) / p.ckpt_interval;
        t.snaps = cv.take<float>(int64_t(t.nchunk) * kSnap * N);
    } else if (p.memory == SWP_MEM_BINOMIAL) {
        SWP_CHECK(p.ckpt_interval >= 1, "binomial checkpointing needs >= 1 snapshot slot");
        t.nchunk = p.ckpt_interval;
        t.snaps = cv.take<float>(int64_t(t.nchunk) * kSnap * N);
    } else if (p.memory == SWP_MEM_BOUNDARY) {
        t.strips = cv.take<float>(int64_t(g.nt) * 5 * g.B * strip_cells(g));
        t.last = cv.take<float>(5 * N);
    }
    return t;
}

// The strip pass over fields [f0, f1) of one step's (5, B, ns) slice.
void strips(const Geo &g, const Fields &src, const Fields &dst, float *slice, int f0, int f1, int mode)
{
    const int64_t ns = strip_cells(g);
    if (ns <= 0) return;
    hipLaunchKernelGGL(strip_kernel, dim3(unsigned((ns + 255) / 256), unsigned(g.B)), dim3(256), 0, stream(), g, src,
                       dst, slice, f0, f1, mode);
    SWP_LAUNCHED();
}

void derive_models(const swp_problem &p, const Geo &g, const Models &md)
{
    const int64_t n = model_cells(g);
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(lame_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), p.models[0], p.models[1],
                       p.models[2], md.mu, md.lam, n);
    SWP_LAUNCHED();
}

PointArgs point_args(const Geo &g)
{
    PointArgs a{};
    a.nt = g.nt;
    a.ndim = 2;
    a.nz = g.nz;
    a.ny = 1;
    a.nx = g.nx;
    a.n = g.n;
    a.sign = 1.0f;
    return a;
}

// One forward step: velocity, stress (capturing v when vcap), sources, records.
void forward_step(const Ctx &x, const swp_problem &p, Work &w, int it, float *vcap, float *record)
{
    const Geo &g = x.g;
    const dim3 grid = grid_for(g.nx - 2 * g.M, g.nz - 2 * g.M, g.B);
    {
        TimedScope ts(K_FWD);
        by_order(g.M, [&](auto ms) {
            hipLaunchKernelGGL(vel_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c, x.pr, w.s,
                               p.models[2]);
            hipLaunchKernelGGL(stress_kernel<decltype(ms)::value>, grid, dim3(64, 4), 0, stream(), g, x.c, x.pr, w.s,
                               w.md.lam, w.md.mu, vcap);
        });
        SWP_LAUNCHED();
    }
    TimedScope ts(K_OTHER);
    for (int k = 0; k < x.nsf; ++k) {
        PointArgs a = point_args(g);
        a.it = it;
        a.u_add = w.s.f[x.sf[k]];
        a.add_w = p.source;
        a.add_loc = p.sources_loc;
        a.nadd = p.nsrc;
        launch_points(a, g.B);
    }
    if (record) {
        for (int k = 0; k < x.nrf; ++k) {
            PointArgs a = point_args(g);
            a.it = it;
            a.u_get = w.s.f[x.rf[k]];
            a.get_out = record + int64_t(k) * g.B * p.nrec * g.nt;
            a.get_loc = p.receivers_loc;
            a.nget = p.nrec;
            launch_points(a, g.B);
        }
    }
}

void zero_state(const Geo &g, Fields &f)
{
    for (auto p : f.f) fill_zero(p, int64_t(g.B) * g.n);
}

void snapshot(const Geo &g, const Fields &s, float *dst)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (int k = 0; k < kSnap; ++k) copy(dst + k * N, s.f[k], N);
}

void restore(const Geo &g, Fields &s, const float *src)
{
    const int64_t N = int64_t(g.B) * g.n;
    for (int k = 0; k < kSnap; ++k) copy(s.f[k], src + k * N, N);
}

// The velocities one reverse step images against: v(it) and v(it + 1), each
// as (vx, vz) fields of (B, n).
struct VelPair {
    const float *vxn, *vzn, *vxp, *vzp;
};

VelPair pair_of(const float *vnow, const float *vnext, int64_t N) { return {vnow, vnow + N, vnext, vnext + N}; }

// One reverse step it: rho correction at the sources, signed residuals,
// imaging against v(it) / v(it+1), the adjoint step (it > 0), rho correction
// at the receivers (sg_driver.cuh:40-50).  With `between` the boundary-saving
// reconstruction of the stresses runs after the residuals and before the
// imaging (driver_traits.cuh:690-733).
template <class Between = std::nullptr_t>
void reverse_step(const Ctx &x, const swp_problem &p, const swp_backward_io &io, Work &w, int it, const VelPair &v,
                  Between between = nullptr)
{
    const Geo &g = x.g;
    float *gvp = io.grad_models[0], *gvs = io.grad_models[1], *grho = io.grad_models[2];
    SWP_CHECK(gvp && gvs && grho, "elastic2d backward: grad_models[0..2] (vp, vs, rho) are required");
    {
        TimedScope ts(K_OTHER);
        for (int k = 0; k < x.nsf; ++k) {
            if (x.sf[k] > 1 || p.nsrc == 0) continue;
            hipLaunchKernelGGL(rho_src_kernel, dim3(unsigned((p.nsrc + 255) / 256), unsigned(g.B)), dim3(256), 0,
                               stream(), g, grho, w.a.f[x.sf[k]], p.models[2], p.source, p.sources_loc, p.nsrc, it);
        }
        for (int k = 0; k < x.nrf; ++k) {
            PointArgs a = point_args(g);
            a.it = it;
            a.u_add = w.a.f[x.rf[k]];
            a.add_w = io.adjoint_source + int64_t(k) * g.B * p.nrec * g.nt;
            a.add_loc = p.receivers_loc;
            a.nadd = p.nrec;
            a.sign = x.rf[k] >= 2 ? -1.f : 1.f;
            launch_points(a, g.B);
        }
        SWP_LAUNCHED();
    }
    if constexpr (!std::is_same_v<Between, std::nullptr_t>) between();
    ImgArgs m{v.vxn, v.vzn, v.vxp, v.vzp, p.models[0], p.models[1], p.models[2], gvp, gvs, grho};
    {
        TimedScope ts(K_RECON);
        by_order(g.M, [&](auto ms) {
            hipLaunchKernelGGL(image_kernel<decltype(ms)::value>, grid_for(g.X1 - g.X0, g.Z1 - g.Z0, g.B),
                               dim3(64, 4), 0, stream(), g, x.c, w.a, m);
        });
        SWP_LAUNCHED();
    }
    if (it > 0) {
        TimedScope ts(K_ADJ);
        const dim3 all = grid_for(g.nx, g.nz, g.B), in = grid_for(g.nx - 2 * g.M, g.nz - 2 * g.M, g.B);
        by_order(g.M, [&](auto ms) {
            constexpr int MS = decltype(ms)::value;
            hipLaunchKernelGGL(stress_prep_kernel, all, dim3(64, 4), 0, stream(), g, x.pr, w.a, w.md.lam, w.md.mu, w.qp);
            hipLaunchKernelGGL(stress_apply_kernel<MS>, in, dim3(64, 4), 0, stream(), g, x.c, w.a, w.qp);
            hipLaunchKernelGGL(vel_prep_kernel, all, dim3(64, 4), 0, stream(), g, x.pr, w.a, p.models[2], w.qp);
            hipLaunchKernelGGL(vel_apply_kernel<MS>, in, dim3(64, 4), 0, stream(), g, x.c, w.a, w.qp);
        });
        SWP_LAUNCHED();
    }
    TimedScope ts(K_OTHER);
    for (int k = 0; k < x.nrf; ++k) {
        if (x.rf[k] > 1 || p.nrec == 0) continue;
        const float *fn = x.rf[k] == 0 ? v.vxn : v.vzn, *fp = x.rf[k] == 0 ? v.vxp : v.vzp;
        hipLaunchKernelGGL(rho_rec_kernel, dim3(unsigned((p.nrec + 255) / 256), unsigned(g.B)), dim3(256), 0, stream(),
                           g, grho, fn, fp, p.models[2],
                           io.adjoint_source + int64_t(k) * g.B * p.nrec * g.nt, p.receivers_loc, p.nrec, it);
    }
    SWP_LAUNCHED();
}

}  // namespace

int64_t workspace_bytes(const swp_problem &p, int kind)
{
    const Geo g = make_geo(p);
    Carver cv(nullptr, 0);
    Work w{};
    carve(cv, p, g, kind, w);
    return cv.used + 256;
}

int64_t tape_bytes(const swp_problem &p)
{
    SWP_CHECK(!(p.memory == SWP_MEM_BOUNDARY && (p.flags & SWP_FLAG_BF16_STRIPS)),
              "elastic2d: boundary saving keeps float32 strips (bf16 strips are acoustic only)");
    const Geo g = make_geo(p);
    Carver cv(nullptr, 0);
    carve_tape(cv, p, g);
    return cv.used + 256;
}

void forward(const swp_problem &p, float *record)
{
    const Ctx x = make_ctx(p);
    const Geo &g = x.g;
    Carver ws(p.workspace, p.workspace_bytes);
    Work w{};
    carve(ws, p, g, 0, w);
    derive_models(p, g, w.md);
    zero_state(g, w.s);
    Carver tc(p.tape, p.tape_bytes);
    Tape t = carve_tape(tc, p, g);
    const int64_t N = int64_t(g.B) * g.n;
    std::vector<SpinePt> spine;
    if (p.memory == SWP_MEM_BINOMIAL) spine = revolve_spine(g.nt, t.nchunk);
    size_t sp = 0;
    for (int it = 0; it < g.nt; ++it) {
        if (p.memory == SWP_MEM_CKPT && it % p.ckpt_interval == 0)
            snapshot(g, w.s, t.snaps + int64_t(it / p.ckpt_interval) * kSnap * N);
        for (; sp < spine.size() && spine[sp].step == it; ++sp)
            snapshot(g, w.s, t.snaps + int64_t(spine[sp].slot) * kSnap * N);
        forward_step(x, p, w, it, t.vel ? t.vel + int64_t(it) * 2 * N : nullptr, record);
        // the five fields' strips after the step's injection (sg_driver.cuh:324-326)
        if (t.strips) strips(g, w.s, w.s, t.strips + int64_t(it) * 5 * g.B * strip_cells(g), 0, 5, STRIP_SAVE);
    }
    if (t.last)
        for (int f = 0; f < 5; ++f) copy(t.last + f * N, w.s.f[f], N);
}

namespace {

// Boundary-saving reverse (sg_driver.cuh:740-812, elastic2d/driver_traits.cuh:
// 634-826): the forward state is rebuilt backwards from the last state, the
// reverse no-CPML steps inside the box and the saved strips at its rim; the
// reverse loop runs it = nt - 1 down to 1 (no step 0).  At step it the
// reconstruction holds the state after forward step it; with the step's
// source removed its velocities are v(it); the stresses are stepped back and
// their strips restored (from forward step it - 1) before the imaging, which
// reads v(it) and the carriers v(it + 1); after the adjoint step the
// velocities are stepped back, leaving v(it) in the carriers.
void backward_bs(const Ctx &x, const swp_problem &p, const swp_backward_io &io, Work &w, const Tape &t)
{
    const Geo &g = x.g;
    const int64_t N = int64_t(g.B) * g.n;
    const int64_t slice = int64_t(5) * g.B * strip_cells(g);
    for (int f = 0; f < 5; ++f) copy(w.s.f[f], t.last + f * N, N);
    fill_zero(w.carry, 2 * N);   // v(nt) = 0
    Fields car{};
    car.f[VX] = w.carry;
    car.f[VZ] = w.carry + N;
    const dim3 box = grid_for(g.X1 - g.X0 - 2, g.Z1 - g.Z0 - 2, g.B);
    for (int it = g.nt - 1; it >= 1; --it) {
        float *sl = t.strips + int64_t(it - 1) * slice;
        auto recon_stress = [&] {
            TimedScope ts(K_OTHER);
            for (int k = 0; k < x.nsf; ++k) {   // remove step it's source (sign -1)
                PointArgs a = point_args(g);
                a.it = it;
                a.u_add = w.s.f[x.sf[k]];
                a.add_w = p.source;
                a.add_loc = p.sources_loc;
                a.nadd = p.nsrc;
                a.sign = -1.f;
                launch_points(a, g.B);
            }
            if (g.X1 - g.X0 > 2 && g.Z1 - g.Z0 > 2)
                by_order(g.M, [&](auto ms) {
                    hipLaunchKernelGGL(stress_nopml_kernel<decltype(ms)::value>, box, dim3(64, 4), 0, stream(), g,
                                       x.c, w.s, w.md.lam, w.md.mu);
                });
            SWP_LAUNCHED();
            strips(g, w.s, w.s, sl, SXX, SXZ + 1, STRIP_RESTORE);
        };
        reverse_step(x, p, io, w, it, VelPair{w.s.f[VX], w.s.f[VZ], car.f[VX], car.f[VZ]}, recon_stress);
        TimedScope ts(K_OTHER);
        strips(g, w.s, car, sl, VX, VZ + 1, STRIP_COPY);   // v(it) at the rim into the carriers
        if (g.X1 - g.X0 > 2 && g.Z1 - g.Z0 > 2)
            by_order(g.M, [&](auto ms) {
                hipLaunchKernelGGL(vel_nopml_kernel<decltype(ms)::value>, box, dim3(64, 4), 0, stream(), g, x.c, w.s,
                                   p.models[2], car.f[VX], car.f[VZ]);
            });
        SWP_LAUNCHED();
        strips(g, w.s, w.s, sl, VX, VZ + 1, STRIP_RESTORE);
    }
}

}  // namespace

void backward(const swp_problem &p, const swp_backward_io &io)
{
    const Ctx x = make_ctx(p);
    const Geo &g = x.g;
    const int64_t N = int64_t(g.B) * g.n;
    Carver ws(p.workspace, p.workspace_bytes);
    Work w{};
    carve(ws, p, g, 1, w);
    derive_models(p, g, w.md);
    zero_state(g, w.a);
    for (int k = 0; k < 8; ++k) fill_zero(w.qp.f[k], N);
    fill_zero(w.zero, 2 * N);
    Carver tc(p.tape, p.tape_bytes);
    Tape t = carve_tape(tc, p, g);

    if (p.memory == SWP_MEM_FULL) {
        for (int it = g.nt - 1; it >= 0; --it) {
            const float *vnow = t.vel + int64_t(it) * 2 * N;
            const float *vnext = it + 1 < g.nt ? t.vel + int64_t(it + 1) * 2 * N : w.zero;
            reverse_step(x, p, io, w, it, pair_of(vnow, vnext, N));
        }
        return;
    }
    if (p.memory == SWP_MEM_BOUNDARY) {
        backward_bs(x, p, io, w, t);
        return;
    }
    if (p.memory == SWP_MEM_BINOMIAL) {
        // the forward left the spine snapshots; each reverse step re-runs its
        // forward step for v(it), v(it + 1) being the previous reverse step's
        fill_zero(w.carry, 2 * N);   // v(nt) = 0
        zero_state(g, w.s);
        for (const RevOp &o : revolve_schedule(g.nt, t.nchunk)) {
            switch (o.code) {
            case REV_STEP: forward_step(x, p, w, o.arg, nullptr, nullptr); break;
            case REV_SAVE: snapshot(g, w.s, t.snaps + int64_t(o.arg) * kSnap * N); break;
            case REV_LOAD: restore(g, w.s, t.snaps + int64_t(o.arg) * kSnap * N); break;
            default:
                forward_step(x, p, w, o.arg, w.chunk, nullptr);
                reverse_step(x, p, io, w, o.arg, pair_of(w.chunk, w.carry, N));
                std::swap(w.chunk, w.carry);
                break;
            }
        }
        return;
    }
    SWP_CHECK(p.memory == SWP_MEM_CKPT, "elastic2d backward: memory mode ", p.memory, " is not available");
    const int I = p.ckpt_interval;
    fill_zero(w.carry, 2 * N);   // v(nt) = 0 for the youngest chunk
    for (int c = t.nchunk - 1; c >= 0; --c) {
        const int it0 = c * I, it1 = std::min(g.nt, it0 + I);
        restore(g, w.s, t.snaps + int64_t(c) * kSnap * N);
        for (int it = it0; it < it1; ++it) forward_step(x, p, w, it, w.chunk + int64_t(it - it0) * 2 * N, nullptr);
        for (int it = it1 - 1; it >= it0; --it) {
            const float *vnow = w.chunk + int64_t(it - it0) * 2 * N;
            const float *vnext = it + 1 < it1 ? w.chunk + int64_t(it + 1 - it0) * 2 * N : w.carry;
            reverse_step(x, p, io, w, it, pair_of(vnow, vnext, N));
        }
        copy(w.carry, w.chunk, 2 * N);   // v(it0) is v(it1) of the next older chunk
    }
}

}  // namespace e2d
}  // namespace swp
