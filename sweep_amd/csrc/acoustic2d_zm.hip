// Wave64 z-marching kernels of the 2-D acoustic path (fixed orders 2..8).
//
// One workgroup = one wave of 64 lanes = one 64-column strip of one shot,
// marching down a segment of rows.  Each lane keeps its column's 2M+1 z-taps
// in a register queue, and the wave stages the current row (plus an M-cell
// halo each side) in its own LDS row for the x-taps.  Per cell that is one
// global load per field read and one store per field written.  The loads of
// row z+1 are issued before row z is computed (software pipelining), so each
// wave keeps a row of loads in flight while it computes; the interior
// kernels, with registers to spare, keep two (rows z+1 and z+2).  The LDS row is
// private to the wave; a workgroup barrier (one wave: no waiting on other
// waves) with its LDS fence orders each row's stores against the reads of
// the row before and after -- a wave-level scheduling barrier alone does not
// order LDS accesses, the compiler may move them across it.
//
// Blocks whose cells and taps stay out of every CPML band run the *_int
// kernels (bare leapfrog, no CPML registers, tall segments); the others run
// the *_band kernels on short segments (see make_plan).
//
// The arithmetic is the per-cell kernels' (acoustic2d.hip), which follow the
// reference's expression trees (csrc/cuda/equations/acoustic2d/
// forward_split.cu:39-168, adjoint_split.cu:27-190, kernels.cuh:472-574).
#include "acoustic2d_zm.h"
#include "fd.h"

#include <algorithm>

namespace swp {
namespace a2d {

#define WAVE_SYNC() __syncthreads()

// Block r's strip and row segment [zb, ze).  The kernels take r and the shot
// from the XCD-local numbering (xcd_block): each XCD marches a contiguous run
// of strips and segments of whole shots, so the x-halo columns and segment
// prologue rows a block re-reads were fetched into the same L2.
__device__ __forceinline__ void block_coords(const BlockSet &bs, int r, int &strip, int &zb, int &ze)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const Rect &q = bs.r[k];
        const int w = q.s1 - q.s0;
        const int nseg = q.tz > 0 ? (q.z1 - q.z0 + q.tz - 1) / q.tz : 0;
        const int cnt = w > 0 ? w * nseg : 0;
        if (r < cnt) {
            strip = q.s0 + r % w;
            zb = q.z0 + (r / w) * q.tz;
            ze = min(zb + q.tz, q.z1);
            return;
        }
        r -= cnt;
    }
    strip = 0;
    zb = ze = 0;
}

using fd::d1_taps;
using fd::lap_taps;

// The lane's halo duty: lanes 0..M-1 load the M columns left of the strip,
// lanes M..2M-1 the M columns right of it (clamped in-grid; an out-of-grid
// halo column is never read by an active lane).
template <int M>
struct Halo {
    int col, slot;
    bool on;
    __device__ Halo(int lane, int xb, int nx)
    {
        if (lane < M) { col = xb - M + lane; slot = lane; }
        else if (lane < 2 * M) { col = xb + 64 + (lane - M); slot = 64 + M + (lane - M); }
        else { col = xb; slot = -1; }
        on = slot >= 0;
        col = min(max(col, 0), nx - 1);
    }
};

// ---------------------------------------------------------------- forward
template <int M, bool UTHIS>
__global__ void __launch_bounds__(64) fwd_int_kernel(Geo g, Coef c, StepFields f,
                                                     float *__restrict__ u_this, BlockSet bs)
{
    __shared__ float row[64 + 2 * M];
    int strip, zb, ze;
    const BlockXYZ blk = xcd_block();
    block_coords(bs, blk.x, strip, zb, ze);
    if (ze <= zb) return;
    const int lane = threadIdx.x;
    const int xb = bs.x0 + strip * 64;
    const int x = xb + lane;
    const int nx = g.nx, px = g.px;
    const int xc = min(x, nx - 1);
    const bool active = x >= M && x < nx - M;
    const int64_t off = int64_t(blk.z) * g.n;
    const float *__restrict__ un = f.u_now + off;
    const float *__restrict__ up = f.u_prev + off;
    float *__restrict__ unx = f.u_next + off;
    const float *__restrict__ vp = f.vp + int64_t(blk.z) * g.mstride;
    const Halo<M> h(lane, xb, nx);
    const float dt = g.dt;

    float q[2 * M + 1];
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) q[k] = un[int64_t(zb - M + k) * px + xc];
    // Two rows of loads in flight: slots A and B alternate, each refilled with
    // row z+2 as soon as row z has taken its values, so a row's loads have two
    // rows of compute to land (the kernel is bound by load latency; one slot
    // reused by register copies would make each row wait for the next).
    struct Ahead {
        float q, h, up, v;
    };
    auto fetch = [&](int zr, Ahead &n) {   // clamped on the last row
        const int64_t rn = int64_t(min(zr, ze - 1)) * px;
        n.q = un[rn + int64_t(M) * px + xc];
        n.h = un[rn + h.col];
        n.up = up[rn + xc];
        n.v = vp[rn + xc];
    };
    auto step = [&](int z, Ahead &n) {
        const int64_t rz = int64_t(z) * px;
        q[2 * M] = n.q;
        const float hv = n.h, upv = n.up, v = n.v;
        fetch(z + 2, n);
        WAVE_SYNC();
        row[M + lane] = q[M];
        if (h.on) row[h.slot] = hv;
        WAVE_SYNC();
        const float lap_x = lap_taps<M>(c, [&](int k) { return row[M + lane + k]; }, g.inv_dx2);
        const float lap_z = lap_taps<M>(c, [&](int k) { return q[M + k]; }, g.inv_dz2);
        const float v2_dt2 = (v * v) * dt * dt;
        if (active) {
            unx[rz + x] = 2.0f * q[M] - upv + v2_dt2 * (lap_x + lap_z);
            if constexpr (UTHIS) u_this[off + rz + x] = (v * v) * (lap_x + lap_z);
        }
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) q[k] = q[k + 1];
    };
    Ahead A, B;
    fetch(zb, A);
    fetch(zb + 1, B);
    for (int z = zb; z < ze; z += 2) {
        step(z, A);
        if (z + 1 < ze) step(z + 1, B);
    }
}

template <int M, bool UTHIS, bool STRIP>
__global__ void __launch_bounds__(64) fwd_band_kernel(Geo g, Coef c, Prof pr, StepFields f,
                                                      float *__restrict__ u_this,
                                                      float *__restrict__ strip_out, BlockSet bs)
{
    __shared__ float row[64 + 2 * M];
    __shared__ float prow[64 + 2 * M];
    int strip, zb, ze;
    const BlockXYZ blk = xcd_block();
    block_coords(bs, blk.x, strip, zb, ze);
    if (ze <= zb) return;
    const int lane = threadIdx.x;
    const int xb = bs.x0 + strip * 64;
    const int x = xb + lane;
    const int nx = g.nx, px = g.px, nz = g.nz;
    const int xc = min(x, nx - 1);
    const bool active = x >= M && x < nx - M;
    const int b = blk.z;
    const int64_t off = int64_t(b) * g.n;
    const float *__restrict__ un = f.u_now + off;
    const float *__restrict__ up = f.u_prev + off;
    float *__restrict__ unx = f.u_next + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const float *__restrict__ psix = f.psix + off;
    const float *__restrict__ psiz = f.psiz + off;
    float *__restrict__ psixn = f.psixn + off;
    float *__restrict__ psizn = f.psizn + off;
    float *__restrict__ zetax = f.zetax + off;
    float *__restrict__ zetaz = f.zetaz + off;
    const Halo<M> h(lane, xb, nx);
    const float dt = g.dt;

    // x band of this lane, and whether the wave has any x-band lane
    const int xlo = g.pxl + 2 * M, xhi = nx - g.pxh - 2 * M;
    const bool cx = x < xlo || x >= xhi;
    const bool wave_x = xb < xlo || xb + 63 >= xhi;
    const int zlo = g.pzl + 2 * M, zhi = nz - g.pzh - 2 * M;
    const bool seg_z = zb < zlo || ze > zhi;
    float ax_ = 0.f, bx_ = 0.f, dbx_ = 0.f;
    if (wave_x) { ax_ = pr.ax[xc]; bx_ = pr.bx[xc]; dbx_ = pr.dbx[xc]; }
    const float hax = wave_x ? pr.ax[h.col] : 0.f;

    float q[2 * M + 1];
    float pq[2 * M + 1];   // az * psiz, z-band segments only
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) {
        const int zr = zb - M + k;
        q[k] = un[int64_t(zr) * px + xc];
        pq[k] = seg_z ? pr.az[zr] * psiz[int64_t(zr) * px + xc] : 0.f;
    }
    float *s_out = STRIP ? strip_at(strip_out, int64_t(b) * g.ns, g.sbf16) : nullptr;

    float n_q, n_pq = 0.f, n_h, n_hp = 0.f, n_px = 0.f, n_up, n_v, n_zx = 0.f, n_zz = 0.f, n_pz = 0.f;
    auto issue = [&](int zr) {
        const int64_t rn = int64_t(zr) * px;
        n_q = un[rn + int64_t(M) * px + xc];
        if (seg_z) {
            n_pq = pr.az[zr + M] * psiz[rn + int64_t(M) * px + xc];
            n_zz = zetaz[rn + xc];
            n_pz = psiz[rn + xc];
        }
        n_h = un[rn + h.col];
        if (wave_x) {
            n_hp = hax * psix[rn + h.col];
            n_px = psix[rn + xc];
            n_zx = zetax[rn + xc];
        }
        n_up = up[rn + xc];
        n_v = vp[rn + xc];
    };
    issue(zb);
    for (int z = zb; z < ze; ++z) {
        const int64_t rz = int64_t(z) * px;
        q[2 * M] = n_q;
        pq[2 * M] = n_pq;
        const float hv = n_h, hp = n_hp, pxv = n_px, upv = n_up, v = n_v, zxc = n_zx, zzc = n_zz, pzc = n_pz;
        issue(min(z + 1, ze - 1));
        WAVE_SYNC();
        row[M + lane] = q[M];
        if (h.on) row[h.slot] = hv;
        if (wave_x) {
            prow[M + lane] = ax_ * pxv;
            if (h.on) prow[h.slot] = hp;
        }
        WAVE_SYNC();
        if (STRIP && active) {   // boundary save of u_now (strip.cuh)
            const float u0 = q[M];
            const bool inx = x >= g.X0 && x < g.X1, inz = z >= g.Z0 && z < g.Z1;
            const bool bf = g.sbf16;
            if (inx && z < g.Z0 + g.W && z >= g.Z0) strip_put(s_out, int64_t(z - g.Z0) * g.nxp + (x - g.X0), u0, bf);
            if (inx && z >= g.Z1 - g.W && z < g.Z1)
                strip_put(s_out, int64_t(g.W) * g.nxp + int64_t(z - (g.Z1 - g.W)) * g.nxp + (x - g.X0), u0, bf);
            if (inz && x >= g.X0 && x < g.X0 + g.W)
                strip_put(s_out, 2LL * g.W * g.nxp + int64_t(z - g.Z0) * g.W + (x - g.X0), u0, bf);
            if (inz && x >= g.X1 - g.W && x < g.X1)
                strip_put(s_out, 2LL * g.W * g.nxp + int64_t(g.nzp) * g.W + int64_t(z - g.Z0) * g.W + (x - (g.X1 - g.W)),
                          u0, bf);
        }
        const float lap_x = lap_taps<M>(c, [&](int k) { return row[M + lane + k]; }, g.inv_dx2);
        const float lap_z = lap_taps<M>(c, [&](int k) { return q[M + k]; }, g.inv_dz2);
        const float v2_dt2 = (v * v) * dt * dt;
        const bool cz = z < zlo || z >= zhi;
        float w_sum;
        if (!cx && !cz) {
            w_sum = lap_x + lap_z;
        } else {
            w_sum = 0.0f;
            if (cx) {
                const float dudx = d1_taps<M>(c, [&](int k) { return row[M + lane + k]; }, g.inv_dx);
                const float dap = d1_taps<M>(c, [&](int k) { return prow[M + lane + k]; }, g.inv_dx);
                const float tmpx = ((1.0f + bx_) * lap_x + dbx_ * dudx) + dap;
                const float zx = zxc;
                w_sum += (1.0f + bx_) * tmpx + ax_ * zx;
                if (active) {
                    psixn[rz + x] = bx_ * dudx + ax_ * pxv;
                    zetax[rz + x] = bx_ * tmpx + ax_ * zx;
                }
            } else {
                w_sum += lap_x;
            }
            if (cz) {
                const float az_ = pr.az[z], bz_ = pr.bz[z], dbz_ = pr.dbz[z];
                const float dudz = d1_taps<M>(c, [&](int k) { return q[M + k]; }, g.inv_dz);
                const float dap = d1_taps<M>(c, [&](int k) { return pq[M + k]; }, g.inv_dz);
                const float tmpz = ((1.0f + bz_) * lap_z + dbz_ * dudz) + dap;
                const float zz = zzc;
                w_sum += (1.0f + bz_) * tmpz + az_ * zz;
                if (active) {
                    psizn[rz + x] = bz_ * dudz + az_ * pzc;
                    zetaz[rz + x] = bz_ * tmpz + az_ * zz;
                }
            } else {
                w_sum += lap_z;
            }
        }
        if (active) {
            unx[rz + x] = 2.0f * q[M] - upv + v2_dt2 * w_sum;
            if constexpr (UTHIS) u_this[off + rz + x] = (v * v) * w_sum;
        }
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) { q[k] = q[k + 1]; pq[k] = pq[k + 1]; }
    }
}

// ---------------------------------------------------------------- adjoint
template <int M, bool IMG>
__global__ void __launch_bounds__(64) adj_int_kernel(Geo g, Coef c, AdjFields f,
                                                     const float *__restrict__ img,
                                                     float *__restrict__ grad, BlockSet bs)
{
    __shared__ float row[64 + 2 * M];
    int strip, zb, ze;
    const BlockXYZ blk = xcd_block();
    block_coords(bs, blk.x, strip, zb, ze);
    if (ze <= zb) return;
    const int lane = threadIdx.x;
    const int xb = bs.x0 + strip * 64;
    const int x = xb + lane;
    const int nx = g.nx, px = g.px;
    const int xc = min(x, nx - 1);
    const bool active = x >= M && x < nx - M;
    const int b = blk.z;
    const int64_t off = int64_t(b) * g.n;
    const float *__restrict__ ln = f.l_now + off;
    const float *__restrict__ lp = f.l_prev + off;
    float *__restrict__ lnx = f.l_next + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const Halo<M> h(lane, xb, nx);
    const float dt2 = g.dt2;

    float q[2 * M + 1];   // gw = (v^2 dt^2) lambda
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) {
        const int64_t r = int64_t(zb - M + k) * px + xc;
        const float v = vp[r];
        q[k] = v * v * dt2 * ln[r];
    }
    // Rows in flight, two slots as in fwd_int_kernel: row z+M of the queue
    // fields, row z's halo, lambda_prev and lambda_now (the centre, used
    // unscaled), and with IMG the imaging operands and the gradient (each
    // cell's gradient is updated by its own lane only).
    struct Ahead {
        float v, l, hv, hl, lp, gun, im, v0, gr;
    };
    auto fetch = [&](int zr, Ahead &n) {   // clamped on the last row
        const int64_t rn = int64_t(min(zr, ze - 1)) * px;
        n.v = vp[rn + int64_t(M) * px + xc];
        n.l = ln[rn + int64_t(M) * px + xc];
        n.hv = vp[rn + h.col];
        n.hl = ln[rn + h.col];
        n.lp = lp[rn + xc];
        n.gun = ln[rn + xc];
        if constexpr (IMG) {
            n.im = img[off + rn + xc];
            n.v0 = vp[rn + xc];
            n.gr = grad[off + rn + xc];
        }
    };
    auto step = [&](int z, Ahead &n) {
        const int64_t rz = int64_t(z) * px;
        q[2 * M] = n.v * n.v * dt2 * n.l;
        const float hv = n.hv * n.hv * dt2 * n.hl;
        const float lpv = n.lp, gun = n.gun, im = n.im, v0 = n.v0, gr = n.gr;
        fetch(z + 2, n);
        WAVE_SYNC();
        row[M + lane] = q[M];
        if (h.on) row[h.slot] = hv;
        WAVE_SYNC();
        const float gw = q[M];
        float lapx = 0.f, lapz = 0.f;
#pragma unroll
        for (int k = 1; k <= M; ++k) {
            lapx += c.lap[k] * ((row[M + lane + k] - gw) + (row[M + lane - k] - gw));
            lapz += c.lap[k] * ((q[M + k] - gw) + (q[M - k] - gw));
        }
        if (active) {
            if constexpr (IMG) grad[off + rz + x] = gr + 2.f * g.dt * g.dt * im * gun / v0;
            lnx[rz + x] = 2.0f * gun - lpv + lapx * g.inv_dx2.inv + lapz * g.inv_dz2.inv;
        }
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) q[k] = q[k + 1];
    };
    Ahead A{}, B{};
    fetch(zb, A);
    fetch(zb + 1, B);
    for (int z = zb; z < ze; z += 2) {
        step(z, A);
        if (z + 1 < ze) step(z + 1, B);
    }
}

template <int M, bool IMG>
__global__ void __launch_bounds__(64) adj_band_kernel(Geo g, Coef c, Prof pr, AdjFields f,
                                                      const float *__restrict__ img,
                                                      float *__restrict__ grad, BlockSet bs)
{
    // LDS rows of the current row: gw, and the x-axis tap quantities
    //   T = b zeta + ((1+b) c) lambda,  A = (1+b) T,  G = b psi + db T
    __shared__ float rg[64 + 2 * M], rT[64 + 2 * M], rA[64 + 2 * M], rG[64 + 2 * M];
    int strip, zb, ze;
    const BlockXYZ blk = xcd_block();
    block_coords(bs, blk.x, strip, zb, ze);
    if (ze <= zb) return;
    const int lane = threadIdx.x;
    const int xb = bs.x0 + strip * 64;
    const int x = xb + lane;
    const int nx = g.nx, px = g.px;
    const int xc = min(x, nx - 1);
    const bool active = x >= M && x < nx - M;
    const int b = blk.z;
    const int64_t off = int64_t(b) * g.n;
    const float *__restrict__ ln = f.l_now + off;
    const float *__restrict__ lp = f.l_prev + off;
    float *__restrict__ lnx = f.l_next + off;
    const float *__restrict__ vp = f.vp + int64_t(b) * g.mstride;
    const float *__restrict__ psix = f.psix + off;
    const float *__restrict__ psiz = f.psiz + off;
    const float *__restrict__ zetax = f.zetax + off;
    const float *__restrict__ zetaz = f.zetaz + off;
    const Halo<M> h(lane, xb, nx);
    const float dt2 = g.dt2;

    // pure interior (adjoint_split.cu:265-270): taps out of every band and off
    // the db leak cell
    const int px0 = g.X0 + M + 1, px1 = g.X1 - M - 1, pz0 = g.Z0 + M + 1, pz1 = g.Z1 - M - 1;
    const bool xin = x >= px0 && x < px1;
    const bool wave_x = xb < px0 || xb + 63 >= px1;
    const bool seg_z = zb < pz0 || ze > pz1;
    float bx_ = 0.f, dbx_ = 0.f, ax_ = 0.f, hb = 0.f, hdb = 0.f;
    if (wave_x) {
        ax_ = pr.ax[xc]; bx_ = pr.bx[xc]; dbx_ = pr.dbx[xc];
        hb = pr.bx[h.col]; hdb = pr.dbx[h.col];
    }

    float qg[2 * M + 1], qT[2 * M + 1], qA[2 * M + 1], qG[2 * M + 1];
    auto put_row = [&](int zr, int k, float v, float l, float zz, float pz) {
        const float cc = v * v * dt2;
        qg[k] = cc * l;
        if (seg_z) {
            const float bz = pr.bz[zr], dbz = pr.dbz[zr];
            const float T = bz * zz + (1.0f + bz) * cc * l;
            qT[k] = T;
            qA[k] = (1.0f + bz) * T;
            qG[k] = bz * pz + dbz * T;
        }
    };
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) {
        const int zr = zb - M + k;
        const int64_t r = int64_t(zr) * px + xc;
        put_row(zr, k, vp[r], ln[r], seg_z ? zetaz[r] : 0.f, seg_z ? psiz[r] : 0.f);
    }
    // row z+M (queue tail) and row z (halo, own x quantities) in flight
    float n_v, n_l, n_zz = 0.f, n_pz = 0.f, n_hv, n_hl, n_hz = 0.f, n_hp = 0.f;
    float n_gun, n_lp, n_v0, n_zx = 0.f, n_px = 0.f, n_im = 0.f, n_zzc = 0.f, n_pzc = 0.f, n_gr = 0.f;
    auto issue = [&](int zr) {
        const int64_t rq = int64_t(zr + M) * px + xc;
        n_v = vp[rq];
        n_l = ln[rq];
        if (seg_z) { n_zz = zetaz[rq]; n_pz = psiz[rq]; }
        const int64_t rn = int64_t(zr) * px;
        n_hv = vp[rn + h.col];
        n_hl = ln[rn + h.col];
        if (wave_x) { n_hz = zetax[rn + h.col]; n_hp = psix[rn + h.col]; n_zx = zetax[rn + xc]; n_px = psix[rn + xc]; }
        n_gun = ln[rn + xc];
        n_lp = lp[rn + xc];
        n_v0 = vp[rn + xc];
        if constexpr (IMG) {   // the gradient one row ahead (own lane only)
            n_im = img[off + rn + xc];
            n_gr = grad[off + rn + xc];
        }
    };
    issue(zb);
    for (int z = zb; z < ze; ++z) {
        const int64_t rz = int64_t(z) * px;
        put_row(z + M, 2 * M, n_v, n_l, n_zz, n_pz);
        const float hvv = n_hv, hl = n_hl, hz = n_hz, hpx = n_hp;
        const float gun = n_gun, lpv = n_lp, v0 = n_v0, zx = n_zx, px = n_px, im = n_im;
        const float zzc = n_zzc, pzc = n_pzc, gr = n_gr;
        issue(min(z + 1, ze - 1));
        const float hcc = hvv * hvv * dt2;
        const float hg = hcc * hl;
        const float cc0 = v0 * v0 * dt2;
        WAVE_SYNC();
        rg[M + lane] = qg[M];
        if (h.on) rg[h.slot] = hg;
        if (wave_x) {
            const float Tx = bx_ * zx + (1.0f + bx_) * cc0 * gun;
            rT[M + lane] = Tx;
            rA[M + lane] = (1.0f + bx_) * Tx;
            rG[M + lane] = bx_ * px + dbx_ * Tx;
            if (h.on) {
                const float hT = hb * hz + (1.0f + hb) * hcc * hl;
                rT[h.slot] = hT;
                rA[h.slot] = (1.0f + hb) * hT;
                rG[h.slot] = hb * hpx + hdb * hT;
            }
        }
        WAVE_SYNC();
        if (IMG && active && x >= g.X0 && x < g.X1 && z >= g.Z0 && z < g.Z1)
            grad[off + rz + x] = gr + 2.f * g.dt * g.dt * im * gun / v0;
        const float gw = cc0 * gun;
        const bool zin = z >= pz0 && z < pz1;
        if (xin && zin) {
            float lapx = 0.f, lapz = 0.f;
#pragma unroll
            for (int k = 1; k <= M; ++k) {
                lapx += c.lap[k] * ((rg[M + lane + k] - gw) + (rg[M + lane - k] - gw));
                lapz += c.lap[k] * ((qg[M + k] - gw) + (qg[M - k] - gw));
            }
            if (active) lnx[rz + x] = 2.0f * gun - lpv + lapx * g.inv_dx2.inv + lapz * g.inv_dz2.inv;
        } else {
            float Lx = 0.f, Dgx = 0.f, Dqx = 0.f, Lz = 0.f, Dgz = 0.f, Dqz = 0.f;
            if (wave_x) {
                const float lx0 = (1.0f + bx_) * (bx_ * zx + (1.0f + bx_) * gw);
#pragma unroll
                for (int k = 1; k <= M; ++k) {
                    Lx += c.lap[k] * ((rA[M + lane + k] - lx0) + (rA[M + lane - k] - lx0));
                    Dgx += c.grad[k] * (rG[M + lane + k] - rG[M + lane - k]);
                    Dqx += c.grad[k] * (rT[M + lane + k] - rT[M + lane - k]);
                }
            } else {
#pragma unroll
                for (int k = 1; k <= M; ++k) {
                    Lx += c.lap[k] * ((rg[M + lane + k] - gw) + (rg[M + lane - k] - gw));
                    Dqx += c.grad[k] * (rg[M + lane + k] - rg[M + lane - k]);
                }
            }
            float zz = 0.f, pzv = 0.f;
            if (seg_z) {
                zz = zetaz[rz + xc];
                pzv = psiz[rz + xc];
                const float bz_ = pr.bz[z];
                const float lz0 = (1.0f + bz_) * (bz_ * zz + (1.0f + bz_) * gw);
#pragma unroll
                for (int k = 1; k <= M; ++k) {
                    Lz += c.lap[k] * ((qA[M + k] - lz0) + (qA[M - k] - lz0));
                    Dgz += c.grad[k] * (qG[M + k] - qG[M - k]);
                    Dqz += c.grad[k] * (qT[M + k] - qT[M - k]);
                }
            } else {
#pragma unroll
                for (int k = 1; k <= M; ++k) {
                    Lz += c.lap[k] * ((qg[M + k] - gw) + (qg[M - k] - gw));
                    Dqz += c.grad[k] * (qg[M + k] - qg[M - k]);
                }
            }
            Lx *= g.inv_dx2.inv; Dgx *= g.inv_dx.inv; Dqx *= g.inv_dx.inv;
            Lz *= g.inv_dz2.inv; Dgz *= g.inv_dz.inv; Dqz *= g.inv_dz.inv;
            if (active) {
                lnx[rz + x] = 2.0f * gun - lpv + (Lx + Lz - Dgx - Dgz);
                if (wave_x) {
                    f.psixo[off + rz + x] = ax_ * (px - Dqx);
                    f.zetaxo[off + rz + x] = ax_ * (zx + gw);
                }
                if (seg_z) {
                    const float az_ = pr.az[z];
                    f.psizo[off + rz + x] = az_ * (pzv - Dqz);
                    f.zetazo[off + rz + x] = az_ * (zz + gw);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) {
            qg[k] = qg[k + 1];
            qT[k] = qT[k + 1];
            qA[k] = qA[k + 1];
            qG[k] = qG[k + 1];
        }
    }
}

// ---------------------------------------------------------------- BS recon
// Reverse leapfrog without PML over [x_lo, x_hi) x [z_lo, z_hi), fused u_tt
// imaging inside the strips' inner edge.  Strips of 64 columns start at x_lo rounded
// down to a 128-byte line; lanes outside [x_lo, x_hi) only stage their column.
template <int M, bool IMG>
__global__ void __launch_bounds__(64) recon_zm_kernel(Geo g, Coef c, const float *__restrict__ vpb,
                                                      const float *__restrict__ u_prev,
                                                      const float *__restrict__ u_now,
                                                      float *__restrict__ u_next,
                                                      const float *__restrict__ adj,
                                                      float *__restrict__ grad, int x_lo, int x_hi,
                                                      int z_lo, int z_hi, int tz)
{
    __shared__ float row[64 + 2 * M];
    const int lane = threadIdx.x;
    const BlockXYZ blk = xcd_block();
    const int xb = x_lo - x_lo % 32 + blk.x * 64;   // strips on 128-byte lines
    const int x = xb + lane;
    const int nx = g.nx, px = g.px;
    const int xc = min(x, nx - 1);
    const bool active = x >= x_lo && x < x_hi;
    const int zb = z_lo + blk.y * tz;
    const int ze = min(zb + tz, z_hi);
    const int b = blk.z;
    const int64_t off = int64_t(b) * g.n;
    const float *__restrict__ un = u_now + off;
    const float *__restrict__ up = u_prev + off;
    const float *__restrict__ vp = vpb + int64_t(b) * g.mstride;
    const Halo<M> h(lane, xb, nx);
    const bool ximg = x >= g.X0 + g.W && x < g.X1 - g.W;
    const float dt = g.dt;

    float q[2 * M + 1];
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) q[k] = un[int64_t(zb - M + k) * px + xc];
    // the gradient is read with the row's other fields, one row ahead: each
    // cell's gradient is updated by its own lane only
    float n_q, n_h, n_up, n_v, n_a = 0.f, n_gr = 0.f;
    auto issue = [&](int zr) {
        const int64_t rn = int64_t(zr) * px;
        n_q = un[rn + int64_t(M) * px + xc];
        n_h = un[rn + h.col];
        n_up = up[rn + xc];
        n_v = vp[rn + xc];
        if (IMG) {
            n_a = adj[off + rn + xc];
            n_gr = grad[off + rn + xc];
        }
    };
    issue(zb);
    for (int z = zb; z < ze; ++z) {
        const int64_t rz = int64_t(z) * px;
        q[2 * M] = n_q;
        const float hv = n_h, upv = n_up, v = n_v, a = n_a, gr = n_gr;
        issue(min(z + 1, ze - 1));
        const bool img_row = IMG && z >= g.Z0 + g.W && z < g.Z1 - g.W;
        WAVE_SYNC();
        row[M + lane] = q[M];
        if (h.on) row[h.slot] = hv;
        WAVE_SYNC();
        const float w_sum = lap_taps<M>(c, [&](int k) { return row[M + lane + k]; }, g.inv_dx2) +
                            lap_taps<M>(c, [&](int k) { return q[M + k]; }, g.inv_dz2);
        const float u0 = q[M];
        const float unx = 2.0f * u0 - upv + (v * v) * dt * dt * w_sum;
        if (active) {
            u_next[off + rz + x] = unx;
            if (img_row && ximg) {
                const float u_tt = fd::scaled(unx - 2 * u0 + upv, g.inv_dt2);
                grad[off + rz + x] = gr + 2.f * dt * dt * u_tt * a / v;
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) q[k] = q[k + 1];
    }
}

// ---------------------------------------------------------------- launches
static int rect_blocks(const Rect &r)
{
    if (r.s1 <= r.s0 || r.z1 <= r.z0 || r.tz <= 0) return 0;
    return (r.s1 - r.s0) * ((r.z1 - r.z0 + r.tz - 1) / r.tz);
}

ZmPlan make_plan(const Geo &g, int tz, int tz_band)
{
    ZmPlan p{};
    const int M = g.M;
    p.tz = tz;
    p.tz_band = tz_band;
    const int x0 = 0;   // strips start on 128-byte lines (rows are line-aligned)
    p.nstrip = (g.nx - M - x0 + 63) / 64;
    p.inner.x0 = p.band.x0 = x0;
    const int zmin = M, zmax = g.nz - M;
    // interior: strips and rows inside [lo, hi) with lo = pad + 2M + 1 on
    // every face (no CPML term, no strip save, no db leak cell reaches them)
    const int xlo = std::max(g.pxl + 2 * M + 1, g.X0 + g.W), xhi = std::min(g.nx - g.pxh - 2 * M - 1, g.X1 - g.W);
    const int zlo = std::max(g.pzl + 2 * M + 1, g.Z0 + g.W), zhi = std::min(g.nz - g.pzh - 2 * M - 1, g.Z1 - g.W);
    int s0 = 0, s1 = p.nstrip;
    while (s0 < p.nstrip && x0 + s0 * 64 < xlo) ++s0;
    while (s1 > s0 && x0 + s1 * 64 > xhi) --s1;
    const int zi0 = std::max(zlo, zmin);
    const int zi1 = zi0 + std::max(0, (zhi - zi0) / tz) * tz;   // whole interior segments
    if (s1 <= s0 || zi1 <= zi0) {
        p.band.r[0] = Rect{0, p.nstrip, zmin, zmax, tz_band};
        p.nband = rect_blocks(p.band.r[0]);
        p.ninner = 0;
        return p;
    }
    p.inner.r[0] = Rect{s0, s1, zi0, zi1, tz};
    p.ninner = rect_blocks(p.inner.r[0]);
    p.band.r[0] = Rect{0, p.nstrip, zmin, zi0, tz_band};
    p.band.r[1] = Rect{0, p.nstrip, zi1, zmax, tz_band};
    p.band.r[2] = Rect{0, s0, zi0, zi1, tz_band};
    p.band.r[3] = Rect{s1, p.nstrip, zi0, zi1, tz_band};
    p.nband = 0;
    for (auto &r : p.band.r) p.nband += rect_blocks(r);
    return p;
}

bool zm_supported(const Geo &g) { return g.M >= 1 && g.M <= 4; }

template <class K>
static void by_m(int M, K &&k)
{
    switch (M) {
    case 1: k(std::integral_constant<int, 1>{}); break;
    case 2: k(std::integral_constant<int, 2>{}); break;
    case 3: k(std::integral_constant<int, 3>{}); break;
    case 4: k(std::integral_constant<int, 4>{}); break;
    default: fail("z-march kernels take M in 1..4, got ", M);
    }
}

void launch_fwd_zm(const Ctx &x, const ZmPlan &p, const StepFields &f, float *u_this, float *strip)
{
    const Geo &g = x.g;
    by_m(g.M, [&](auto mc) {
        constexpr int M = decltype(mc)::value;
        if (p.nband) {
            const dim3 grid(unsigned(p.nband), 1, unsigned(g.B));
            census_add("a2d.fwd_band", int64_t(p.nband) * g.B);
            if (u_this) {
                if (strip) hipLaunchKernelGGL((fwd_band_kernel<M, true, true>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, u_this, strip, p.band);
                else hipLaunchKernelGGL((fwd_band_kernel<M, true, false>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, u_this, strip, p.band);
            } else {
                if (strip) hipLaunchKernelGGL((fwd_band_kernel<M, false, true>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, u_this, strip, p.band);
                else hipLaunchKernelGGL((fwd_band_kernel<M, false, false>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, u_this, strip, p.band);
            }
        }
        if (p.ninner) {
            const dim3 grid(unsigned(p.ninner), 1, unsigned(g.B));
            census_add("a2d.fwd_int", int64_t(p.ninner) * g.B);
            if (u_this) hipLaunchKernelGGL((fwd_int_kernel<M, true>), grid, dim3(64), 0, stream(), g, x.c, f, u_this, p.inner);
            else hipLaunchKernelGGL((fwd_int_kernel<M, false>), grid, dim3(64), 0, stream(), g, x.c, f, u_this, p.inner);
        }
    });
    SWP_LAUNCHED();
}

void launch_adj_zm(const Ctx &x, const ZmPlan &p, const AdjFields &f, const float *img, float *grad)
{
    const Geo &g = x.g;
    const bool im = img != nullptr && grad != nullptr;
    by_m(g.M, [&](auto mc) {
        constexpr int M = decltype(mc)::value;
        if (p.nband) {
            const dim3 grid(unsigned(p.nband), 1, unsigned(g.B));
            census_add("a2d.adj_band", int64_t(p.nband) * g.B);
            if (im) hipLaunchKernelGGL((adj_band_kernel<M, true>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, img, grad, p.band);
            else hipLaunchKernelGGL((adj_band_kernel<M, false>), grid, dim3(64), 0, stream(), g, x.c, x.pr, f, img, grad, p.band);
        }
        if (p.ninner) {
            const dim3 grid(unsigned(p.ninner), 1, unsigned(g.B));
            census_add("a2d.adj_int", int64_t(p.ninner) * g.B);
            if (im) hipLaunchKernelGGL((adj_int_kernel<M, true>), grid, dim3(64), 0, stream(), g, x.c, f, img, grad, p.inner);
            else hipLaunchKernelGGL((adj_int_kernel<M, false>), grid, dim3(64), 0, stream(), g, x.c, f, img, grad, p.inner);
        }
    });
    SWP_LAUNCHED();
}

void launch_recon_zm(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                     const float *adj, float *grad, int x_lo, int x_hi, int z_lo, int z_hi, int tz)
{
    const Geo &g = x.g;
    const dim3 grid(unsigned((x_hi - (x_lo - x_lo % 32) + 63) / 64), unsigned((z_hi - z_lo + tz - 1) / tz),
                    unsigned(g.B));
    census_add("a2d.recon_zm", int64_t(grid.x) * grid.y * grid.z);
    by_m(g.M, [&](auto mc) {
        constexpr int M = decltype(mc)::value;
        if (grad) hipLaunchKernelGGL((recon_zm_kernel<M, true>), grid, dim3(64), 0, stream(), g, x.c, vp, u_prev, u_now, u_next, adj, grad, x_lo, x_hi, z_lo, z_hi, tz);
        else hipLaunchKernelGGL((recon_zm_kernel<M, false>), grid, dim3(64), 0, stream(), g, x.c, vp, u_prev, u_now, u_next, adj, grad, x_lo, x_hi, z_lo, z_hi, tz);
    });
    SWP_LAUNCHED();
}

}  // namespace a2d
}  // namespace swp
