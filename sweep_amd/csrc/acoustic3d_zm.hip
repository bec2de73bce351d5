// 2.5-D z-marching kernels of the 3-D acoustic path (fixed orders 2..8).
//
// A workgroup of 256 lanes (4 wave64s) owns a 64 (x) by TY (y) tile of
// columns and marches down a segment of z planes.  Each lane keeps, for each
// of its RY = TY/4 rows, the column's 2M+1 z-taps in a register queue; the
// current (x, y) plane of the tile plus an M-cell halo on each side sits in
// LDS for the x- and y-taps.  The LDS plane is double-buffered, so one
// workgroup barrier per plane orders its writes against the reads of the
// plane before.  The loads of plane z+1 (queue head, halos, the other
// fields) are issued before plane z is computed.  Per cell that is one HBM
// read per field read and one write per field written; halo cells come from
// L2.
//
// The computed box is cut into a lean interior -- [lo + M + 1, hi - M - 1)
// per axis: no forward CPML term, no strip cell, no adjoint CPML term (the
// adjoint's pure region) -- run by the *_int kernels, and six slabs around
// it run by the *_shell kernels, which evaluate each cell's CPML bands.
// Block indices are remapped so that each XCD (workgroups go round-robin
// over the eight) takes a contiguous run of tiles, whose halos then meet in
// the same L2.
//
// The arithmetic is that of the per-cell kernels (acoustic3d.hip), which
// follow the reference's expression trees (csrc/cuda/equations/acoustic3d/
// forward_split.cu:81-328, adjoint_split.cu:56-318).
#include "acoustic3d_zm.h"
#include "fd.h"

#include <algorithm>

namespace swp {
namespace a3d {

using fd::d1_taps;
using fd::lap_taps;

namespace {

constexpr int kRectMax = 8;

// A box of cells [x0, x1) x [y0, y1) x [z0, z1) cut into tiles of 64 x TY
// columns (partial tiles masked) and z segments of tz planes.
// Tiles start at xs <= x0 (x0 rounded down to a 128-byte line when rows are
// line-aligned), so each wave row covers whole lines; lanes outside [x0, x1)
// are masked.
struct ZRect {
    int x0, x1, y0, y1, z0, z1, tz;
    int xs;
    int ntx, nty, nseg, nblk;
};

struct ZSet {
    ZRect r[kRectMax];
    int n;
    int per_shot;   // blocks per shot
};

// Calls f(integral_constant<int, I>) for I = B..N-1 in order while it returns true.
template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        if (f(std::integral_constant<int, I>{})) static_for<I + 1, N>(f);
    }
}

// One shot's field seen from plane zbase on through a raw buffer resource
// (base and extent in SGPRs).  A load or store takes the lane's in-plane byte
// offset (a loop-invariant VGPR) and the plane's byte offset (an SGPR), so a
// plane step costs one scalar multiply instead of a 64-bit address per lane.
// Accesses past the field's end read 0 and write nothing.  The caller keeps
// (planes used) * sxy * 4 below 2^32.
struct Fld {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Fld(const float *field, int64_t base_cell, int64_t end_cell)
    {
        const int64_t bytes = (end_cell - base_cell) * 4;
        r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(field + base_cell), short(0),
                                              int(uint32_t(bytes < 0xffffffffLL ? bytes : 0xffffffffLL)),
                                              0x00020000);   // gfx9 dword 3: 32-bit raw data
    }
    __device__ __forceinline__ float ld(uint32_t voff, uint32_t soff) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    __device__ __forceinline__ void st(float v, uint32_t voff, uint32_t soff) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
    }
    // Streaming forms (non-temporal cache policy) for data no other lane of
    // the step reads again, so it does not push the shared halo planes out of L2.
    __device__ __forceinline__ float ld_nt(uint32_t voff, uint32_t soff) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    __device__ __forceinline__ void st_nt(float v, uint32_t voff, uint32_t soff) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
    }
};

// This block's shot, tile origin and z segment; false for a block with no
// work (never happens for a well-formed set).
template <int TY>
__device__ __forceinline__ bool zcoords(const ZSet &s, int &b, int &k, int &xo, int &yo, int &zb, int &ze)
{
    const int total = int(gridDim.x);
    const int f = int(xcd_local_id(blockIdx.x, unsigned(total)));
    b = f / s.per_shot;
    int t = f % s.per_shot;
    for (k = 0; k < s.n; ++k) {
        const ZRect &R = s.r[k];
        if (t < R.nblk) {
            const int tx = t % R.ntx;
            const int ty = (t / R.ntx) % R.nty;
            const int sg = t / (R.ntx * R.nty);
            xo = R.xs + 64 * tx;
            yo = R.y0 + TY * ty;
            zb = R.z0 + sg * R.tz;
            ze = min(zb + R.tz, R.z1);
            return true;
        }
        t -= R.nblk;
    }
    return false;
}

// ---------------------------------------------------------------- plane I/O
// The halo duties of a lane: y-halo rows j = w + 4t (t < YH) of the 2M rows
// above and below the tile, x-halo columns for lanes < 2M on each of its rows.
template <int M, int TY, int NW>
struct PlaneMap {
    static constexpr int LW = 64 + 2 * M;
    static constexpr int LH = TY + 2 * M;
    static constexpr int YH = (2 * M + NW - 1) / NW;
    int yh_row[YH];      // global row of each y-halo duty (clamped), or -1
    int yh_slot[YH];     // LDS row of each duty
    int xh_col, xh_slot; // x-halo column (clamped) and LDS column, xh_slot < 0: none
    __device__ PlaneMap(int lane, int w, int xo, int yo, int nx, int ny)
    {
#pragma unroll
        for (int t = 0; t < YH; ++t) {
            const int j = w + NW * t;
            if (j < 2 * M) {
                const int gy = j < M ? yo - M + j : yo + TY + (j - M);
                yh_row[t] = min(max(gy, 0), ny - 1);
                yh_slot[t] = j < M ? j : TY + M + (j - M);
            } else {
                yh_row[t] = -1;
                yh_slot[t] = 0;
            }
        }
        if (lane < M) {
            xh_col = xo - M + lane;
            xh_slot = lane;
        } else if (lane < 2 * M) {
            xh_col = xo + 64 + (lane - M);
            xh_slot = 64 + M + (lane - M);
        } else {
            xh_col = xo;
            xh_slot = -1;
        }
        xh_col = min(max(xh_col, 0), nx - 1);
    }
};

template <int M, int RY, int NW>
struct Lanes {
    int x, xc;
    bool xa;
    int y[RY], yc[RY];
    bool act[RY];
    __device__ Lanes(const ZRect &R, int lane, int w, int xo, int yo, int nx, int ny)
    {
        x = xo + lane;
        xa = x >= R.x0 && x < R.x1;
        xc = min(x, nx - 1);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            y[r] = yo + w + NW * r;
            yc[r] = min(y[r], ny - 1);
            act[r] = xa && y[r] < R.y1;
        }
    }
};

// ---------------------------------------------------------------- halo duties
// The x-halo of a tile plane is RY rows x 2M columns per wave; lane l < 2M RY
// of wave w takes row w + 4 (l / 2M) and halo column l mod 2M of it (M on the
// left of the tile, then M on the right), so each lane holds at most one
// x-halo value per plane.  The y-halo rows are spread as in
// PlaneMap (rows j = w + 4t of the 2M rows above and below the tile).
template <int M, int RY, int NW>
struct XHalo {
    static_assert(2 * M * RY <= 64, "x-halo duties exceed one wave");
    uint32_t off;   // in-plane byte offset (clamped into the grid)
    int col;        // global column (clamped)
    int lds;        // index in the LDS plane, -1: no duty
    __device__ XHalo(int lane, int w, int xo, int yo, int nx, int px, int ny)
    {
        if (lane < 2 * M * RY) {
            const int r = lane / (2 * M), c = lane % (2 * M);
            const int row = w + NW * r;
            const int gx = c < M ? xo - M + c : xo + 64 + (c - M);
            const int gy = min(yo + row, ny - 1);
            col = min(max(gx, 0), nx - 1);
            off = uint32_t(gy * px + col) * 4u;
            lds = (row + M) * (64 + 2 * M) + (c < M ? c : 64 + M + (c - M));
        } else {
            col = 0;
            off = 0;
            lds = -1;
        }
    }
};

// ---------------------------------------------------------------- forward
// One template for every region of the computed box: BX/BY/BZ switch on the
// CPML machinery of an axis (the region lies in that axis' CPML margin);
// the lean interior runs it with all three off.  With an axis on, its psi
// taps come from an LDS plane (x, y; halo on that axis only) or a register
// queue (z), so each psi is read once per step; each cell then applies the
// product-rule update on the axes whose band holds it (pos < lo + 1 or
// pos >= hi - 1), the plain second derivative on the others, and saves its
// strip cells.
//
// The z loop runs in rounds of R = 2M + 2 planes with every queue slot fixed
// at compile time.  At step S of a round (plane z), the u queue slot
// (S + M + d) mod R holds plane z + d for |d| <= M, and the free slot takes
// the head z + M + 1.  The x/y halo of u is fetched together with the head
// (when the neighbouring tiles fetch that plane as their own head, so the
// shared lines are still in L2) and waits in an (M + 1)-deep queue until the
// plane is staged in LDS; halo slot p mod (M + 1) holds plane zb + p.
template <int M, int RY, bool BX, bool BY, bool BZ, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) fwd3_kernel(Geo g, Coef c, Prof pr, StepFields f, float *__restrict__ u_this,
                                                   float *__restrict__ strip, ZSet zs)
{
    constexpr bool SHELL = BX || BY || BZ;
    constexpr int TY = NW * RY;
    using PM = PlaneMap<M, TY, NW>;
    constexpr int LW = PM::LW, PS = PM::LH * PM::LW, YH = PM::YH;
    constexpr int NPL = 1 + (BX ? 1 : 0) + (BY ? 1 : 0);
    constexpr int SX = 1, SY = BX ? 2 : 1;   // plane slots of psi_x, psi_y
    constexpr int R = 2 * M + 2, H = M + 1;
    __shared__ float lds[2][NPL * PS];
    int b, k, xo, yo, zb, ze;
    if (!zcoords<TY>(zs, b, k, xo, yo, zb, ze)) return;
    const ZRect Rc = zs.r[k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nx = g.nx, ny = g.ny;
    const int64_t sxy = g.sxy;
    const Lanes<M, RY, NW> L(Rc, lane, w, xo, yo, nx, ny);
    const PM pm(lane, w, xo, yo, nx, ny);
    const XHalo<M, RY, NW> xh(lane, w, xo, yo, nx, g.px, ny);
    const int64_t off = int64_t(b) * g.n;
    const float dt = g.dt;
    // fields from plane zb - M on; plane z sits at byte offset pl(z)
    const int zbase = zb - M;
    const int64_t fb = off + int64_t(zbase) * sxy, fe = off + g.n;
    const int64_t mb = int64_t(b) * g.mstride + int64_t(zbase) * sxy, me = int64_t(b) * g.mstride + g.n;
    const uint32_t sxy4 = uint32_t(sxy) * 4u;
    auto pl = [&](int z) { return uint32_t(z - zbase) * sxy4; };
    const Fld un(f.u_now, fb, fe), up(f.u_prev, fb, fe), unx(f.u_next, fb, fe), vp(f.vp, mb, me);
    const Fld psx(f.psi[0], fb, fe), psy(f.psi[1], fb, fe), psz(f.psi[2], fb, fe);
    const Fld pnx(f.psin[0], fb, fe), pny(f.psin[1], fb, fe), pnz(f.psin[2], fb, fe);
    const Fld zex(f.zeta[0], fb, fe), zey(f.zeta[1], fb, fe), zez(f.zeta[2], fb, fe);

    uint32_t col[RY], yhc[YH];   // in-plane byte offsets
#pragma unroll
    for (int r = 0; r < RY; ++r) col[r] = uint32_t(L.yc[r] * g.px + L.xc) * 4u;
#pragma unroll
    for (int t = 0; t < YH; ++t) yhc[t] = uint32_t((pm.yh_row[t] >= 0 ? pm.yh_row[t] : 0) * g.px + L.xc) * 4u;

    float ax = 0.f, bx = 0.f, dbx = 0.f, dax = 0.f;
    bool inbx = false;
    float ay[RY], by[RY], dby[RY], day[RY];
    bool inby[RY];
    if constexpr (BX) {
        const int xs = min(max(L.x, M), nx - 1 - M);
        inbx = L.x < g.X0 + 1 || L.x >= g.X1 - 1;
        ax = pr.a[0][xs];
        bx = pr.b[0][xs];
        dbx = pr.db[0][xs];
        dax = d1_taps<M>(c, [&](int q) { return pr.a[0][xs + q]; }, g.inv_dx);
    }
    if constexpr (BY) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int ys = min(max(L.y[r], M), ny - 1 - M);
            inby[r] = L.y[r] < g.Y0 + 1 || L.y[r] >= g.Y1 - 1;
            ay[r] = pr.a[1][ys];
            by[r] = pr.b[1][ys];
            dby[r] = pr.db[1][ys];
            day[r] = d1_taps<M>(c, [&](int q) { return pr.a[1][ys + q]; }, g.inv_dy);
        }
    }

    float q[RY][R], qp[RY][BZ ? R : 1];   // u and psi_z z-queues
    float hx[H], hy[H][YH];               // u halo queue
    // plane-(z+1) registers (single-use fields, fetched one plane ahead)
    float nup[RY], nv[RY], npx[RY], nzx[RY], npy[RY], nzy[RY], nzz[RY];
    float npxh = 0.f, npyh[YH];           // psi_x x-halo, psi_y y-halo

    auto load_halo = [&](int zh, int slot) {   // the u halo of plane zh into halo slot `slot`
        const uint32_t so = pl(zh);
        if (xh.lds >= 0) hx[slot] = un.ld(xh.off, so);
#pragma unroll
        for (int t = 0; t < YH; ++t)
            if (pm.yh_row[t] >= 0) hy[slot][t] = un.ld(yhc[t], so);
    };
    auto load_plane = [&](int zn) {   // the single-use fields of plane zn
        const uint32_t so = pl(zn);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            nup[r] = up.ld_nt(col[r], so);
            nv[r] = vp.ld_nt(col[r], so);
            if constexpr (BX) {
                npx[r] = psx.ld(col[r], so);
                nzx[r] = zex.ld_nt(col[r], so);
            }
            if constexpr (BY) {
                npy[r] = psy.ld(col[r], so);
                nzy[r] = zey.ld_nt(col[r], so);
            }
            if constexpr (BZ) nzz[r] = zez.ld_nt(col[r], so);
        }
        if constexpr (BX)
            if (xh.lds >= 0) npxh = psx.ld(xh.off, so);
        if constexpr (BY) {
#pragma unroll
            for (int t = 0; t < YH; ++t)
                if (pm.yh_row[t] >= 0) npyh[t] = psy.ld(yhc[t], so);
        }
    };
    // stage plane z: its u from queue slot cs, its u halo from halo slot hs
    auto store_plane = [&](float *P, int cs, int hs) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int row = (w + NW * r + M) * LW;
            P[row + lane + M] = q[r][cs];
            if constexpr (BX) P[SX * PS + row + lane + M] = npx[r];
            if constexpr (BY) P[SY * PS + row + lane + M] = npy[r];
        }
        if (xh.lds >= 0) {
            P[xh.lds] = hx[hs];
            if constexpr (BX) P[SX * PS + xh.lds] = npxh;
        }
#pragma unroll
        for (int t = 0; t < YH; ++t) {
            if (pm.yh_row[t] >= 0) {
                P[pm.yh_slot[t] * LW + lane + M] = hy[hs][t];
                if constexpr (BY) P[SY * PS + pm.yh_slot[t] * LW + lane + M] = npyh[t];
            }
        }
    };

    // prologue: u slots t = 0..2M hold planes zb - M + t, halo slots p hold
    // planes zb + p (p <= M), plane zb staged in buffer 0
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
        for (int t = 0; t <= 2 * M; ++t) {
            q[r][t] = un.ld(col[r], pl(zb - M + t));
            if constexpr (BZ) qp[r][t] = psz.ld(col[r], pl(zb - M + t));
        }
    }
#pragma unroll
    for (int p = 0; p < H; ++p) load_halo(min(zb + p, ze - 1 + M), p);
    load_plane(zb);
    store_plane(lds[0], M, 0);
    __syncthreads();

    for (int z0 = zb; z0 < ze; z0 += R) {
        static_for<0, R>([&](auto Sc) -> bool {
        constexpr int S = decltype(Sc)::value;
        constexpr int C = (S + M) % R;       // slot of plane z
        const int z = z0 + S;
        if (z >= ze) return false;
        const float *P = lds[S & 1];
        float cup[RY], cv[RY], czx[RY], czy[RY], czz[RY], cpx[RY], cpy[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            cup[r] = nup[r];
            cv[r] = nv[r];
            czx[r] = nzx[r];
            czy[r] = nzy[r];
            czz[r] = nzz[r];
            cpx[r] = npx[r];
            cpy[r] = npy[r];
        }
        {   // plane z + M + 1: u head and u halo; plane z + 1: the rest
            // (clamped at the segment end, where they are not used)
            const int zh = min(z + M + 1, ze - 1 + M);
            const uint32_t so = pl(zh);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                q[r][(S + 2 * M + 1) % R] = un.ld(col[r], so);
                if constexpr (BZ) qp[r][(S + 2 * M + 1) % R] = psz.ld(col[r], so);
            }
            load_halo(zh, S % H);
            load_plane(min(z + 1, ze - 1));
        }
        const uint32_t so = pl(z);
        bool inbz = false;
        float az = 0.f, bz = 0.f, dbz = 0.f, daz = 0.f;
        if constexpr (BZ) {
            inbz = z < g.Z0 + 1 || z >= g.Z1 - 1;
            az = pr.a[2][z];
            bz = pr.b[2][z];
            dbz = pr.db[2][z];
            daz = d1_taps<M>(c, [&](int t) { return pr.a[2][z + t]; }, g.inv_dz);
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!L.act[r]) continue;
            const int row = (w + NW * r + M) * LW + lane + M;
            const float *Rw = P + row;
            const float u0 = q[r][C];
            auto qz = [&](int t) { return q[r][(C + t + R) % R]; };
            const float lx = lap_taps<M>(c, [&](int t) { return Rw[t]; }, g.inv_dx2);
            const float ly = lap_taps<M>(c, [&](int t) { return Rw[t * LW]; }, g.inv_dy2);
            const float lz = lap_taps<M>(c, qz, g.inv_dz2);
            float w_sum;
            if constexpr (!SHELL) {
                w_sum = lx + ly + lz;
            } else {
                if (strip) save_strips(g, strip + int64_t(b) * g.ns, L.x, L.y[r], z, u0);
                const bool bandx = BX && inbx, bandy = BY && inby[r], bandz = BZ && inbz;
                if (!bandx && !bandy && !bandz) {
                    w_sum = lx + ly + lz;
                } else {
                    w_sum = 0.f;
                    if (bandx) {
                        const float *Px = P + SX * PS + row;
                        const float du = d1_taps<M>(c, [&](int t) { return Rw[t]; }, g.inv_dx);
                        const float dpsi = d1_taps<M>(c, [&](int t) { return Px[t]; }, g.inv_dx);
                        const float psi0 = cpx[r];
                        const float tmp = ((1.0f + bx) * lx + dbx * du) + ax * dpsi + dax * psi0;
                        const float z0v = czx[r];
                        w_sum += (1.0f + bx) * tmp + ax * z0v;
                        pnx.st_nt(bx * du + ax * psi0, col[r], so);
                        zex.st_nt(bx * tmp + ax * z0v, col[r], so);
                    } else {
                        w_sum += lx;
                    }
                    if (bandy) {
                        const float *Py = P + SY * PS + row;
                        const float du = d1_taps<M>(c, [&](int t) { return Rw[t * LW]; }, g.inv_dy);
                        const float dpsi = d1_taps<M>(c, [&](int t) { return Py[t * LW]; }, g.inv_dy);
                        const float psi0 = cpy[r];
                        const float tmp = ((1.0f + by[r]) * ly + dby[r] * du) + ay[r] * dpsi + day[r] * psi0;
                        const float z0v = czy[r];
                        w_sum += (1.0f + by[r]) * tmp + ay[r] * z0v;
                        pny.st_nt(by[r] * du + ay[r] * psi0, col[r], so);
                        zey.st_nt(by[r] * tmp + ay[r] * z0v, col[r], so);
                    } else {
                        w_sum += ly;
                    }
                    if constexpr (BZ) {
                        if (bandz) {
                            auto qpz = [&](int t) { return qp[r][(C + t + R) % R]; };
                            const float du = d1_taps<M>(c, qz, g.inv_dz);
                            const float dpsi = d1_taps<M>(c, qpz, g.inv_dz);
                            const float psi0 = qp[r][C];
                            const float tmp = ((1.0f + bz) * lz + dbz * du) + az * dpsi + daz * psi0;
                            const float z0v = czz[r];
                            w_sum += (1.0f + bz) * tmp + az * z0v;
                            pnz.st_nt(bz * du + az * psi0, col[r], so);
                            zez.st_nt(bz * tmp + az * z0v, col[r], so);
                        } else {
                            w_sum += lz;
                        }
                    } else {
                        w_sum += lz;
                    }
                }
            }
            const float v = cv[r];
            unx.st_nt(2.0f * u0 - cup[r] + (v * v) * dt * dt * w_sum, col[r], so);
            if (u_this) u_this[off + int64_t(z) * sxy + col[r] / 4u] = (v * v) * w_sum;
        }
        if (z + 1 < ze) store_plane(lds[(S + 1) & 1], (C + 1) % R, (S + 1) % H);
        __syncthreads();
        return true;
        });
    }
}

// ---------------------------------------------------------------- adjoint
// The LDS plane G and the z queue hold gw = v^2 dt^2 lambda; a FIFO keeps
// lambda for the centre.  With an axis on (the region lies in that axis'
// adjoint band, outside phys_box(M + 1)), the transposed CPML update of
// that axis is taken through three per-cell fields of the neighbours,
//   T = b zeta + (1 + b) gw,  U = (1 + b) T,  V = b psi + db T,  W = a T,
// (profiles at the neighbour's own position), staged in LDS planes (x, y)
// or register queues (z):
//   L = sum_k lap_k ((U(+k) - U(0)) + (U(-k) - U(0))),
//   Dg = sum_k grad_k (V(+k) - V(-k)),  Dq = sum_k grad_k (W(+k) - W(-k)),
//   psi' = a psi + (da) T(0) - Dq,  zeta' = a (zeta + gw)
// -- the expression tree of the per-cell adjoint (adjoint_split.cu:56-318)
// with each neighbour's T formed once.  img != nullptr fuses the lagged
// full-mode imaging grad += 2 dt^2 img lambda / v.
template <int M, int RY, bool BX, bool BY, bool BZ, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 4 : 1))) adj3_kernel(Geo g, Coef c, Prof pr, AdjFields f, const float *__restrict__ img, float *__restrict__ grad, ZSet zs)
{
    constexpr bool SHELL = BX || BY || BZ;
    constexpr int TY = NW * RY;
    using PM = PlaneMap<M, TY, NW>;
    constexpr int LW = PM::LW, PS = PM::LH * PM::LW, YH = PM::YH;
    constexpr int NPL = 1 + (BX ? 3 : 0) + (BY ? 3 : 0);
    constexpr int SX = 1, SY = BX ? 4 : 1;   // first plane slot of (U, V, W) for x, y
    constexpr int R = 2 * M + 2, H = M + 1;
    __shared__ float lds[2][NPL * PS];
    int b, k, xo, yo, zb, ze;
    if (!zcoords<TY>(zs, b, k, xo, yo, zb, ze)) return;
    const ZRect Rc = zs.r[k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nx = g.nx, ny = g.ny;
    const int64_t sxy = g.sxy;
    const Lanes<M, RY, NW> L(Rc, lane, w, xo, yo, nx, ny);
    const PM pm(lane, w, xo, yo, nx, ny);
    const XHalo<M, RY, NW> xh(lane, w, xo, yo, nx, g.px, ny);
    const int64_t off = int64_t(b) * g.n;
    const float dt2 = g.dt2;
    // fields from plane zb - M on; plane z sits at byte offset pl(z)
    const int zbase = zb - M;
    const int64_t fb = off + int64_t(zbase) * sxy, fe = off + g.n;
    const int64_t mb = int64_t(b) * g.mstride + int64_t(zbase) * sxy, me = int64_t(b) * g.mstride + g.n;
    const uint32_t sxy4 = uint32_t(sxy) * 4u;
    auto pl = [&](int z) { return uint32_t(z - zbase) * sxy4; };
    const Fld ln(f.l_now, fb, fe), lp(f.l_prev, fb, fe), lnx(f.l_next, fb, fe), vp(f.vp, mb, me);
    const Fld psx(f.psi[0], fb, fe), psy(f.psi[1], fb, fe), psz(f.psi[2], fb, fe);
    const Fld zex(f.zeta[0], fb, fe), zey(f.zeta[1], fb, fe), zez(f.zeta[2], fb, fe);
    const Fld pox(f.psio[0], fb, fe), poy(f.psio[1], fb, fe), poz(f.psio[2], fb, fe);
    const Fld zox(f.zetao[0], fb, fe), zoy(f.zetao[1], fb, fe), zoz(f.zetao[2], fb, fe);

    uint32_t col[RY], yhc[YH];   // in-plane byte offsets
#pragma unroll
    for (int r = 0; r < RY; ++r) col[r] = uint32_t(L.yc[r] * g.px + L.xc) * 4u;
#pragma unroll
    for (int t = 0; t < YH; ++t) yhc[t] = uint32_t((pm.yh_row[t] >= 0 ? pm.yh_row[t] : 0) * g.px + L.xc) * 4u;

    // profiles: the lane's column / x-halo duty column (x), rows / halo rows (y)
    float ax = 0.f, bx = 0.f, dbx = 0.f, dax = 0.f, hax = 0.f, hbx = 0.f, hdbx = 0.f;
    float ay[RY], by[RY], dby[RY], day[RY], hay[YH], hby[YH], hdby[YH];
    if constexpr (BX) {
        const int xs = min(max(L.x, M), nx - 1 - M);
        ax = pr.a[0][L.xc];
        bx = pr.b[0][L.xc];
        dbx = pr.db[0][L.xc];
        dax = d1_taps<M>(c, [&](int q) { return pr.a[0][xs + q]; }, g.inv_dx);
        hax = pr.a[0][xh.col];
        hbx = pr.b[0][xh.col];
        hdbx = pr.db[0][xh.col];
    }
    if constexpr (BY) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int ys = min(max(L.y[r], M), ny - 1 - M);
            ay[r] = pr.a[1][L.yc[r]];
            by[r] = pr.b[1][L.yc[r]];
            dby[r] = pr.db[1][L.yc[r]];
            day[r] = d1_taps<M>(c, [&](int q) { return pr.a[1][ys + q]; }, g.inv_dy);
        }
#pragma unroll
        for (int t = 0; t < YH; ++t) {
            const int yr = pm.yh_row[t] >= 0 ? pm.yh_row[t] : 0;
            hay[t] = pr.a[1][yr];
            hby[t] = pr.b[1][yr];
            hdby[t] = pr.db[1][yr];
        }
    }
    auto gwof = [&](float lv, float v) { return v * v * dt2 * lv; };
    // U, V, W of one cell from its raw values and its profile (b, db, a)
    auto uvw = [&](float zeta, float psi, float gwv, float bb, float dd, float aa, float &U, float &V, float &W) {
        const float T = bb * zeta + (1.0f + bb) * gwv;
        U = (1.0f + bb) * T;
        V = bb * psi + dd * T;
        W = aa * T;
    };

    // z queues over rounds of R = 2M + 2 planes (slot (S + M + d) mod R holds
    // plane z + d): gw (z taps), lambda (centre), U/V/W of the z axis; the gw
    // halo of a plane is formed with its queue head and waits in an
    // (M + 1)-deep queue (slot p mod (M + 1) holds plane zb + p) until staged.
    float qg[RY][R], ql[RY][R];
    float qu[RY][BZ ? R : 1], qv[RY][BZ ? R : 1], qw[RY][BZ ? R : 1];
    float hgx[H], hgy[H][YH];
    // heads of plane z + M + 1 in flight
    float nl[RY], nvh[RY], nzh[RY], nph[RY], nhl = 0.f, nhv = 0.f, nyl[YH], nyv[YH];
    // single-use fields of plane z + 1 in flight
    float nlp[RY], nzx[RY], npx[RY], nzxh = 0.f, npxh = 0.f, nzy[RY], npy[RY], nzyh[YH], npyh[YH], nzz[RY], npz[RY];

    auto load_head = [&](int zh) {   // queue heads and the lambda / vp halo of plane zh
        const uint32_t so = pl(zh);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            nl[r] = ln.ld(col[r], so);
            nvh[r] = vp.ld(col[r], so);
            if constexpr (BZ) {
                nzh[r] = zez.ld_nt(col[r], so);
                nph[r] = psz.ld_nt(col[r], so);
            }
        }
        if (xh.lds >= 0) {
            nhl = ln.ld(xh.off, so);
            nhv = vp.ld(xh.off, so);
        }
#pragma unroll
        for (int t = 0; t < YH; ++t) {
            if (pm.yh_row[t] >= 0) {
                nyl[t] = ln.ld(yhc[t], so);
                nyv[t] = vp.ld(yhc[t], so);
            }
        }
    };
    auto push_head = [&](int zh, int slot, int hslot) {   // the heads of plane zh into their queue slots
        float bz_ = 0.f, dbz_ = 0.f, az_ = 0.f;
        if constexpr (BZ) {
            bz_ = pr.b[2][zh];
            dbz_ = pr.db[2][zh];
            az_ = pr.a[2][zh];
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float gh = gwof(nl[r], nvh[r]);
            qg[r][slot] = gh;
            ql[r][slot] = nl[r];
            if constexpr (BZ) uvw(nzh[r], nph[r], gh, bz_, dbz_, az_, qu[r][slot], qv[r][slot], qw[r][slot]);
        }
        hgx[hslot] = gwof(nhl, nhv);
#pragma unroll
        for (int t = 0; t < YH; ++t) hgy[hslot][t] = gwof(nyl[t], nyv[t]);
    };
    auto load_plane = [&](int zn) {   // single-use fields of plane zn
        const uint32_t so = pl(zn);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            nlp[r] = lp.ld_nt(col[r], so);
            if constexpr (BX) {
                nzx[r] = zex.ld(col[r], so);
                npx[r] = psx.ld(col[r], so);
            }
            if constexpr (BY) {
                nzy[r] = zey.ld(col[r], so);
                npy[r] = psy.ld(col[r], so);
            }
            if constexpr (BZ) {
                nzz[r] = zez.ld(col[r], so);
                npz[r] = psz.ld(col[r], so);
            }
        }
        if constexpr (BX) {
            if (xh.lds >= 0) {
                nzxh = zex.ld(xh.off, so);
                npxh = psx.ld(xh.off, so);
            }
        }
        if constexpr (BY) {
#pragma unroll
            for (int t = 0; t < YH; ++t) {
                if (pm.yh_row[t] >= 0) {
                    nzyh[t] = zey.ld(yhc[t], so);
                    npyh[t] = psy.ld(yhc[t], so);
                }
            }
        }
    };
    // stage plane z: centre gw from queue slot cs, gw halo from halo slot hs
    auto store_plane = [&](float *P, int cs, int hs) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int row = (w + NW * r + M) * LW;
            const float gc = qg[r][cs];
            P[row + lane + M] = gc;
            if constexpr (BX) {
                float U, V, W;
                uvw(nzx[r], npx[r], gc, bx, dbx, ax, U, V, W);
                P[SX * PS + row + lane + M] = U;
                P[(SX + 1) * PS + row + lane + M] = V;
                P[(SX + 2) * PS + row + lane + M] = W;
            }
            if constexpr (BY) {
                float U, V, W;
                uvw(nzy[r], npy[r], gc, by[r], dby[r], ay[r], U, V, W);
                P[SY * PS + row + lane + M] = U;
                P[(SY + 1) * PS + row + lane + M] = V;
                P[(SY + 2) * PS + row + lane + M] = W;
            }
        }
        if (xh.lds >= 0) {
            const float gh = hgx[hs];
            P[xh.lds] = gh;
            if constexpr (BX) {
                float U, V, W;
                uvw(nzxh, npxh, gh, hbx, hdbx, hax, U, V, W);
                P[SX * PS + xh.lds] = U;
                P[(SX + 1) * PS + xh.lds] = V;
                P[(SX + 2) * PS + xh.lds] = W;
            }
        }
#pragma unroll
        for (int t = 0; t < YH; ++t) {
            if (pm.yh_row[t] >= 0) {
                const int row = pm.yh_slot[t] * LW + lane + M;
                const float gh = hgy[hs][t];
                P[row] = gh;
                if constexpr (BY) {
                    float U, V, W;
                    uvw(nzyh[t], npyh[t], gh, hby[t], hdby[t], hay[t], U, V, W);
                    P[SY * PS + row] = U;
                    P[(SY + 1) * PS + row] = V;
                    P[(SY + 2) * PS + row] = W;
                }
            }
        }
    };

    // prologue: queue slot t holds plane zb - M + t (t <= 2M), halo slot p
    // plane zb + p (p <= M); plane zb staged in buffer 0
#pragma unroll
    for (int t = 0; t <= 2 * M; ++t) {
        load_head(min(zb - M + t, ze - 1 + M));
        push_head(min(zb - M + t, ze - 1 + M), t, t >= M ? t - M : 0);
    }
    load_plane(zb);
    store_plane(lds[0], M, 0);
    __syncthreads();

    for (int z0 = zb; z0 < ze; z0 += R) {
        static_for<0, R>([&](auto Sc) -> bool {
        constexpr int S = decltype(Sc)::value;
        constexpr int C = (S + M) % R;
        const int z = z0 + S;
        if (z >= ze) return false;
        const float *P = lds[S & 1];
        float clp[RY], czx[RY], cpx[RY], czy[RY], cpy[RY], czz[RY], cpz[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            clp[r] = nlp[r];
            czx[r] = nzx[r];
            cpx[r] = npx[r];
            czy[r] = nzy[r];
            cpy[r] = npy[r];
            czz[r] = nzz[r];
            cpz[r] = npz[r];
        }
        const int zh = min(z + M + 1, ze - 1 + M);   // clamped at the segment end, unused there
        load_head(zh);
        load_plane(min(z + 1, ze - 1));
        const uint32_t so = pl(z);
        float az = 0.f, bz = 0.f, daz = 0.f;
        if constexpr (BZ) {
            az = pr.a[2][z];
            bz = pr.b[2][z];
            daz = d1_taps<M>(c, [&](int t) { return pr.a[2][z + t]; }, g.inv_dz);
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!L.act[r]) continue;
            const int row = (w + NW * r + M) * LW + lane + M;
            const float *Gw = P + row;
            const float gun = ql[r][C];
            const float gw = qg[r][C];
            auto zq = [&](const float (&qq)[R], int t) { return qq[(C + t + R) % R]; };
            if (grad && img && L.x >= g.X0 && L.x < g.X1 && L.y[r] >= g.Y0 && L.y[r] < g.Y1 && z >= g.Z0 &&
                z < g.Z1) {
                const int64_t i = off + int64_t(z) * sxy + col[r] / 4u;
                grad[i] += 2.f * g.dt * g.dt * img[i] * gun / vp.ld(col[r], so);
            }
            float out;
            if constexpr (!SHELL) {
                float lx = 0.f, ly = 0.f, lz = 0.f;
#pragma unroll
                for (int t = 1; t <= M; ++t) {
                    lx += c.lap[t] * ((Gw[t] - gw) + (Gw[-t] - gw));
                    ly += c.lap[t] * ((Gw[t * LW] - gw) + (Gw[-t * LW] - gw));
                    lz += c.lap[t] * ((zq(qg[r], t) - gw) + (zq(qg[r], -t) - gw));
                }
                out = 2.0f * gun - clp[r] + lx * g.inv_dx2.inv + ly * g.inv_dy2.inv + lz * g.inv_dz2.inv;
            } else {
                float Lt[3], Dg[3];
                if constexpr (BX) {
                    const float *Ux = P + SX * PS + row, *Vx = P + (SX + 1) * PS + row, *Wx = P + (SX + 2) * PS + row;
                    const float gt0 = bx * czx[r] + (1.0f + bx) * gw;
                    const float l0 = (1.0f + bx) * gt0;
                    float Ls = 0.f, Dgg = 0.f, Dq = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) {
                        Ls += c.lap[t] * ((Ux[t] - l0) + (Ux[-t] - l0));
                        Dgg += c.grad[t] * (Vx[t] - Vx[-t]);
                        Dq += c.grad[t] * (Wx[t] - Wx[-t]);
                    }
                    Lt[0] = Ls * g.inv_dx2.inv;
                    Dg[0] = Dgg * g.inv_dx.inv;
                    Dq *= g.inv_dx.inv;
                    pox.st_nt(ax * cpx[r] + dax * gt0 - Dq, col[r], so);
                    zox.st_nt(ax * (czx[r] + gw), col[r], so);
                } else {
                    float Ls = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) Ls += c.lap[t] * ((Gw[t] - gw) + (Gw[-t] - gw));
                    Lt[0] = Ls * g.inv_dx2.inv;
                    Dg[0] = 0.f;
                }
                if constexpr (BY) {
                    const float *Uy = P + SY * PS + row, *Vy = P + (SY + 1) * PS + row, *Wy = P + (SY + 2) * PS + row;
                    const float gt0 = by[r] * czy[r] + (1.0f + by[r]) * gw;
                    const float l0 = (1.0f + by[r]) * gt0;
                    float Ls = 0.f, Dgg = 0.f, Dq = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) {
                        Ls += c.lap[t] * ((Uy[t * LW] - l0) + (Uy[-t * LW] - l0));
                        Dgg += c.grad[t] * (Vy[t * LW] - Vy[-t * LW]);
                        Dq += c.grad[t] * (Wy[t * LW] - Wy[-t * LW]);
                    }
                    Lt[1] = Ls * g.inv_dy2.inv;
                    Dg[1] = Dgg * g.inv_dy.inv;
                    Dq *= g.inv_dy.inv;
                    poy.st_nt(ay[r] * cpy[r] + day[r] * gt0 - Dq, col[r], so);
                    zoy.st_nt(ay[r] * (czy[r] + gw), col[r], so);
                } else {
                    float Ls = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) Ls += c.lap[t] * ((Gw[t * LW] - gw) + (Gw[-t * LW] - gw));
                    Lt[1] = Ls * g.inv_dy2.inv;
                    Dg[1] = 0.f;
                }
                if constexpr (BZ) {
                    const float gt0 = bz * czz[r] + (1.0f + bz) * gw;
                    const float l0 = (1.0f + bz) * gt0;
                    float Ls = 0.f, Dgg = 0.f, Dq = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) {
                        Ls += c.lap[t] * ((zq(qu[r], t) - l0) + (zq(qu[r], -t) - l0));
                        Dgg += c.grad[t] * (zq(qv[r], t) - zq(qv[r], -t));
                        Dq += c.grad[t] * (zq(qw[r], t) - zq(qw[r], -t));
                    }
                    Lt[2] = Ls * g.inv_dz2.inv;
                    Dg[2] = Dgg * g.inv_dz.inv;
                    Dq *= g.inv_dz.inv;
                    poz.st_nt(az * cpz[r] + daz * gt0 - Dq, col[r], so);
                    zoz.st_nt(az * (czz[r] + gw), col[r], so);
                } else {
                    float Ls = 0.f;
#pragma unroll
                    for (int t = 1; t <= M; ++t) Ls += c.lap[t] * ((zq(qg[r], t) - gw) + (zq(qg[r], -t) - gw));
                    Lt[2] = Ls * g.inv_dz2.inv;
                    Dg[2] = 0.f;
                }
                out = 2.0f * gun - clp[r] + (Lt[0] + Lt[1] + Lt[2] - Dg[0] - Dg[1] - Dg[2]);
            }
            lnx.st_nt(out, col[r], so);
        }
        push_head(zh, (S + 2 * M + 1) % R, S % H);
        if (z + 1 < ze) store_plane(lds[(S + 1) & 1], (C + 1) % R, (S + 1) % H);
        __syncthreads();
        return true;
        });
    }
}

// ---------------------------------------------------------------- BS recon
// Reverse leapfrog without PML over one box, fused u_tt imaging on the cells
// of [lo + W, hi - W) per axis (the cells the strip restore leaves alone).
template <int M, int RY, int NW>
__global__ void __launch_bounds__(64 * NW) recon_zm_kernel(Geo g, Coef c, const float *__restrict__ vpb,
                                                       const float *__restrict__ u_prev, const float *__restrict__ u_now,
                                                       float *__restrict__ u_next, const float *__restrict__ adj,
                                                       float *__restrict__ grad, ZSet zs)
{
    constexpr int TY = NW * RY;
    using PM = PlaneMap<M, TY, NW>;
    constexpr int LW = PM::LW;
    __shared__ float lds[2][PM::LH * PM::LW];
    int b, k, xo, yo, zb, ze;
    if (!zcoords<TY>(zs, b, k, xo, yo, zb, ze)) return;
    const ZRect R = zs.r[k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nx = g.nx, ny = g.ny;
    const int64_t sxy = g.sxy;
    const Lanes<M, RY, NW> L(R, lane, w, xo, yo, nx, ny);
    const PM pm(lane, w, xo, yo, nx, ny);
    const int64_t off = int64_t(b) * g.n;
    const float dt = g.dt;
    const int zbase = zb - M;
    const int64_t fb = off + int64_t(zbase) * sxy, fe = off + g.n;
    const int64_t mb = int64_t(b) * g.mstride + int64_t(zbase) * sxy, me = int64_t(b) * g.mstride + g.n;
    const uint32_t sxy4 = uint32_t(sxy) * 4u;
    auto pl = [&](int z) { return uint32_t(z - zbase) * sxy4; };
    const Fld un(u_now, fb, fe), up(u_prev, fb, fe), unx(u_next, fb, fe), vp(vpb, mb, me);
    const Fld ad(adj, fb, fe), gr(grad, fb, fe);   // unused without grad
    const bool img_x = L.x >= g.X0 + g.W && L.x < g.X1 - g.W;
    bool img_y[RY];
    uint32_t col[RY], xhc[RY], yhc[PM::YH];   // in-plane byte offsets
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        col[r] = uint32_t(L.yc[r] * g.px + L.xc) * 4u;
        xhc[r] = uint32_t(L.yc[r] * g.px + pm.xh_col) * 4u;
        img_y[r] = L.y[r] >= g.Y0 + g.W && L.y[r] < g.Y1 - g.W;
    }
#pragma unroll
    for (int t = 0; t < PM::YH; ++t) yhc[t] = uint32_t((pm.yh_row[t] >= 0 ? pm.yh_row[t] : 0) * g.px + L.xc) * 4u;

    constexpr int Q = 2 * M + 1;
    float q[RY][Q];
    float nq[RY], nup[RY], nv[RY], nxh[RY], na[RY], ng[RY];
    float nyh[PM::YH];
    auto load_plane = [&](int zn) {
        const uint32_t so = pl(zn);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            nup[r] = up.ld(col[r], so);
            nv[r] = vp.ld(col[r], so);
            if (grad) {
                na[r] = ad.ld(col[r], so);
                ng[r] = gr.ld(col[r], so);
            }
            nxh[r] = un.ld(xhc[r], so);
        }
#pragma unroll
        for (int t = 0; t < PM::YH; ++t)
            if (pm.yh_row[t] >= 0) nyh[t] = un.ld(yhc[t], so);
    };
    auto store_plane = [&](float *P, int cs) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            P[(w + NW * r + M) * LW + lane + M] = q[r][cs];
            if (pm.xh_slot >= 0) P[(w + NW * r + M) * LW + pm.xh_slot] = nxh[r];
        }
#pragma unroll
        for (int t = 0; t < PM::YH; ++t)
            if (pm.yh_row[t] >= 0) P[pm.yh_slot[t] * LW + lane + M] = nyh[t];
    };
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        na[r] = ng[r] = 0.f;
#pragma unroll
        for (int t = 0; t < Q; ++t) q[r][t] = un.ld(col[r], pl(zb - M + t));
    }
    load_plane(zb);
    store_plane(lds[zb & 1], M);
    __syncthreads();

    // rounds of Q planes with compile-time queue slots (as in fwd3_kernel)
    for (int z0 = zb; z0 < ze; z0 += Q) {
        static_for<0, Q>([&](auto Sc) -> bool {
        constexpr int S = decltype(Sc)::value;
        constexpr int C = (S + M) % Q;
        const int z = z0 + S;
        if (z >= ze) return false;
        const float *P = lds[z & 1];
        float cup[RY], cv[RY], ca[RY], cg[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            cup[r] = nup[r];
            cv[r] = nv[r];
            ca[r] = na[r];
            cg[r] = ng[r];
        }
        {
            const int zn = min(z + 1, ze - 1);
            const uint32_t so = pl(zn + M);
#pragma unroll
            for (int r = 0; r < RY; ++r) nq[r] = un.ld(col[r], so);
            load_plane(zn);
        }
        const uint32_t so = pl(z);
        const bool img_z = z >= g.Z0 + g.W && z < g.Z1 - g.W;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!L.act[r]) continue;
            const float *Rw = P + (w + NW * r + M) * LW + lane + M;
            const float u0 = q[r][C];
            const float w_sum = lap_taps<M>(c, [&](int t) { return Rw[t]; }, g.inv_dx2) +
                                lap_taps<M>(c, [&](int t) { return Rw[t * LW]; }, g.inv_dy2) +
                                lap_taps<M>(c, [&](int t) { return q[r][(C + t + Q) % Q]; }, g.inv_dz2);
            const float v = cv[r];
            const float unxt = 2.0f * u0 - cup[r] + (v * v) * dt * dt * w_sum;
            unx.st(unxt, col[r], so);
            if (grad && img_x && img_y[r] && img_z) {
                const float u_tt = fd::scaled(unxt - 2 * u0 + cup[r], g.inv_dt2);
                gr.st(cg[r] + 2.f * dt * dt * u_tt * ca[r] / v, col[r], so);
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) q[r][S] = nq[r];
        if (z + 1 < ze) store_plane(lds[(z + 1) & 1], (C + 1) % Q);
        __syncthreads();
        return true;
        });
    }
}

// ---------------------------------------------------------------- host plans
// Rows per lane and waves per workgroup of the interior kernels (a block's
// tile is 64 x NW*RY cells; each plane re-reads 2M halo rows per tile, so a
// taller tile of more waves cuts the y-halo share without more registers per
// lane).  The margin kernels keep 4 waves.
constexpr int kRyFwd = 4;
constexpr int kRyAdj = 4;
constexpr int kRyRecon = 4;
constexpr int kNwFwd = 8;
constexpr int kNwAdj = 4;
constexpr int kNwRecon = 8;

ZRect make_rect(int x0, int x1, int y0, int y1, int z0, int z1, int tz, int TY, int px)
{
    ZRect R{};
    R.x0 = x0; R.x1 = x1; R.y0 = y0; R.y1 = y1; R.z0 = z0; R.z1 = z1; R.tz = tz;
    // line-aligned tile origin unless that costs an extra column of tiles
    const int xa = px % 32 == 0 ? x0 - x0 % 32 : x0;
    R.xs = (x1 - xa + 63) / 64 <= (x1 - x0 + 63) / 64 ? xa : x0;
    R.ntx = (x1 - R.xs + 63) / 64;
    R.nty = (y1 - y0 + TY - 1) / TY;
    R.nseg = (z1 - z0 + tz - 1) / tz;
    R.nblk = R.ntx * R.nty * R.nseg;
    return R;
}

void add_rect(ZSet &s, int x0, int x1, int y0, int y1, int z0, int z1, int tz, int TY, int px)
{
    if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
    SWP_CHECK(s.n < kRectMax, "acoustic3d: too many block rectangles");
    s.r[s.n] = make_rect(x0, x1, y0, y1, z0, z1, tz, TY, px);
    s.per_shot += s.r[s.n].nblk;
    ++s.n;
}

// Per axis, the computed range [M, n-M) cut at the lean interior
// [lo + M + 1, hi - M - 1): low margin [c0, a), middle [a, b), high margin
// [b, c1).  The margins are exactly the cells in that axis' adjoint CPML band
// and hold its forward band and strip cells.
struct Cuts {
    int c0, a, b, c1;
};

// align > 1 (the x axis on line-aligned rows): the interior starts at the
// next multiple of align, so its tiles cover whole 128-byte lines; the low
// margin takes the few extra cells (its kernels run them without CPML terms).
Cuts cuts(int n, int M, int lo, int hi, int align = 1)
{
    Cuts k;
    k.c0 = M;
    k.c1 = n - M;
    k.a = std::min(std::max(lo + M + 1, k.c0), k.c1);
    k.b = std::min(std::max(hi - M - 1, k.a), k.c1);
    if (align > 1) k.a = std::min((k.a + align - 1) / align * align, k.b);
    return k;
}

// The regions whose axes lie in a margin exactly where (ex, ey, ez) say so.
ZSet region_set(const Geo &g, bool ex, bool ey, bool ez, int tz, int TY)
{
    const Cuts X = cuts(g.nx, g.M, g.X0, g.X1, 32), Y = cuts(g.ny, g.M, g.Y0, g.Y1), Z = cuts(g.nz, g.M, g.Z0, g.Z1);
    int xs[2][2], ys[2][2], zs[2][2];
    int nxr = 0, nyr = 0, nzr = 0;
    auto fill = [](const Cuts &k, bool edge, int out[2][2], int &n) {
        if (edge) {
            out[0][0] = k.c0; out[0][1] = k.a;
            out[1][0] = k.b; out[1][1] = k.c1;
            n = 2;
        } else {
            out[0][0] = k.a; out[0][1] = k.b;
            n = 1;
        }
    };
    fill(X, ex, xs, nxr);
    fill(Y, ey, ys, nyr);
    fill(Z, ez, zs, nzr);
    ZSet s{};
    for (int i = 0; i < nzr; ++i)
        for (int j = 0; j < nyr; ++j)
            for (int l = 0; l < nxr; ++l)
                add_rect(s, xs[l][0], xs[l][1], ys[j][0], ys[j][1], zs[i][0], zs[i][1], tz, TY, g.px);
    return s;
}

template <class K>
void by_order(int M, K &&k)
{
    switch (M) {
    case 1: k(std::integral_constant<int, 1>{}); break;
    case 2: k(std::integral_constant<int, 2>{}); break;
    case 3: k(std::integral_constant<int, 3>{}); break;
    case 4: k(std::integral_constant<int, 4>{}); break;
    default: fail("acoustic3d z-march: order ", 2 * M, " not instantiated");
    }
}

// The region set with its segment height: tz, halved (down to 16 planes)
// while the launch would have fewer than `min_waves` waves (default 4096, 16
// per CU; SWP_REGION_WAVES overrides, 1 keeps tz).  A block walks its
// segment one plane (one load latency) at a time, so a region of few blocks
// -- the edges and corners of the CPML margins -- runs for about tz plane
// latencies however few cells it holds; shorter segments give it more
// blocks that march fewer planes, at the cost of more 2M + 1 plane segment
// prologues (which the larger regions avoid).
ZSet region_set_sized(const Geo &g, bool ex, bool ey, bool ez, int tz, int TY, int NW)
{
    const int64_t min_waves = tuning(-1, "SWP_REGION_WAVES", 4096);
    ZSet s = region_set(g, ex, ey, ez, tz, TY);
    while (tz > 16 && int64_t(s.per_shot) * g.B * NW < min_waves) {
        tz /= 2;
        s = region_set(g, ex, ey, ez, tz, TY);
    }
    return s;
}

dim3 grid_of(const ZSet &s, int B)
{
    const int64_t n = int64_t(s.per_shot) * B;
    SWP_CHECK(n < (int64_t(1) << 31), "acoustic3d: launch too large");
    return dim3(unsigned(n));
}

// Rows per lane of the margin kernels, by the number of axes they carry.
template <bool BX, bool BY, bool BZ>
constexpr int ry_fwd() { return (int(BX) + int(BY) + int(BZ)) <= 1 ? 2 : 1; }
template <bool BX, bool BY, bool BZ>
constexpr int ry_adj() { return (int(BX) + int(BY) + int(BZ)) <= 1 ? 2 : 1; }

template <int M, bool BX, bool BY, bool BZ>
void fwd_region(const Ctx &x, const StepFields &f, float *u_this, float *strip)
{
    constexpr bool SHELL = BX || BY || BZ;
    constexpr int RY = SHELL ? ry_fwd<BX, BY, BZ>() : kRyFwd;
    constexpr int NW = SHELL ? 4 : kNwFwd;
    const ZSet s = region_set_sized(x.g, BX, BY, BZ, x.tz, NW * RY, NW);
    if (!s.per_shot) return;
    if (census_on()) {
        const char site[] = {'a', '3', 'd', '.', 'f', 'w', 'd', '3', '.', char('0' + BX), char('0' + BY), char('0' + BZ), 0};
        census_add(site, int64_t(s.per_shot) * x.g.B);
    }
    hipLaunchKernelGGL((fwd3_kernel<M, RY, BX, BY, BZ, NW>), grid_of(s, x.g.B), dim3(64 * NW), 0, stream(), x.g, x.c, x.pr,
                       f, u_this, strip, s);
}

template <int M, bool BX, bool BY, bool BZ>
void adj_region(const Ctx &x, const AdjFields &f, const float *img, float *grad)
{
    constexpr bool SHELL = BX || BY || BZ;
    constexpr int RY = SHELL ? ry_adj<BX, BY, BZ>() : kRyAdj;
    constexpr int NW = SHELL ? 4 : kNwAdj;
    const ZSet s = region_set_sized(x.g, BX, BY, BZ, x.tz, NW * RY, NW);
    if (!s.per_shot) return;
    if (census_on()) {
        const char site[] = {'a', '3', 'd', '.', 'a', 'd', 'j', '3', '.', char('0' + BX), char('0' + BY), char('0' + BZ), 0};
        census_add(site, int64_t(s.per_shot) * x.g.B);
    }
    hipLaunchKernelGGL((adj3_kernel<M, RY, BX, BY, BZ, NW>), grid_of(s, x.g.B), dim3(64 * NW), 0, stream(), x.g, x.c, x.pr,
                       f, img, grad, s);
}

}  // namespace

bool zm_supported(const Geo &g, int &tz)
{
    if (g.M < 1 || g.M > 4) return false;
    const int64_t planes = ((int64_t(1) << 32) - 1) / (4 * g.sxy);   // planes addressable from a segment base
    const int64_t fit = planes - 2 * g.M - 1;
    if (fit < 1) return false;
    tz = int(std::min<int64_t>(tz, fit));
    return true;
}

int shell_boxes(const Geo &g, Box out[6])
{
    // the six slabs around the lean interior (per-cell shell launches)
    const Cuts X = cuts(g.nx, g.M, g.X0, g.X1, 32), Y = cuts(g.ny, g.M, g.Y0, g.Y1), Z = cuts(g.nz, g.M, g.Z0, g.Z1);
    const Box cand[6] = {
        {X.c0, X.c1, Y.c0, Y.c1, Z.c0, Z.a}, {X.c0, X.c1, Y.c0, Y.c1, Z.b, Z.c1},
        {X.c0, X.c1, Y.c0, Y.a, Z.a, Z.b},   {X.c0, X.c1, Y.b, Y.c1, Z.a, Z.b},
        {X.c0, X.a, Y.a, Y.b, Z.a, Z.b},     {X.b, X.c1, Y.a, Y.b, Z.a, Z.b},
    };
    int n = 0;
    for (const Box &b : cand)
        if (b.x1 > b.x0 && b.y1 > b.y0 && b.z1 > b.z0) out[n++] = b;
    return n;
}

void launch_fwd_zm(const Ctx &x, const StepFields &f, float *u_this, float *strip, bool shell)
{
    by_order(x.g.M, [&](auto ms) {
        constexpr int M = decltype(ms)::value;
        if (shell) {
            fwd_region<M, true, true, true>(x, f, u_this, strip);
            fwd_region<M, true, true, false>(x, f, u_this, strip);
            fwd_region<M, true, false, true>(x, f, u_this, strip);
            fwd_region<M, false, true, true>(x, f, u_this, strip);
            fwd_region<M, true, false, false>(x, f, u_this, strip);
            fwd_region<M, false, true, false>(x, f, u_this, strip);
            fwd_region<M, false, false, true>(x, f, u_this, strip);
        }
        fwd_region<M, false, false, false>(x, f, u_this, strip);
    });
    SWP_LAUNCHED();
}

void launch_adj_zm(const Ctx &x, const AdjFields &f, const float *img, float *grad, bool shell)
{
    by_order(x.g.M, [&](auto ms) {
        constexpr int M = decltype(ms)::value;
        if (shell) {
            adj_region<M, true, true, true>(x, f, img, grad);
            adj_region<M, true, true, false>(x, f, img, grad);
            adj_region<M, true, false, true>(x, f, img, grad);
            adj_region<M, false, true, true>(x, f, img, grad);
            adj_region<M, true, false, false>(x, f, img, grad);
            adj_region<M, false, true, false>(x, f, img, grad);
            adj_region<M, false, false, true>(x, f, img, grad);
        }
        adj_region<M, false, false, false>(x, f, img, grad);
    });
    SWP_LAUNCHED();
}

void launch_recon_zm(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                     const float *adj, float *grad, int x0, int x1, int y0, int y1, int z0, int z1)
{
    const Geo &g = x.g;
    ZSet s{};
    add_rect(s, x0, x1, y0, y1, z0, z1, x.tz, kNwRecon * kRyRecon, x.g.px);
    if (!s.per_shot) return;
    census_add("a3d.recon_zm", int64_t(s.per_shot) * g.B);
    by_order(g.M, [&](auto ms) {
        constexpr int M = decltype(ms)::value;
        hipLaunchKernelGGL((recon_zm_kernel<M, kRyRecon, kNwRecon>), grid_of(s, g.B), dim3(64 * kNwRecon), 0, stream(), g, x.c, vp, u_prev,
                           u_now, u_next, adj, grad, s);
    });
    SWP_LAUNCHED();
}

}  // namespace a3d
}  // namespace swp
