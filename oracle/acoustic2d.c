/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, CPU restatement of the reference's 2-D constant-density acoustic
 * time-stepping path (forward, exact discrete adjoint, boundary-saving
 * reconstruction, imaging).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker: the
 * product path (sweep_amd) never calls it.
 *
 * Build: oracle/Makefile, gcc -O3 -fopenmp -ffp-contract=off (every float
 * operation rounds once, as IEEE binary32; no FMA contraction), so the oracle
 * is deterministic and thread-count independent: each cell is written by one
 * iteration and reads only the previous time level.
 *
 * Expression trees follow the reference's kernels operand for operand:
 *   forward stencil + cpmlr ring ...... csrc/cuda/equations/acoustic2d/forward_split.cu:39-168
 *                                       (= kernels.cuh:121-288)
 *   Laplacian forms ................... csrc/cuda/operators/laplace.cuh:68-457
 *                                       (order 4 summed in the forward, difference otherwise)
 *   first derivative / d(a psi) ....... csrc/cuda/operators/gradient.cuh:46-88,
 *                                       acoustic2d/kernels.cuh:23-77
 *   exact adjoint (interior + ring) ... acoustic2d/adjoint_split.cu:27-190 (= kernels.cuh:311-459)
 *   BS reconstruction + fused imaging . acoustic2d/kernels.cuh:472-574
 *   strip imaging / full imaging ...... acoustic2d/kernels.cu:3-131
 *   illumination ...................... acoustic2d/kernels.cu:133-189
 *   source / record / strip un-inject . csrc/cuda/common/common.cu:11-37,119-144,252-282
 *   rim zeroing of the seed ........... common.cu:362-400
 *   strip geometry .................... csrc/cuda/common/boundary/strip.cuh:46-120
 *   forward loop ...................... csrc/cuda/common/eq_driver.cuh:377-441
 *   full backward loop ................ eq_driver.cuh:689-707
 *   BS backward loop + it=0 tail ...... eq_driver.cuh:833-880,
 *                                       acoustic2d/driver_traits.cuh:493-635
 *
 * Pinning (tests/test_oracle_acoustic2d.py): golden records and gradients
 * written by the reference itself (tests/golden/), the dot-product adjoint
 * test, central finite differences, the analytic 2-D Green's function, and
 * Full == BS agreement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IDX(iz, ix) ((long)(iz) * nx + (ix))

/* ------------------------------------------------------------------ stencils */

/* Forward Laplacian along one axis (laplace.cuh): order 4 summed form, other
 * fixed orders difference form with the constexpr weights, runtime order the
 * table (k = +k tap first). */
static inline float lap_fwd(const orc_a2d *g, const float *u, long i, long s, float h)
{
    const int M = g->M;
    const float u0 = u[i];
    if (M == 1)
        return ((u[i - s] - u0) + (u[i + s] - u0)) / (h * h);
    if (M == 2) {
        const float c0 = g->slap_c0, c1 = g->slap[0], c2 = g->slap[1];
        return (c1 * (u[i - s] + u[i + s]) + c2 * (u[i - 2 * s] + u[i + 2 * s]) - c0 * u0) / (h * h);
    }
    if (M == 3) {
        const float *c = g->slap;
        return (c[0] * ((u[i - s] - u0) + (u[i + s] - u0)) +
                c[1] * ((u[i - 2 * s] - u0) + (u[i + 2 * s] - u0)) +
                c[2] * ((u[i - 3 * s] - u0) + (u[i + 3 * s] - u0))) / (h * h);
    }
    if (M == 4) {
        const float *c = g->slap;
        return (c[0] * ((u[i - s] - u0) + (u[i + s] - u0)) +
                c[1] * ((u[i - 2 * s] - u0) + (u[i + 2 * s] - u0)) +
                c[2] * ((u[i - 3 * s] - u0) + (u[i + 3 * s] - u0)) +
                c[3] * ((u[i - 4 * s] - u0) + (u[i + 4 * s] - u0))) / (h * h);
    }
    float acc = 0.f;
    for (int k = 1; k <= M; ++k)
        acc += g->lap[k] * ((u[i + k * s] - u0) + (u[i - k * s] - u0));
    return acc / (h * h);
}

/* Centred first derivative of v(k) (gradient.cuh:47-88). */
#define CENTRED(M, cst, tab, VAL, h, out)                                              \
    do {                                                                               \
        if ((M) == 1) (out) = (VAL(1) - VAL(-1)) / (2.f * (h));                        \
        else if ((M) == 2) (out) = ((cst)[0] * (VAL(1) - VAL(-1)) +                     \
                                    (cst)[1] * (VAL(2) - VAL(-2))) / (h);               \
        else if ((M) == 3) (out) = ((cst)[0] * (VAL(1) - VAL(-1)) +                     \
                                    (cst)[1] * (VAL(2) - VAL(-2)) +                     \
                                    (cst)[2] * (VAL(3) - VAL(-3))) / (h);               \
        else if ((M) == 4) (out) = ((cst)[0] * (VAL(1) - VAL(-1)) +                     \
                                    (cst)[1] * (VAL(2) - VAL(-2)) +                     \
                                    (cst)[2] * (VAL(3) - VAL(-3)) +                     \
                                    (cst)[3] * (VAL(4) - VAL(-4))) / (h);               \
        else {                                                                         \
            float acc_ = 0.f;                                                          \
            for (int k_ = 1; k_ <= (M); ++k_) acc_ += (tab)[k_] * (VAL(k_) - VAL(-k_)); \
            (out) = acc_ / (h);                                                        \
        }                                                                              \
    } while (0)

/* ------------------------------------------------------------------ geometry */

static inline int phys_x0(const orc_a2d *g) { return g->pad[2] + g->M; }
static inline int phys_x1(const orc_a2d *g) { return g->nx - g->pad[3] - g->M; }
static inline int phys_z0(const orc_a2d *g) { return g->pad[0] + g->M; }
static inline int phys_z1(const orc_a2d *g) { return g->nz - g->pad[1] - g->M; }

/* ------------------------------------------------------------------ forward */

/* Edge-treatment switch used only to explain the residual difference against
 * the reference's eager path (tests/test_oracle_acoustic2d.py): the compiled
 * path never updates the outer M cells of the runtime grid (computed box
 * [M, n-M), kernels.cuh:144-146, launch/box.cuh computed_box), while the
 * eager path convolves the whole grid with zero padding and zeroes only the
 * first-derivative halo (operators/torch.py:229-241,451-470).  Mode 0 (the
 * default) is the compiled path's; mode 1 mirrors the eager edge, for that
 * one check. */
static int g_edge_mode = 0;
void orc_set_edge_mode(int mode) { g_edge_mode = mode; }

int orc_g_strip_bf16 = 0;
void orc_set_strip_bf16(int on) { orc_g_strip_bf16 = on; }

float orc_bf16_round(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) u = (u | 0x00400000u) & 0xffff0000u;   /* quiet NaN */
    else u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

static inline float tap0(const float *u, long i, int ok) { return ok ? u[i] : 0.f; }

/* Eager-edge forward step: all cells, out-of-grid taps read 0, first
 * derivatives zero within M of the edge of their own axis; w_sum over z then
 * x (commutative with the compiled x-then-z for one addition). */
static void fwd_step_eager_edge(const orc_a2d *g, const float *vp, const float *u_prev,
                                const float *u_now, float *u_next, const float *psix,
                                const float *psiz, float *psixn, float *psizn, float *zetax,
                                float *zetaz)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt, dx = g->dx, dz = g->dz;
    const float *lc = g->lap;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; ++iz) {
        for (int ix = 0; ix < nx; ++ix) {
            const long i = IDX(iz, ix);
            float sx = -lc[0] * u_now[i], sz = -lc[0] * u_now[i];
            for (int k = 1; k <= M; ++k) {
                sx += lc[k] * (tap0(u_now, i - k, ix - k >= 0) + tap0(u_now, i + k, ix + k < nx));
                sz += lc[k] * (tap0(u_now, i - (long)k * nx, iz - k >= 0) +
                               tap0(u_now, i + (long)k * nx, iz + k < nz));
            }
            const float lap_x = sx / (dx * dx), lap_z = sz / (dz * dz);
            float dudx = 0.f, dudz = 0.f, dapx = 0.f, dapz = 0.f;
            if (ix >= M && ix < nx - M) {
                for (int k = 1; k <= M; ++k) {
                    dudx += g->grad[k] * (u_now[i + k] - u_now[i - k]);
                    dapx += g->grad[k] * (g->ax[ix + k] * psix[i + k] - g->ax[ix - k] * psix[i - k]);
                }
                dudx /= dx; dapx /= dx;
            }
            if (iz >= M && iz < nz - M) {
                for (int k = 1; k <= M; ++k) {
                    dudz += g->grad[k] * (u_now[i + (long)k * nx] - u_now[i - (long)k * nx]);
                    dapz += g->grad[k] * (g->az[iz + k] * psiz[i + (long)k * nx] -
                                          g->az[iz - k] * psiz[i - (long)k * nx]);
                }
                dudz /= dz; dapz /= dz;
            }
            const float tz = ((1.0f + g->bz[iz]) * lap_z + g->dbz[iz] * dudz) + dapz;
            const float tx = ((1.0f + g->bx[ix]) * lap_x + g->dbx[ix] * dudx) + dapx;
            const float cz = (1.0f + g->bz[iz]) * tz + g->az[iz] * zetaz[i];
            const float cx = (1.0f + g->bx[ix]) * tx + g->ax[ix] * zetax[i];
            psizn[i] = g->bz[iz] * dudz + g->az[iz] * psiz[i];
            zetaz[i] = g->bz[iz] * tz + g->az[iz] * zetaz[i];
            psixn[i] = g->bx[ix] * dudx + g->ax[ix] * psix[i];
            zetax[i] = g->bx[ix] * tx + g->ax[ix] * zetax[i];
            const float v = vp[i];
            u_next[i] = 2.0f * u_now[i] - u_prev[i] + (v * v) * (dt * dt) * (cz + cx);
        }
    }
}

/* One forward step of one shot over the computed box [M, n-M).  psixn/psizn
 * receive the next psi; zeta is updated in place (each cell reads only its
 * own zeta).  u_this (may be NULL) receives v^2 * w_sum. */
static void fwd_step(const orc_a2d *g, const float *vp, const float *u_prev, const float *u_now,
                     float *u_next, const float *psix, const float *psiz, float *psixn,
                     float *psizn, float *zetax, float *zetaz, float *u_this)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt, dx = g->dx, dz = g->dz;
    const int fx0 = phys_x0(g) + M, fx1 = phys_x1(g) - M;
    const int fz0 = phys_z0(g) + M, fz1 = phys_z1(g) - M;
    const int pml_x0 = g->pad[2] + M, pml_x1 = nx - (g->pad[3] + M);
    const int pml_z0 = g->pad[0] + M, pml_z1 = nz - (g->pad[1] + M);
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz) {
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = IDX(iz, ix);
            const float lap_x = lap_fwd(g, u_now, i, 1, dx);
            const float lap_z = lap_fwd(g, u_now, i, nx, dz);
            const float v = vp[i];
            const float v2_dt2 = (v * v) * dt * dt;
            if (ix >= fx0 && ix < fx1 && iz >= fz0 && iz < fz1) {
                u_next[i] = 2.0f * u_now[i] - u_prev[i] + v2_dt2 * (lap_x + lap_z);
                if (u_this) u_this[i] = (v * v) * (lap_x + lap_z);
                continue;
            }
            const int cx = (ix < pml_x0 + M) || (ix >= pml_x1 - M);
            const int cz = (iz < pml_z0 + M) || (iz >= pml_z1 - M);
            float w_sum = 0.0f;
            if (cx) {
                const float ax_ = g->ax[ix], bx_ = g->bx[ix], dbx_ = g->dbx[ix];
                float dudx, dapsi;
#define UX(k) (u_now[i + (k)])
#define APX(k) (g->ax[ix + (k)] * psix[i + (k)])
                CENTRED(M, g->sgrad, g->grad, UX, dx, dudx);
                CENTRED(M, g->sgrad, g->grad, APX, dx, dapsi);
#undef UX
#undef APX
                const float tmpx = ((1.0f + bx_) * lap_x + dbx_ * dudx) + dapsi;
                w_sum += (1.0f + bx_) * tmpx + ax_ * zetax[i];
                psixn[i] = bx_ * dudx + ax_ * psix[i];
                zetax[i] = bx_ * tmpx + ax_ * zetax[i];
            } else {
                w_sum += lap_x;
            }
            if (cz) {
                const float az_ = g->az[iz], bz_ = g->bz[iz], dbz_ = g->dbz[iz];
                float dudz, dapsi;
#define UZ(k) (u_now[i + (long)(k) * nx])
#define APZ(k) (g->az[iz + (k)] * psiz[i + (long)(k) * nx])
                CENTRED(M, g->sgrad, g->grad, UZ, dz, dudz);
                CENTRED(M, g->sgrad, g->grad, APZ, dz, dapsi);
#undef UZ
#undef APZ
                const float tmpz = ((1.0f + bz_) * lap_z + dbz_ * dudz) + dapsi;
                w_sum += (1.0f + bz_) * tmpz + az_ * zetaz[i];
                psizn[i] = bz_ * dudz + az_ * psiz[i];
                zetaz[i] = bz_ * tmpz + az_ * zetaz[i];
            } else {
                w_sum += lap_z;
            }
            u_next[i] = 2.0f * u_now[i] - u_prev[i] + v2_dt2 * w_sum;
            if (u_this) u_this[i] = (v * v) * w_sum;
        }
    }
}

/* ------------------------------------------------------------------ adjoint */

/* One exact-adjoint step of one shot over the computed box.  grad/img (may
 * be NULL) carry the full-mode fused imaging of the lambda at entry. */
static void adj_step(const orc_a2d *g, const float *vp, const float *l_prev, const float *l_now,
                     float *l_next, const float *psix, const float *psiz, const float *zetax,
                     const float *zetaz, float *psixo, float *psizo, float *zetaxo, float *zetazo,
                     const float *img, float *grad)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt;
    const float dt2 = dt * dt;
    const float *lc = g->lap, *gc = g->grad;
    const float invdx2 = 1.0f / (g->dx * g->dx), invdz2 = 1.0f / (g->dz * g->dz);
    const float invdx = 1.0f / g->dx, invdz = 1.0f / g->dz;
    const int X0 = phys_x0(g), X1 = phys_x1(g), Z0 = phys_z0(g), Z1 = phys_z1(g);
    const int ix0 = X0 + M + 1, ix1 = X1 - M - 1, iz0 = Z0 + M + 1, iz1 = Z1 - M - 1;
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz) {
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = IDX(iz, ix);
            const float gun = l_now[i];
            if (grad && ix >= X0 && ix < X1 && iz >= Z0 && iz < Z1)
                grad[i] += 2.f * dt * dt * img[i] * gun / vp[i];
            float c0c = vp[i];
            c0c = c0c * c0c * dt2;
            const float gw = c0c * gun;
#define GW(n) (vp[n] * vp[n] * dt2 * l_now[n])
            if (ix >= ix0 && ix < ix1 && iz >= iz0 && iz < iz1) {
                float lapx = 0.0f, lapz = 0.0f;
                for (int k = 1; k <= M; ++k) {
                    lapx += lc[k] * ((GW(i + k) - gw) + (GW(i - k) - gw));
                    lapz += lc[k] * ((GW(i + (long)k * nx) - gw) + (GW(i - (long)k * nx) - gw));
                }
                l_next[i] = 2.0f * gun - l_prev[i] + lapx * invdx2 + lapz * invdz2;
                continue;
            }
#undef GW
            const float ax_ = g->ax[ix], az_ = g->az[iz];
            const float bx_ = g->bx[ix], bz_ = g->bz[iz];
            const float gtx0 = bx_ * zetax[i] + (1.0f + bx_) * gw;
            const float gtz0 = bz_ * zetaz[i] + (1.0f + bz_) * gw;
#define GTX(n, c) (g->bx[c] * zetax[n] + (1.0f + g->bx[c]) * (vp[n] * vp[n] * dt2) * l_now[n])
#define GTZ(n, c) (g->bz[c] * zetaz[n] + (1.0f + g->bz[c]) * (vp[n] * vp[n] * dt2) * l_now[n])
            const float lx0 = (1.0f + bx_) * gtx0;
            float Lx = 0.0f, Dgx = 0.0f, Dqx = 0.0f;
            for (int k = 1; k <= M; ++k) {
                const long np = i + k, nm = i - k;
                const int cp = ix + k, cm = ix - k;
                const float tp = GTX(np, cp), tm = GTX(nm, cm);
                const float bp = g->bx[cp], bm = g->bx[cm];
                Lx += lc[k] * (((1.0f + bp) * tp - lx0) + ((1.0f + bm) * tm - lx0));
                Dgx += gc[k] * ((bp * psix[np] + g->dbx[cp] * tp) - (bm * psix[nm] + g->dbx[cm] * tm));
                Dqx += gc[k] * (tp - tm);
            }
            Lx *= invdx2; Dgx *= invdx; Dqx *= invdx;
            const float lz0 = (1.0f + bz_) * gtz0;
            float Lz = 0.0f, Dgz = 0.0f, Dqz = 0.0f;
            for (int k = 1; k <= M; ++k) {
                const long np = i + (long)k * nx, nm = i - (long)k * nx;
                const int cp = iz + k, cm = iz - k;
                const float tp = GTZ(np, cp), tm = GTZ(nm, cm);
                const float bp = g->bz[cp], bm = g->bz[cm];
                Lz += lc[k] * (((1.0f + bp) * tp - lz0) + ((1.0f + bm) * tm - lz0));
                Dgz += gc[k] * ((bp * psiz[np] + g->dbz[cp] * tp) - (bm * psiz[nm] + g->dbz[cm] * tm));
                Dqz += gc[k] * (tp - tm);
            }
            Lz *= invdz2; Dgz *= invdz; Dqz *= invdz;
#undef GTX
#undef GTZ
            l_next[i] = 2.0f * gun - l_prev[i] + (Lx + Lz - Dgx - Dgz);
            psixo[i] = ax_ * (psix[i] - Dqx);
            psizo[i] = az_ * (psiz[i] - Dqz);
            zetaxo[i] = ax_ * (zetax[i] + gw);
            zetazo[i] = az_ * (zetaz[i] + gw);
        }
    }
}

/* -------------------------------------------------- boundary-saving pieces */

static inline int save_width(const orc_a2d *g) { return g->M + 1; }

/* Strip slot layout of one shot, one step: top (W, nxp), bottom (W, nxp),
 * left (nzp, W), right (nzp, W) back to back. */
static long strip_cells(const orc_a2d *g)
{
    const int W = save_width(g);
    const long nxp = phys_x1(g) - phys_x0(g), nzp = phys_z1(g) - phys_z0(g);
    return 2L * W * nxp + 2L * W * nzp;
}

static void strips_io(const orc_a2d *g, float *u, float *slot, int restore)
{
    const int nx = g->nx, W = save_width(g);
    const int x0 = phys_x0(g), x1 = phys_x1(g), z0 = phys_z0(g), z1 = phys_z1(g);
    const int nxp = x1 - x0, nzp = z1 - z0;
    float *top = slot, *bot = top + (long)W * nxp, *lef = bot + (long)W * nxp,
          *rig = lef + (long)nzp * W;
    for (int r = 0; r < W; ++r)
        for (int c = 0; c < nxp; ++c) {
            float *t = &top[(long)r * nxp + c], *b = &bot[(long)r * nxp + c];
            float *ut = &u[IDX(z0 + r, x0 + c)], *ub = &u[IDX(z1 - W + r, x0 + c)];
            if (restore) { *ut = *t; *ub = *b; }
            else { *t = orc_g_strip_bf16 ? orc_bf16_round(*ut) : *ut; *b = orc_g_strip_bf16 ? orc_bf16_round(*ub) : *ub; }
        }
    for (int r = 0; r < nzp; ++r)
        for (int c = 0; c < W; ++c) {
            float *l = &lef[(long)r * W + c], *rr = &rig[(long)r * W + c];
            float *ul = &u[IDX(z0 + r, x0 + c)], *ur = &u[IDX(z0 + r, x1 - W + c)];
            if (restore) { *ul = *l; *ur = *rr; }
            else { *l = orc_g_strip_bf16 ? orc_bf16_round(*ul) : *ul; *rr = orc_g_strip_bf16 ? orc_bf16_round(*ur) : *ur; }
        }
}

static inline int in_strip(const orc_a2d *g, int ix, int iz)
{
    const int W = save_width(g);
    const int x0 = phys_x0(g), x1 = phys_x1(g), z0 = phys_z0(g), z1 = phys_z1(g);
    if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) return 0;
    const int inx = ix >= x0 && ix < x1, inz = iz >= z0 && iz < z1;
    return (iz >= z0 && iz < z0 + W && inx) || (iz >= z1 - W && iz < z1 && inx) ||
           (ix >= x0 && ix < x0 + W && inz) || (ix >= x1 - W && ix < x1 && inz);
}

/* BS reverse reconstruction of one shot: u_next = 2 u_now - u_prev + v^2 dt^2
 * Lap(u_now) on the cells inside the restore band, fused u_tt imaging there. */
static void nopml_step(const orc_a2d *g, const float *vp, const float *u_prev, const float *u_now,
                       float *u_next, const float *adj, float *grad)
{
    const int nx = g->nx, nz = g->nz, M = g->M, W = save_width(g);
    const float dt = g->dt;
    const int bxl = g->pad[2] > 0 ? g->pad[2] + 2 * M + 1 : 2 * M;
    const int bxh = g->pad[3] > 0 ? g->pad[3] + 2 * M + 1 : 2 * M;
    const int bzl = g->pad[0] > 0 ? g->pad[0] + 2 * M + 1 : 2 * M;
    const int bzh = g->pad[1] > 0 ? g->pad[1] + 2 * M + 1 : 2 * M;
    const int gx0 = phys_x0(g) + W, gx1 = phys_x1(g) - W;
    const int gz0 = phys_z0(g) + W, gz1 = phys_z1(g) - W;
#pragma omp parallel for schedule(static)
    for (int iz = bzl; iz < nz - bzh; ++iz) {
        for (int ix = bxl; ix < nx - bxh; ++ix) {
            const long i = IDX(iz, ix);
            const float w_sum = lap_fwd(g, u_now, i, 1, g->dx) + lap_fwd(g, u_now, i, nx, g->dz);
            const float v = vp[i], u0 = u_now[i], up = u_prev[i];
            const float un = 2.0f * u0 - up + (v * v) * dt * dt * w_sum;
            u_next[i] = un;
            if (grad && ix >= gx0 && ix < gx1 && iz >= gz0 && iz < gz1) {
                const float u_tt = (un - 2 * u0 + up) / (dt * dt);
                grad[i] += 2.f * dt * dt * u_tt * adj[i] / v;
            }
        }
    }
}

/* u_tt imaging of the restore strips of the physical box (kernels.cu:83-131);
 * fnext = w^{it+1}, fnow = w^{it-1} - s^{it}, fprev = w^{it}. */
static void band_image(const orc_a2d *g, const float *fnext, const float *fnow, const float *fprev,
                       const float *adj, const float *vp, float *grad)
{
    const int nx = g->nx, W = save_width(g);
    const float dt = g->dt;
    const int x0 = phys_x0(g), x1 = phys_x1(g), z0 = phys_z0(g), z1 = phys_z1(g);
    int wzl = W, wzh = W, wxl = W, wxh = W;
    const int bw = x1 - x0;
    /* bs_strip_widths: clip to a thin box, no cell twice */
    { int e = z1 - z0; wzl = wzl < (e > 0 ? e : 0) ? wzl : (e > 0 ? e : 0);
      int r = e - wzl; r = r > 0 ? r : 0; wzh = wzh < r ? wzh : r; }
    { int e = bw; wxl = wxl < (e > 0 ? e : 0) ? wxl : (e > 0 ? e : 0);
      int r = e - wxl; r = r > 0 ? r : 0; wxh = wxh < r ? wxh : r; }
    const int zi0 = z0 + wzl, zi1 = z1 - wzh;
    for (int iz = z0; iz < z1; ++iz)
        for (int ix = x0; ix < x1; ++ix) {
            const int strip = (iz < zi0) || (iz >= zi1) || (ix < x0 + wxl) || (ix >= x1 - wxh);
            if (!strip) continue;
            const long i = IDX(iz, ix);
            const float u_tt = (fnow[i] - 2 * fprev[i] + fnext[i]) / (dt * dt);
            grad[i] += 2.f * dt * dt * u_tt * adj[i] / vp[i];
        }
}

/* accumulate_illumination_2d over the physical box; fnow == NULL means fnext
 * already holds u_tt; S == NULL is the receiver-only call. */
static void illum(const orc_a2d *g, const float *fnext, const float *fnow, const float *fprev,
                  const float *adj, float *S, float *R)
{
    const int nx = g->nx;
    const float dt = g->dt;
    for (int iz = phys_z0(g); iz < phys_z1(g); ++iz)
        for (int ix = phys_x0(g); ix < phys_x1(g); ++ix) {
            const long i = IDX(iz, ix);
            if (R) { const float ub = adj[i]; R[i] += ub * ub; }
            if (!S) continue;
            float u_tt;
            if (!fnow) u_tt = fnext[i];
            else u_tt = (fnow[i] - 2 * fprev[i] + fnext[i]) / (dt * dt);
            S[i] += u_tt * u_tt;
        }
}

static void add_points(const orc_a2d *g, float *u, const float *w, const int *loc, int n, int it,
                       float sign)
{
    for (int s = 0; s < n; ++s) {
        const int ix = loc[2 * s], iz = loc[2 * s + 1];
        if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
        u[(long)iz * g->nx + ix] += sign * w[(long)s * g->nt + it];
    }
}

static void sub_points_in_strip(const orc_a2d *g, float *u, const float *w, const int *loc, int n,
                                int it)
{
    for (int s = 0; s < n; ++s) {
        const int ix = loc[2 * s], iz = loc[2 * s + 1];
        if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
        if (!in_strip(g, ix, iz)) continue;
        u[(long)iz * g->nx + ix] += -w[(long)s * g->nt + it];
    }
}

static void gather_points(const orc_a2d *g, const float *u, float *out, const int *loc, int n,
                          int it, int accumulate)
{
    for (int r = 0; r < n; ++r) {
        const int ix = loc[2 * r], iz = loc[2 * r + 1];
        if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
        const float v = u[(long)iz * g->nx + ix];
        if (accumulate) out[(long)r * g->nt + it] += v;
        else out[(long)r * g->nt + it] = v;
    }
}

static void zero_rim(const orc_a2d *g, float *u)
{
    const int nx = g->nx, nz = g->nz, wdt = g->abcn + g->M;
    for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix)
            if (ix < wdt || ix >= nx - wdt || iz < wdt || iz >= nz - wdt)
                u[IDX(iz, ix)] = 0.f;
}

/* ------------------------------------------------------------------ drivers */

typedef struct {
    float *u[3];          /* prev, now, next */
    float *psix, *psixn, *psiz, *psizn, *zetax, *zetaz, *zetaxn, *zetazn;
} fields;

static float *zalloc(long n) { return (float *)calloc((size_t)n, sizeof(float)); }

static void fields_alloc(fields *f, long n)
{
    for (int k = 0; k < 3; ++k) f->u[k] = zalloc(n);
    f->psix = zalloc(n); f->psixn = zalloc(n); f->psiz = zalloc(n); f->psizn = zalloc(n);
    f->zetax = zalloc(n); f->zetaz = zalloc(n); f->zetaxn = zalloc(n); f->zetazn = zalloc(n);
}

static void fields_free(fields *f)
{
    for (int k = 0; k < 3; ++k) free(f->u[k]);
    free(f->psix); free(f->psixn); free(f->psiz); free(f->psizn);
    free(f->zetax); free(f->zetaz); free(f->zetaxn); free(f->zetazn);
}

#define SWAPP(a, b) do { float *t_ = (a); (a) = (b); (b) = t_; } while (0)
static void rotate_u(fields *f)
{
    float *p = f->u[0];
    f->u[0] = f->u[1]; f->u[1] = f->u[2]; f->u[2] = p;
}

/* Forward of one shot.  rec (nrec, nt); u_allt (nt, grid) or NULL; strips
 * (nt, strip_cells) or NULL; last_two (2, grid) or NULL. */
static void forward_shot(const orc_a2d *g, const float *vp, const float *wav, const int *sloc,
                         int nsrc, const int *rloc, int nrec, float *rec, float *u_allt,
                         float *strips, float *last_two)
{
    const long n = (long)g->nz * g->nx, ns = strip_cells(g);
    fields f;
    fields_alloc(&f, n);
    for (int it = 0; it < g->nt; ++it) {
        if (g_edge_mode)
            fwd_step_eager_edge(g, vp, f.u[0], f.u[1], f.u[2], f.psix, f.psiz, f.psixn, f.psizn,
                                f.zetax, f.zetaz);
        else
            fwd_step(g, vp, f.u[0], f.u[1], f.u[2], f.psix, f.psiz, f.psixn, f.psizn, f.zetax,
                     f.zetaz, u_allt ? u_allt + (long)it * n : NULL);
        if (strips) strips_io(g, f.u[1], strips + (long)it * ns, 0);
        add_points(g, f.u[2], wav, sloc, nsrc, it, 1.0f);
        if (rec) gather_points(g, f.u[2], rec, rloc, nrec, it, 0);
        rotate_u(&f);
        SWAPP(f.psix, f.psixn);
        SWAPP(f.psiz, f.psizn);
    }
    if (last_two) {
        memcpy(last_two, f.u[0], (size_t)n * sizeof(float));
        memcpy(last_two + n, f.u[1], (size_t)n * sizeof(float));
    }
    fields_free(&f);
}

static void adjoint_once(const orc_a2d *g, const float *vp, fields *a, const float *res,
                         const int *rloc, int nrec, const int *sloc, int nsrc, float *gwav, int it,
                         const float *img, float *grad)
{
    adj_step(g, vp, a->u[0], a->u[1], a->u[2], a->psix, a->psiz, a->zetax, a->zetaz, a->psixn,
             a->psizn, a->zetaxn, a->zetazn, img, grad);
    add_points(g, a->u[2], res, rloc, nrec, it, 1.0f);
    rotate_u(a);
    SWAPP(a->psix, a->psixn);
    SWAPP(a->psiz, a->psizn);
    SWAPP(a->zetax, a->zetaxn);
    SWAPP(a->zetaz, a->zetazn);
    if (gwav) gather_points(g, a->u[1], gwav, sloc, nsrc, it, 1);
}

void orc_a2d_forward(const orc_a2d *g, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, float *rec)
{
    const long n = (long)g->nz * g->nx;
    for (int b = 0; b < g->B; ++b)
        forward_shot(g, vp + (g->vp_batched ? b * n : 0), wav + (long)b * nsrc * g->nt,
                     sloc + (long)b * nsrc * 2, nsrc, rloc + (long)b * nrec * 2, nrec,
                     rec + (long)b * nrec * g->nt, NULL, NULL, NULL);
}

long orc_a2d_strip_cells(const orc_a2d *g) { return strip_cells(g); }

int orc_a2d_gradient(const orc_a2d *g, int mode, const float *vp, const float *wav, const int *sloc,
                     int nsrc, const int *rloc, int nrec, const float *res, float *rec,
                     float *grad_vp, float *grad_wav, float *illum_s, float *illum_r)
{
    const long n = (long)g->nz * g->nx, ns = strip_cells(g);
    for (int b = 0; b < g->B; ++b) {
        const float *vpb = vp + (g->vp_batched ? b * n : 0);
        const float *wb = wav + (long)b * nsrc * g->nt;
        const int *sl = sloc + (long)b * nsrc * 2, *rl = rloc + (long)b * nrec * 2;
        const float *rb = res + (long)b * nrec * g->nt;
        float *gv = grad_vp + b * n, *gw = grad_wav ? grad_wav + (long)b * nsrc * g->nt : NULL;
        float *S = illum_s ? illum_s + b * n : NULL, *R = illum_r ? illum_r + b * n : NULL;
        float *recb = rec ? rec + (long)b * nrec * g->nt : NULL;
        fields a;
        fields_alloc(&a, n);
        if (mode == 0) {
            float *u_allt = zalloc((long)g->nt * n);
            if (!u_allt) { fields_free(&a); return -1; }
            forward_shot(g, vpb, wb, sl, nsrc, rl, nrec, recb, u_allt, NULL, NULL);
            for (int it = g->nt - 1; it >= 0; --it) {
                const float *img = (it + 1 < g->nt) ? u_allt + (long)(it + 1) * n : NULL;
                adjoint_once(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, it, img, img ? gv : NULL);
                if (S || R) illum(g, u_allt + (long)it * n, NULL, NULL, a.u[1], S, R);
            }
            /* trailing image_step at it == 0 */
            {
                const float dt = g->dt;
                const float *u0 = u_allt, *lam = a.u[1];
                const int nx = g->nx;
                for (int iz = phys_z0(g); iz < phys_z1(g); ++iz)
                    for (int ix = phys_x0(g); ix < phys_x1(g); ++ix) {
                        const long i = IDX(iz, ix);
                        gv[i] += 2.f * dt * dt * u0[i] * lam[i] / vpb[i];
                    }
            }
            free(u_allt);
        } else {
            float *strips = zalloc((long)g->nt * ns), *last_two = zalloc(2 * n);
            if (!strips || !last_two) { free(strips); free(last_two); fields_free(&a); return -1; }
            forward_shot(g, vpb, wb, sl, nsrc, rl, nrec, recb, NULL, strips, last_two);
            float *r[3] = {zalloc(n), zalloc(n), zalloc(n)};  /* prev, now, next */
            memcpy(r[0], last_two + n, (size_t)n * sizeof(float));
            memcpy(r[1], last_two, (size_t)n * sizeof(float));
            zero_rim(g, r[0]);
            zero_rim(g, r[1]);
            for (int it = g->nt - 1; it >= 1; --it) {
                adjoint_once(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, it, NULL, NULL);
                nopml_step(g, vpb, r[0], r[1], r[2], a.u[1], gv);
                strips_io(g, r[2], strips + (long)(it - 1) * ns, 1);
                sub_points_in_strip(g, r[2], wb, sl, nsrc, it);
                band_image(g, r[0], r[2], r[1], a.u[1], vpb, gv);
                if (S || R) illum(g, r[0], r[2], r[1], a.u[1], S, R);
                add_points(g, r[2], wb, sl, nsrc, it, 1.0f);
                float *p = r[0]; r[0] = r[1]; r[1] = r[2]; r[2] = p;
            }
            if (g->nt > 0) {
                adjoint_once(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, 0, NULL, NULL);
                if (R) illum(g, NULL, NULL, NULL, a.u[1], NULL, R);
            }
            free(r[0]); free(r[1]); free(r[2]);
            free(strips); free(last_two);
        }
        fields_free(&a);
    }
    return 0;
}

/* Single-step entry points for kernel-level parity tests. */
void orc_a2d_fwd_step(const orc_a2d *g, const float *vp, const float *u_prev, const float *u_now,
                      float *u_next, const float *psix, const float *psiz, float *psixn,
                      float *psizn, float *zetax, float *zetaz, float *u_this)
{
    fwd_step(g, vp, u_prev, u_now, u_next, psix, psiz, psixn, psizn, zetax, zetaz, u_this);
}

void orc_a2d_adj_step(const orc_a2d *g, const float *vp, const float *l_prev, const float *l_now,
                      float *l_next, const float *psix, const float *psiz, const float *zetax,
                      const float *zetaz, float *psixo, float *psizo, float *zetaxo,
                      float *zetazo, const float *img, float *grad)
{
    adj_step(g, vp, l_prev, l_now, l_next, psix, psiz, zetax, zetaz, psixo, psizo, zetaxo, zetazo,
             img, grad);
}
