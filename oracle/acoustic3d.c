/*
 * ORACLE -- test infrastructure only (see acoustic2d.c for the rules).
 *
 * CPU restatement of the reference's 3-D constant-density acoustic path:
 *   forward interior + product-rule CPML shell ... csrc/cuda/equations/acoustic3d/forward_split.cu:81-249
 *   (per axis: tmp = ((1+b) L + db du) + a dpsi + (da) psi on the band plus the db leak cell)
 *   exact adjoint interior + shell ............... acoustic3d/adjoint_split.cu:56-318
 *   BS reconstruction + fused imaging ............ acoustic3d/forward_split.cu:256-328, 505-523
 *   strip imaging, illumination .................. acoustic3d/kernels.cu
 *   driver order (as the 2-D twin; no seed rim-zeroing) acoustic3d/driver_traits.cuh:449-560
 * Shared operators (Laplacian forms, centred first derivative) as in acoustic2d.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

extern int orc_g_strip_bf16;   /* acoustic2d.c, set by orc_set_strip_bf16 */

#define I3(iz, iy, ix) (((long)(iz) * ny + (iy)) * nx + (ix))

static inline float lapf(const orc_a3d *g, const float *u, long i, long s, float h)
{
    const int M = g->M;
    const float u0 = u[i];
    if (M == 1) return ((u[i - s] - u0) + (u[i + s] - u0)) / (h * h);
    if (M == 2) {
        const float c0 = g->slap_c0, c1 = g->slap[0], c2 = g->slap[1];
        return (c1 * (u[i - s] + u[i + s]) + c2 * (u[i - 2 * s] + u[i + 2 * s]) - c0 * u0) / (h * h);
    }
    if (M == 3) {
        const float *c = g->slap;
        return (c[0] * ((u[i - s] - u0) + (u[i + s] - u0)) + c[1] * ((u[i - 2 * s] - u0) + (u[i + 2 * s] - u0)) +
                c[2] * ((u[i - 3 * s] - u0) + (u[i + 3 * s] - u0))) / (h * h);
    }
    if (M == 4) {
        const float *c = g->slap;
        return (c[0] * ((u[i - s] - u0) + (u[i + s] - u0)) + c[1] * ((u[i - 2 * s] - u0) + (u[i + 2 * s] - u0)) +
                c[2] * ((u[i - 3 * s] - u0) + (u[i + 3 * s] - u0)) +
                c[3] * ((u[i - 4 * s] - u0) + (u[i + 4 * s] - u0))) / (h * h);
    }
    float acc = 0.f;
    for (int k = 1; k <= M; ++k) acc += g->lap[k] * ((u[i + k * s] - u0) + (u[i - k * s] - u0));
    return acc / (h * h);
}

/* centred first derivative of a[i + k*s] (gradient.cuh:47-88) */
static inline float d1(const orc_a3d *g, const float *a, long i, long s, float h)
{
    const int M = g->M;
    const float *c = g->sgrad;
#define V(k) (a[i + (long)(k) * s])
    if (M == 1) return (V(1) - V(-1)) / (2.f * h);
    if (M == 2) return (c[0] * (V(1) - V(-1)) + c[1] * (V(2) - V(-2))) / h;
    if (M == 3) return (c[0] * (V(1) - V(-1)) + c[1] * (V(2) - V(-2)) + c[2] * (V(3) - V(-3))) / h;
    if (M == 4)
        return (c[0] * (V(1) - V(-1)) + c[1] * (V(2) - V(-2)) + c[2] * (V(3) - V(-3)) + c[3] * (V(4) - V(-4))) / h;
    float acc = 0.f;
    for (int k = 1; k <= M; ++k) acc += g->grad[k] * (V(k) - V(-k));
    return acc / h;
#undef V
}

static inline int X0(const orc_a3d *g) { return g->pad[4] + g->M; }
static inline int X1(const orc_a3d *g) { return g->nx - g->pad[5] - g->M; }
static inline int Y0(const orc_a3d *g) { return g->pad[2] + g->M; }
static inline int Y1(const orc_a3d *g) { return g->ny - g->pad[3] - g->M; }
static inline int Z0(const orc_a3d *g) { return g->pad[0] + g->M; }
static inline int Z1(const orc_a3d *g) { return g->nz - g->pad[1] - g->M; }

typedef struct {
    float *u[3];
    float *psi[3], *psin[3], *zeta[3], *zetan[3];   /* axis 0 = x, 1 = y, 2 = z */
} f3;

/* ------------------------------------------------------------------ forward */
static void fwd_step(const orc_a3d *g, const float *vp, f3 *f, float *u_this)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, M = g->M;
    const float dt = g->dt;
    const float *u_prev = f->u[0], *u_now = f->u[1];
    float *u_next = f->u[2];
    const int ix0 = X0(g) + 1, ix1 = X1(g) - 1, iy0 = Y0(g) + 1, iy1 = Y1(g) - 1, iz0 = Z0(g) + 1, iz1 = Z1(g) - 1;
    const float *A[3] = {g->ax, g->ay, g->az}, *Bp[3] = {g->bx, g->by, g->bz}, *DB[3] = {g->dbx, g->dby, g->dbz};
    const long S[3] = {1, nx, (long)nx * ny};
    const float H[3] = {g->dx, g->dy, g->dz};
#pragma omp parallel for schedule(static) collapse(2)
    for (int iz = M; iz < nz - M; ++iz)
        for (int iy = M; iy < ny - M; ++iy)
            for (int ix = M; ix < nx - M; ++ix) {
                const long i = I3(iz, iy, ix);
                const float lx = lapf(g, u_now, i, 1, g->dx);
                const float ly = lapf(g, u_now, i, nx, g->dy);
                const float lz = lapf(g, u_now, i, (long)nx * ny, g->dz);
                float w_sum;
                const int interior = ix >= ix0 && ix < ix1 && iy >= iy0 && iy < iy1 && iz >= iz0 && iz < iz1;
                if (interior) {
                    w_sum = lx + ly + lz;
                } else {
                    w_sum = 0.f;
                    const int pos[3] = {ix, iy, iz};
                    const int in[3] = {ix < X0(g) + 1 || ix >= X1(g) - 1, iy < Y0(g) + 1 || iy >= Y1(g) - 1,
                                       iz < Z0(g) + 1 || iz >= Z1(g) - 1};
                    const float L[3] = {lx, ly, lz};
                    for (int a = 0; a < 3; ++a) {
                        if (!in[a]) { w_sum += L[a]; continue; }
                        const float du = d1(g, u_now, i, S[a], H[a]);
                        const float dpsi = d1(g, f->psi[a], i, S[a], H[a]);
                        const float a_ = A[a][pos[a]], b_ = Bp[a][pos[a]], db_ = DB[a][pos[a]];
                        const float da = d1(g, A[a], pos[a], 1, H[a]);
                        const float psi0 = f->psi[a][i];
                        const float tmp = ((1.f + b_) * L[a] + db_ * du) + a_ * dpsi + da * psi0;
                        w_sum += (1.f + b_) * tmp + a_ * f->zeta[a][i];
                        f->psin[a][i] = b_ * du + a_ * psi0;
                        f->zeta[a][i] = b_ * tmp + a_ * f->zeta[a][i];
                    }
                }
                const float u0 = u_now[i], v = vp[i];
                u_next[i] = 2.f * u0 - u_prev[i] + (v * v) * dt * dt * w_sum;
                if (u_this) u_this[i] = (v * v) * w_sum;
            }
}

/* ------------------------------------------------------------------ adjoint */
static void adj_step(const orc_a3d *g, const float *vp, f3 *f, const float *img, float *grad)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, M = g->M;
    const float dt = g->dt, dt2 = dt * dt;
    const float *lc = g->lap, *gc = g->grad;
    const float *l_prev = f->u[0], *l_now = f->u[1];
    float *l_next = f->u[2];
    const int c0x = X0(g) + M + 1, c1x = X1(g) - M - 1, c0y = Y0(g) + M + 1, c1y = Y1(g) - M - 1,
              c0z = Z0(g) + M + 1, c1z = Z1(g) - M - 1;
    const float *A[3] = {g->ax, g->ay, g->az}, *Bp[3] = {g->bx, g->by, g->bz}, *DB[3] = {g->dbx, g->dby, g->dbz};
    const long S[3] = {1, nx, (long)nx * ny};
    const float H[3] = {g->dx, g->dy, g->dz};
#pragma omp parallel for schedule(static) collapse(2)
    for (int iz = M; iz < nz - M; ++iz)
        for (int iy = M; iy < ny - M; ++iy)
            for (int ix = M; ix < nx - M; ++ix) {
                const long i = I3(iz, iy, ix);
                const float gun = l_now[i];
                if (grad && ix >= X0(g) && ix < X1(g) && iy >= Y0(g) && iy < Y1(g) && iz >= Z0(g) && iz < Z1(g))
                    grad[i] += 2.f * dt * dt * img[i] * gun / vp[i];
                float c0c = vp[i];
                c0c = c0c * c0c * dt2;
                const float gw = c0c * gun;
#define GW(n) (vp[n] * vp[n] * dt2 * l_now[n])
                const int inx = ix < c0x || ix >= c1x, iny = iy < c0y || iy >= c1y, inz = iz < c0z || iz >= c1z;
                if (!inx && !iny && !inz) {
                    float lapx = 0.f, lapy = 0.f, lapz = 0.f;
                    for (int k = 1; k <= M; ++k) {
                        lapx += lc[k] * ((GW(i + k) - gw) + (GW(i - k) - gw));
                        lapy += lc[k] * ((GW(i + (long)k * nx) - gw) + (GW(i - (long)k * nx) - gw));
                        lapz += lc[k] * ((GW(i + (long)k * nx * ny) - gw) + (GW(i - (long)k * nx * ny) - gw));
                    }
                    l_next[i] = 2.0f * gun - l_prev[i] + lapx * (1.0f / (g->dx * g->dx)) +
                                lapy * (1.0f / (g->dy * g->dy)) + lapz * (1.0f / (g->dz * g->dz));
                    continue;
                }
                const int pos[3] = {ix, iy, iz};
                const int in[3] = {inx, iny, inz};
                float Lt[3], Dg[3];
                for (int a = 0; a < 3; ++a) {
                    const float ih2 = 1.0f / (H[a] * H[a]), ih = 1.0f / H[a];
                    const long s = S[a];
                    if (!in[a]) {
                        float L = 0.f;
                        for (int k = 1; k <= M; ++k) L += lc[k] * ((GW(i + k * s) - gw) + (GW(i - k * s) - gw));
                        Lt[a] = L * ih2;
                        Dg[a] = 0.f;
                        continue;
                    }
                    const float *Aa = A[a], *Ba = Bp[a], *Da = DB[a];
                    const int p = pos[a];
                    const float a_ = Aa[p], b_ = Ba[p];
                    const float da = d1(g, Aa, p, 1, H[a]);
                    const float *zeta = f->zeta[a], *psi = f->psi[a];
                    const float gt0 = b_ * zeta[i] + (1.0f + b_) * gw;
                    const float l0 = (1.0f + b_) * gt0;
                    float L = 0.f, Dgg = 0.f, Dq = 0.f;
                    for (int k = 1; k <= M; ++k) {
                        const long np = i + k * s, nm = i - k * s;
                        const int cp = p + k, cm = p - k;
                        const float tp = Ba[cp] * zeta[np] + (1.0f + Ba[cp]) * (vp[np] * vp[np] * dt2) * l_now[np];
                        const float tm = Ba[cm] * zeta[nm] + (1.0f + Ba[cm]) * (vp[nm] * vp[nm] * dt2) * l_now[nm];
                        L += lc[k] * (((1.0f + Ba[cp]) * tp - l0) + ((1.0f + Ba[cm]) * tm - l0));
                        Dgg += gc[k] * ((Ba[cp] * psi[np] + Da[cp] * tp) - (Ba[cm] * psi[nm] + Da[cm] * tm));
                        Dq += gc[k] * (Aa[cp] * tp - Aa[cm] * tm);
                    }
                    Lt[a] = L * ih2;
                    Dg[a] = Dgg * ih;
                    Dq *= ih;
                    f->psin[a][i] = a_ * psi[i] + da * gt0 - Dq;
                    f->zetan[a][i] = a_ * (zeta[i] + gw);
                }
#undef GW
                l_next[i] = 2.0f * gun - l_prev[i] + (Lt[0] + Lt[1] + Lt[2] - Dg[0] - Dg[1] - Dg[2]);
            }
}

/* -------------------------------------------------- boundary-saving pieces */
static long strip_cells3(const orc_a3d *g)
{
    const long W = g->M + 1, nxp = X1(g) - X0(g), nyp = Y1(g) - Y0(g), nzp = Z1(g) - Z0(g);
    return 2 * W * nyp * nxp + 2 * W * nzp * nxp + 2 * W * nzp * nyp;
}

/* faces: z-lo, z-hi (W, nyp, nxp); y-lo, y-hi (nzp, W, nxp); x-lo, x-hi (nzp, nyp, W) */
static void strips3(const orc_a3d *g, float *u, float *slot, int restore)
{
    const int nx = g->nx, ny = g->ny, W = g->M + 1;
    const int x0 = X0(g), x1 = X1(g), y0 = Y0(g), y1 = Y1(g), z0 = Z0(g), z1 = Z1(g);
    const long nxp = x1 - x0, nyp = y1 - y0, nzp = z1 - z0;
    float *p = slot;
    for (int face = 0; face < 6; ++face) {
        const int axis = face / 2, hi = face % 2;
        long cnt = 0;
        for (int iz = z0; iz < z1; ++iz)
            for (int iy = y0; iy < y1; ++iy)
                for (int ix = x0; ix < x1; ++ix) {
                    int in;
                    if (axis == 0) in = hi ? iz >= z1 - W : iz < z0 + W;
                    else if (axis == 1) in = hi ? iy >= y1 - W : iy < y0 + W;
                    else in = hi ? ix >= x1 - W : ix < x0 + W;
                    if (!in) continue;
                    float *cell = &u[I3(iz, iy, ix)];
                    if (restore) *cell = p[cnt];
                    else p[cnt] = orc_g_strip_bf16 ? orc_bf16_round(*cell) : *cell;
                    ++cnt;
                }
        p += cnt;
    }
    (void)nzp; (void)nyp; (void)nxp;
}

static inline int in_strip3(const orc_a3d *g, int ix, int iy, int iz)
{
    const int W = g->M + 1;
    const int x0 = X0(g), x1 = X1(g), y0 = Y0(g), y1 = Y1(g), z0 = Z0(g), z1 = Z1(g);
    if (ix < x0 || ix >= x1 || iy < y0 || iy >= y1 || iz < z0 || iz >= z1) return 0;
    return iz < z0 + W || iz >= z1 - W || iy < y0 + W || iy >= y1 - W || ix < x0 + W || ix >= x1 - W;
}

static void nopml3(const orc_a3d *g, const float *vp, const float *u_prev, const float *u_now,
                   float *u_next, const float *adj, float *grad)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, M = g->M, W = M + 1;
    const float dt = g->dt;
    const int wide = g->abcn > 0 ? g->abcn + 2 * M + 1 : 2 * M;
    const int gx0 = X0(g) + W, gx1 = X1(g) - W, gy0 = Y0(g) + W, gy1 = Y1(g) - W, gz0 = Z0(g) + W, gz1 = Z1(g) - W;
#pragma omp parallel for schedule(static) collapse(2)
    for (int iz = wide; iz < nz - wide; ++iz)
        for (int iy = wide; iy < ny - wide; ++iy)
            for (int ix = wide; ix < nx - wide; ++ix) {
                const long i = I3(iz, iy, ix);
                const float w_sum = lapf(g, u_now, i, 1, g->dx) + lapf(g, u_now, i, nx, g->dy) +
                                    lapf(g, u_now, i, (long)nx * ny, g->dz);
                const float u0 = u_now[i], v = vp[i], up = u_prev[i];
                const float un = 2.0f * u0 - up + (v * v) * dt * dt * w_sum;
                u_next[i] = un;
                if (grad && ix >= gx0 && ix < gx1 && iy >= gy0 && iy < gy1 && iz >= gz0 && iz < gz1) {
                    const float u_tt = (un - 2 * u0 + up) / (dt * dt);
                    grad[i] += 2.f * dt * dt * u_tt * adj[i] / v;
                }
            }
}

static void band3(const orc_a3d *g, const float *fnext, const float *fnow, const float *fprev,
                  const float *adj, const float *vp, float *grad)
{
    const int nx = g->nx, ny = g->ny;
    const float dt = g->dt;
    for (int iz = Z0(g); iz < Z1(g); ++iz)
        for (int iy = Y0(g); iy < Y1(g); ++iy)
            for (int ix = X0(g); ix < X1(g); ++ix) {
                if (!in_strip3(g, ix, iy, iz)) continue;
                const long i = I3(iz, iy, ix);
                const float u_tt = (fnow[i] - 2 * fprev[i] + fnext[i]) / (dt * dt);
                grad[i] += 2.f * dt * dt * u_tt * adj[i] / vp[i];
            }
}

static void illum3(const orc_a3d *g, const float *fnext, const float *fnow, const float *fprev,
                   const float *adj, float *S, float *R)
{
    const int nx = g->nx, ny = g->ny;
    const float dt = g->dt;
    for (int iz = Z0(g); iz < Z1(g); ++iz)
        for (int iy = Y0(g); iy < Y1(g); ++iy)
            for (int ix = X0(g); ix < X1(g); ++ix) {
                const long i = I3(iz, iy, ix);
                if (R) { const float ub = adj[i]; R[i] += ub * ub; }
                if (!S) continue;
                const float u_tt = fnow ? (fnow[i] - 2 * fprev[i] + fnext[i]) / (dt * dt) : fnext[i];
                S[i] += u_tt * u_tt;
            }
}

static void pts_add(const orc_a3d *g, float *u, const float *w, const int *loc, int n, int it, float sign,
                    int strip_only)
{
    for (int s = 0; s < n; ++s) {
        const int ix = loc[3 * s], iy = loc[3 * s + 1], iz = loc[3 * s + 2];
        if (ix < 0 || ix >= g->nx || iy < 0 || iy >= g->ny || iz < 0 || iz >= g->nz) continue;
        if (strip_only && !in_strip3(g, ix, iy, iz)) continue;
        u[((long)iz * g->ny + iy) * g->nx + ix] += sign * w[(long)s * g->nt + it];
    }
}

static void pts_get(const orc_a3d *g, const float *u, float *out, const int *loc, int n, int it, int acc)
{
    for (int r = 0; r < n; ++r) {
        const int ix = loc[3 * r], iy = loc[3 * r + 1], iz = loc[3 * r + 2];
        if (ix < 0 || ix >= g->nx || iy < 0 || iy >= g->ny || iz < 0 || iz >= g->nz) continue;
        const float v = u[((long)iz * g->ny + iy) * g->nx + ix];
        if (acc) out[(long)r * g->nt + it] += v; else out[(long)r * g->nt + it] = v;
    }
}

/* ------------------------------------------------------------------ drivers */
static float *zal(long n) { return (float *)calloc((size_t)n, sizeof(float)); }

static void f3_alloc(f3 *f, long n)
{
    for (int k = 0; k < 3; ++k) {
        f->u[k] = zal(n);
        f->psi[k] = zal(n); f->psin[k] = zal(n); f->zeta[k] = zal(n); f->zetan[k] = zal(n);
    }
}

static void f3_free(f3 *f)
{
    for (int k = 0; k < 3; ++k) {
        free(f->u[k]); free(f->psi[k]); free(f->psin[k]); free(f->zeta[k]); free(f->zetan[k]);
    }
}

static void rot(f3 *f, int swap_zeta)
{
    float *p = f->u[0];
    f->u[0] = f->u[1]; f->u[1] = f->u[2]; f->u[2] = p;
    for (int a = 0; a < 3; ++a) {
        float *t = f->psi[a]; f->psi[a] = f->psin[a]; f->psin[a] = t;
        if (swap_zeta) { t = f->zeta[a]; f->zeta[a] = f->zetan[a]; f->zetan[a] = t; }
    }
}

static void fwd_shot3(const orc_a3d *g, const float *vp, const float *wav, const int *sloc, int nsrc,
                      const int *rloc, int nrec, float *rec, float *u_allt, float *strips, float *last_two)
{
    const long n = (long)g->nz * g->ny * g->nx, ns = strip_cells3(g);
    f3 f;
    f3_alloc(&f, n);
    for (int it = 0; it < g->nt; ++it) {
        fwd_step(g, vp, &f, u_allt ? u_allt + (long)it * n : NULL);
        if (strips) strips3(g, f.u[1], strips + (long)it * ns, 0);
        pts_add(g, f.u[2], wav, sloc, nsrc, it, 1.0f, 0);
        if (rec) pts_get(g, f.u[2], rec, rloc, nrec, it, 0);
        rot(&f, 0);
    }
    if (last_two) {
        memcpy(last_two, f.u[0], (size_t)n * sizeof(float));
        memcpy(last_two + n, f.u[1], (size_t)n * sizeof(float));
    }
    f3_free(&f);
}

static void adj_once3(const orc_a3d *g, const float *vp, f3 *a, const float *res, const int *rloc, int nrec,
                      const int *sloc, int nsrc, float *gw, int it, const float *img, float *grad)
{
    adj_step(g, vp, a, img, grad);
    pts_add(g, a->u[2], res, rloc, nrec, it, 1.0f, 0);
    rot(a, 1);
    if (gw) pts_get(g, a->u[1], gw, sloc, nsrc, it, 1);
}

void orc_a3d_forward(const orc_a3d *g, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, float *rec)
{
    const long n = (long)g->nz * g->ny * g->nx;
    for (int b = 0; b < g->B; ++b)
        fwd_shot3(g, vp + (g->vp_batched ? b * n : 0), wav + (long)b * nsrc * g->nt, sloc + (long)b * nsrc * 3,
                  nsrc, rloc + (long)b * nrec * 3, nrec, rec + (long)b * nrec * g->nt, NULL, NULL, NULL);
}

int orc_a3d_gradient(const orc_a3d *g, int mode, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, const float *res, float *rec, float *grad_vp, float *grad_wav,
                     float *illum_s, float *illum_r)
{
    const long n = (long)g->nz * g->ny * g->nx, ns = strip_cells3(g);
    const int nx = g->nx, ny = g->ny;
    for (int b = 0; b < g->B; ++b) {
        const float *vpb = vp + (g->vp_batched ? b * n : 0);
        const float *wb = wav + (long)b * nsrc * g->nt;
        const int *sl = sloc + (long)b * nsrc * 3, *rl = rloc + (long)b * nrec * 3;
        const float *rb = res + (long)b * nrec * g->nt;
        float *gv = grad_vp + b * n, *gw = grad_wav ? grad_wav + (long)b * nsrc * g->nt : NULL;
        float *S = illum_s ? illum_s + b * n : NULL, *R = illum_r ? illum_r + b * n : NULL;
        float *recb = rec ? rec + (long)b * nrec * g->nt : NULL;
        f3 a;
        f3_alloc(&a, n);
        if (mode == 0) {
            float *u_allt = zal((long)g->nt * n);
            if (!u_allt) { f3_free(&a); return -1; }
            fwd_shot3(g, vpb, wb, sl, nsrc, rl, nrec, recb, u_allt, NULL, NULL);
            for (int it = g->nt - 1; it >= 0; --it) {
                const float *img = (it + 1 < g->nt) ? u_allt + (long)(it + 1) * n : NULL;
                adj_once3(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, it, img, img ? gv : NULL);
                if (S || R) illum3(g, u_allt + (long)it * n, NULL, NULL, a.u[1], S, R);
            }
            for (int iz = Z0(g); iz < Z1(g); ++iz)
                for (int iy = Y0(g); iy < Y1(g); ++iy)
                    for (int ix = X0(g); ix < X1(g); ++ix) {
                        const long i = I3(iz, iy, ix);
                        gv[i] += 2.f * g->dt * g->dt * u_allt[i] * a.u[1][i] / vpb[i];
                    }
            free(u_allt);
        } else {
            float *strips = zal((long)g->nt * ns), *last_two = zal(2 * n);
            float *r[3] = {zal(n), zal(n), zal(n)};
            if (!strips || !last_two || !r[0] || !r[1] || !r[2]) {
                free(strips); free(last_two); free(r[0]); free(r[1]); free(r[2]); f3_free(&a);
                return -1;
            }
            fwd_shot3(g, vpb, wb, sl, nsrc, rl, nrec, recb, NULL, strips, last_two);
            memcpy(r[0], last_two + n, (size_t)n * sizeof(float));
            memcpy(r[1], last_two, (size_t)n * sizeof(float));
            for (int it = g->nt - 1; it >= 1; --it) {
                adj_once3(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, it, NULL, NULL);
                nopml3(g, vpb, r[0], r[1], r[2], a.u[1], gv);
                strips3(g, r[2], strips + (long)(it - 1) * ns, 1);
                pts_add(g, r[2], wb, sl, nsrc, it, -1.0f, 1);
                band3(g, r[0], r[2], r[1], a.u[1], vpb, gv);
                if (S || R) illum3(g, r[0], r[2], r[1], a.u[1], S, R);
                pts_add(g, r[2], wb, sl, nsrc, it, 1.0f, 0);
                float *p = r[0]; r[0] = r[1]; r[1] = r[2]; r[2] = p;
            }
            if (g->nt > 0) {
                adj_once3(g, vpb, &a, rb, rl, nrec, sl, nsrc, gw, 0, NULL, NULL);
                if (R) illum3(g, NULL, NULL, NULL, a.u[1], NULL, R);
            }
            free(r[0]); free(r[1]); free(r[2]); free(strips); free(last_two);
        }
        f3_free(&a);
    }
    return 0;
}
