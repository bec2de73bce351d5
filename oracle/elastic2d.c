/*
 * ORACLE -- test infrastructure only (see acoustic2d.c for the rules).
 *
 * CPU restatement of the reference's 2-D isotropic elastic path
 * (velocity-stress, Virieux staggered grid, cpmls CPML), no free surface:
 *   velocity step ........................ csrc/cuda/equations/elastic2d/kernels.cuh:48-146
 *   stress step ........................... kernels.cuh:182-330
 *   staggered first derivatives ........... csrc/cuda/operators/staggered.cuh:22-520
 *   Lame parameters ....................... csrc/cuda/common/derived_models.cu:27-37
 *   source / record / signed residual ..... csrc/cuda/common/common.cu:11-37, 119-144, 182-210
 *   stress / velocity adjoint prepare+apply kernels.cuh:547-1055
 *   vp / vs / rho imaging ................. kernels.cuh:1183-1260 (calculate_grad_elastic_nobs)
 *   body-force rho corrections ............ common.cu:40-116
 *   forward / full-storage backward order . csrc/cuda/common/sg_driver.cuh:23-50, 285-340,
 *                                           equations/elastic2d/driver_traits.cuh:380-630
 *   boundary-saving backward ............. sg_driver.cuh:812: the same reverse steps for
 *                                           it = nt-1 .. 1 (no step 0); its reconstruction of
 *                                           the forward state is exact in exact arithmetic, so
 *                                           the oracle images against the stored velocities
 * Field indices: 0 vx, 1 vz, 2 sxx, 3 szz, 4 sxz.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define I2(iz, ix) ((long)(iz) * nx + (ix))

enum { VX = 0, VZ, SXX, SZZ, SXZ, M_VXX, M_VZZ, M_VXZ, M_VZX, M_SXXX, M_SZZZ, M_SXZX, M_SXZZ, NSTATE };

/* staggered first derivative along stride s at cell i; fwd: forward type
 * (u[i+1..] - u[i..]), else backward type (staggered.cuh) */
static inline float sg(const orc_e2d *g, const float *u, long i, long s, int fwd, float h)
{
    const int M = g->M;
    const float inv = 1.f / h;
#define U(k) (u[i + (long)(k) * s])
    if (M == 1) return fwd ? (U(1) - U(0)) * inv : (U(0) - U(-1)) * inv;
    if (M == 2) {
        const float c1 = 9.f / 8.f, c2 = -1.f / 24.f;
        return fwd ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2))) * inv;
    }
    if (M == 3) {
        const float c1 = 75.f / 64.f, c2 = -25.f / 384.f, c3 = 3.f / 640.f;
        return fwd ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1)) + c3 * (U(3) - U(-2))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2)) + c3 * (U(2) - U(-3))) * inv;
    }
    if (M == 4) {
        const float c1 = 1225.f / 1024.f, c2 = -245.f / 3072.f, c3 = 49.f / 5120.f, c4 = -5.f / 7168.f;
        return fwd ? (c1 * (U(1) - U(0)) + c2 * (U(2) - U(-1)) + c3 * (U(3) - U(-2)) + c4 * (U(4) - U(-3))) * inv
                   : (c1 * (U(0) - U(-1)) + c2 * (U(1) - U(-2)) + c3 * (U(2) - U(-3)) + c4 * (U(3) - U(-4))) * inv;
    }
    float acc = 0.f;
    for (int m = 0; m < M; ++m)
        acc += fwd ? g->coef[m] * (U(m + 1) - U(-m)) : g->coef[m] * (U(m) - U(-m - 1));
    return acc * inv;
#undef U
}

static inline int in_pml(const orc_e2d *g, int ix, int iz)
{
    const int M = g->M;
    return ix < g->pad[2] + M || ix >= g->nx - g->pad[3] - M || iz < g->pad[0] + M || iz >= g->nz - g->pad[1] - M;
}

static void lame(long n, const float *vp, const float *vs, const float *rho, float *mu, float *lam)
{
    for (long i = 0; i < n; ++i) {
        const float v = vp[i], s = vs[i], r = rho[i];
        mu[i] = (r * s) * s;
        lam[i] = r * ((v * v) - (2.0f * s) * s);
    }
}

/* ------------------------------------------------------------------ forward */
static void velocity_step(const orc_e2d *g, float **f, const float *rho)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt;
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz)
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = I2(iz, ix);
            float dsxx_dx = sg(g, f[SXX], i, 1, 1, g->dx);
            float dsxz_dz = sg(g, f[SXZ], i, nx, 0, g->dz);
            float dsxz_dx = sg(g, f[SXZ], i, 1, 0, g->dx);
            float dszz_dz = sg(g, f[SZZ], i, nx, 1, g->dz);
            const float inv_rho = 1.f / rho[i];
            if (in_pml(g, ix, iz)) {
                float m;
                m = g->azh[iz] * f[M_SZZZ][i] + g->bzh[iz] * dszz_dz;
                f[M_SZZZ][i] = m;
                dszz_dz += m;
                m = g->ax[ix] * f[M_SXZX][i] + g->bx[ix] * dsxz_dx;
                f[M_SXZX][i] = m;
                dsxz_dx += m;
                m = g->az[iz] * f[M_SXZZ][i] + g->bz[iz] * dsxz_dz;
                f[M_SXZZ][i] = m;
                dsxz_dz += m;
                m = g->axh[ix] * f[M_SXXX][i] + g->bxh[ix] * dsxx_dx;
                f[M_SXXX][i] = m;
                dsxx_dx += m;
            }
            f[VX][i] += dt * inv_rho * (dsxx_dx + dsxz_dz);
            f[VZ][i] += dt * inv_rho * (dsxz_dx + dszz_dz);
        }
}

static void stress_step(const orc_e2d *g, float **f, const float *lam, const float *mu, float *u_this_x,
                        float *u_this_z)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt;
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz)
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = I2(iz, ix);
            float dvx_dx = sg(g, f[VX], i, 1, 0, g->dx);
            float dvz_dz = sg(g, f[VZ], i, nx, 0, g->dz);
            float dvx_dz = sg(g, f[VX], i, nx, 1, g->dz);
            float dvz_dx = sg(g, f[VZ], i, 1, 1, g->dx);
            const float la = lam[i], m_ = mu[i];
            if (!in_pml(g, ix, iz)) {
                f[SXX][i] += dt * ((la + 2.f * m_) * dvx_dx + la * dvz_dz);
                f[SZZ][i] += dt * ((la + 2.f * m_) * dvz_dz + la * dvx_dx);
                f[SXZ][i] += dt * m_ * (dvx_dz + dvz_dx);
            } else {
                float m;
                m = g->az[iz] * f[M_VZZ][i] + g->bz[iz] * dvz_dz;
                f[M_VZZ][i] = m;
                dvz_dz += m;
                m = g->ax[ix] * f[M_VXX][i] + g->bx[ix] * dvx_dx;
                f[M_VXX][i] = m;
                dvx_dx += m;
                f[SXX][i] += dt * ((la + 2.f * m_) * dvx_dx + la * dvz_dz);
                f[SZZ][i] += dt * ((la + 2.f * m_) * dvz_dz + la * dvx_dx);
                m = g->azh[iz] * f[M_VXZ][i] + g->bzh[iz] * dvx_dz;
                f[M_VXZ][i] = m;
                dvx_dz += m;
                m = g->axh[ix] * f[M_VZX][i] + g->bxh[ix] * dvz_dx;
                f[M_VZX][i] = m;
                dvz_dx += m;
                f[SXZ][i] += dt * m_ * (dvx_dz + dvz_dx);
            }
            if (u_this_x) {
                u_this_x[i] = f[VX][i];
                u_this_z[i] = f[VZ][i];
            }
        }
}

static void add_points(const orc_e2d *g, float *u, const float *w, const int *loc, int n, int it, float sign)
{
    for (int s = 0; s < n; ++s) {
        const int ix = loc[2 * s], iz = loc[2 * s + 1];
        if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
        const float v = w[(long)s * g->nt + it];
        u[(long)iz * g->nx + ix] += sign < 0.f ? -v : v;
    }
}

static void get_points(const orc_e2d *g, const float *u, float *out, const int *loc, int n, int it)
{
    for (int r = 0; r < n; ++r) {
        const int ix = loc[2 * r], iz = loc[2 * r + 1];
        if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
        out[(long)r * g->nt + it] = u[(long)iz * g->nx + ix];
    }
}

static int alloc_state(float **f, long n)
{
    for (int k = 0; k < NSTATE; ++k) {
        f[k] = (float *)calloc((size_t)n, sizeof(float));
        if (!f[k]) return -1;
    }
    return 0;
}

static void free_state(float **f)
{
    for (int k = 0; k < NSTATE; ++k) free(f[k]);
}

/* forward of one shot; rec (nrf, nrec, nt) of this shot with field stride
 * rec_fstride; u_allt (nt, 2, n) when non-null */
static int fwd_shot(const orc_e2d *g, const float *rho, const float *lam, const float *mu, const float *wav,
                    const int *sloc, int nsrc, const int *rloc, int nrec, float *rec, long rec_fstride,
                    float *u_allt)
{
    const long n = (long)g->nz * g->nx;
    float *f[NSTATE];
    if (alloc_state(f, n)) { free_state(f); return -1; }
    for (int it = 0; it < g->nt; ++it) {
        velocity_step(g, f, rho);
        stress_step(g, f, lam, mu, u_allt ? u_allt + (long)it * 2 * n : NULL,
                    u_allt ? u_allt + (long)it * 2 * n + n : NULL);
        for (int k = 0; k < g->nsf; ++k) add_points(g, f[g->src_fields[k]], wav, sloc, nsrc, it, 1.f);
        if (rec)
            for (int k = 0; k < g->nrf; ++k)
                get_points(g, f[g->rec_fields[k]], rec + (long)k * rec_fstride, rloc, nrec, it);
    }
    free_state(f);
    return 0;
}

/* ------------------------------------------------------------------ adjoint */
/* stress adjoint prepare + apply, velocity adjoint prepare + apply
 * (kernels.cuh:547-1055; the prepares cover every cell, the applies the
 * computed box [M, n-M)) */
static void adjoint_step(const orc_e2d *g, float **a, float **qp, const float *rho, const float *lam,
                         const float *mu)
{
    const int nx = g->nx, nz = g->nz, M = g->M;
    const float dt = g->dt;
    float *qxx = qp[0], *qzz = qp[1], *qxz = qp[2], *qzx = qp[3];
    float *pxx = qp[4], *pzz = qp[5], *pxz = qp[6], *pzx = qp[7];
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix) {
            const long i = I2(iz, ix);
            const float la = lam[i], m_ = mu[i];
            const float bar_sxx = a[SXX][i], bar_szz = a[SZZ][i], bar_sxz = a[SXZ][i];
            const float bar_dvx_dx = dt * ((la + 2.f * m_) * bar_sxx + la * bar_szz);
            const float bar_dvz_dz = dt * ((la + 2.f * m_) * bar_szz + la * bar_sxx);
            const float bar_dvx_dz = dt * m_ * bar_sxz;
            const float bar_dvz_dx = dt * m_ * bar_sxz;
            if (!in_pml(g, ix, iz)) {
                qxx[i] = bar_dvx_dx;
                qzz[i] = bar_dvz_dz;
                qxz[i] = bar_dvx_dz;
                qzx[i] = bar_dvz_dx;
                continue;
            }
            const float t_vxx = a[M_VXX][i] + bar_dvx_dx, t_vzz = a[M_VZZ][i] + bar_dvz_dz;
            const float t_vxz = a[M_VXZ][i] + bar_dvx_dz, t_vzx = a[M_VZX][i] + bar_dvz_dx;
            qxx[i] = bar_dvx_dx + g->bx[ix] * t_vxx;
            qzz[i] = bar_dvz_dz + g->bz[iz] * t_vzz;
            qxz[i] = bar_dvx_dz + g->bzh[iz] * t_vxz;
            qzx[i] = bar_dvz_dx + g->bxh[ix] * t_vzx;
            a[M_VXX][i] = g->ax[ix] * t_vxx;
            a[M_VZX][i] = g->axh[ix] * t_vzx;
            a[M_VZZ][i] = g->az[iz] * t_vzz;
            a[M_VXZ][i] = g->azh[iz] * t_vxz;
        }
    /* stress apply: the transposes of the stress step's derivatives */
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz)
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = I2(iz, ix);
            const float dqxx_dx = sg(g, qxx, i, 1, 1, g->dx);
            const float dqxz_dz = sg(g, qxz, i, nx, 0, g->dz);
            const float dqzx_dx = sg(g, qzx, i, 1, 0, g->dx);
            const float dqzz_dz = sg(g, qzz, i, nx, 1, g->dz);
            a[VX][i] += dqxx_dx + dqxz_dz;
            a[VZ][i] += dqzx_dx + dqzz_dz;
        }
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix) {
            const long i = I2(iz, ix);
            const float inv_rho = 1.f / rho[i];
            const float bar_dsxx_dx = dt * inv_rho * a[VX][i], bar_dsxz_dz = dt * inv_rho * a[VX][i];
            const float bar_dsxz_dx = dt * inv_rho * a[VZ][i], bar_dszz_dz = dt * inv_rho * a[VZ][i];
            if (!in_pml(g, ix, iz)) {
                pxx[i] = bar_dsxx_dx;
                pxz[i] = bar_dsxz_dz;
                pzx[i] = bar_dsxz_dx;
                pzz[i] = bar_dszz_dz;
                continue;
            }
            const float t_sxxx = a[M_SXXX][i] + bar_dsxx_dx, t_sxzz = a[M_SXZZ][i] + bar_dsxz_dz;
            const float t_sxzx = a[M_SXZX][i] + bar_dsxz_dx, t_szzz = a[M_SZZZ][i] + bar_dszz_dz;
            pxx[i] = bar_dsxx_dx + g->bxh[ix] * t_sxxx;
            pxz[i] = bar_dsxz_dz + g->bz[iz] * t_sxzz;
            pzx[i] = bar_dsxz_dx + g->bx[ix] * t_sxzx;
            pzz[i] = bar_dszz_dz + g->bzh[iz] * t_szzz;
            a[M_SXXX][i] = g->axh[ix] * t_sxxx;
            a[M_SXZX][i] = g->ax[ix] * t_sxzx;
            a[M_SXZZ][i] = g->az[iz] * t_sxzz;
            a[M_SZZZ][i] = g->azh[iz] * t_szzz;
        }
#pragma omp parallel for schedule(static)
    for (int iz = M; iz < nz - M; ++iz)
        for (int ix = M; ix < nx - M; ++ix) {
            const long i = I2(iz, ix);
            const float dpxx_dx = sg(g, pxx, i, 1, 0, g->dx);
            const float dpzz_dz = sg(g, pzz, i, nx, 0, g->dz);
            const float dpxz_dz = sg(g, pxz, i, nx, 1, g->dz);
            const float dpzx_dx = sg(g, pzx, i, 1, 1, g->dx);
            a[SXX][i] += dpxx_dx;
            a[SZZ][i] += dpzz_dz;
            a[SXZ][i] += dpxz_dz + dpzx_dx;
        }
}

/* vp / vs / rho imaging over the physical box (calculate_grad_elastic_nobs) */
static void image(const orc_e2d *g, float **a, const float *fvx, const float *fvz, const float *fvx_prev,
                  const float *fvz_prev, const float *vp, const float *vs, const float *rho, float *gvp,
                  float *gvs, float *grho)
{
    const int nx = g->nx, M = g->M;
    const float dt = g->dt;
    const int x0 = g->pad[2] + M, x1 = g->nx - g->pad[3] - M, z0 = g->pad[0] + M, z1 = g->nz - g->pad[1] - M;
#pragma omp parallel for schedule(static)
    for (int iz = z0; iz < z1; ++iz)
        for (int ix = x0; ix < x1; ++ix) {
            const long i = I2(iz, ix);
            const float fvx_x = sg(g, fvx, i, 1, 0, g->dx);
            const float fvz_z = sg(g, fvz, i, nx, 0, g->dz);
            const float fvx_z = sg(g, fvx, i, nx, 1, g->dz);
            const float fvz_x = sg(g, fvz, i, 1, 1, g->dx);
            const float bar_sxx = a[SXX][i], bar_szz = a[SZZ][i], bar_sxz = a[SXZ][i];
            const float grad_lambda = (bar_sxx + bar_szz) * (fvx_x + fvz_z);
            float grad_mu = 2.f * (bar_sxx * fvx_x + bar_szz * fvz_z);
            grad_mu += bar_sxz * (fvx_z + fvz_x);
            gvp[i] += -2.f * rho[i] * vp[i] * grad_lambda * dt;
            gvs[i] += -(-4.f * rho[i] * vs[i] * grad_lambda + 2.f * rho[i] * vs[i] * grad_mu) * dt;
            grho[i] += (a[VX][i] * (fvx[i] - fvx_prev[i]) + a[VZ][i] * (fvz[i] - fvz_prev[i])) / rho[i];
            grho[i] -= grad_lambda * (vp[i] * vp[i] - 2.f * vs[i] * vs[i]) * dt + grad_mu * (vs[i] * vs[i]) * dt;
        }
}

void orc_e2d_forward(const orc_e2d *g, const float *vp, const float *vs, const float *rho, const float *wav,
                     const int *sloc, int nsrc, const int *rloc, int nrec, float *rec)
{
    const long n = (long)g->nz * g->nx;
    float *mu = (float *)malloc((size_t)n * sizeof(float)), *lam = (float *)malloc((size_t)n * sizeof(float));
    for (int b = 0; b < g->B; ++b) {
        const long mo = g->vp_batched ? (long)b * n : 0;
        lame(n, vp + mo, vs + mo, rho + mo, mu, lam);
        fwd_shot(g, rho + mo, lam, mu, wav + (long)b * nsrc * g->nt, sloc + (long)b * nsrc * 2, nsrc,
                 rloc + (long)b * nrec * 2, nrec, rec + (long)b * nrec * g->nt, (long)g->B * nrec * g->nt, NULL);
    }
    free(mu);
    free(lam);
}

int orc_e2d_gradient(const orc_e2d *g, const float *vp, const float *vs, const float *rho, const float *wav,
                     const int *sloc, int nsrc, const int *rloc, int nrec, const float *res, float *rec,
                     float *gvp, float *gvs, float *grho)
{
    const long n = (long)g->nz * g->nx;
    const int nt = g->nt;
    const long fstride = (long)g->B * nrec * nt;   /* (nfield, B, nrec, nt) */
    float *mu = (float *)malloc((size_t)n * sizeof(float)), *lam = (float *)malloc((size_t)n * sizeof(float));
    float *u_allt = (float *)calloc((size_t)nt * 2 * n, sizeof(float));
    float *zero = (float *)calloc((size_t)n, sizeof(float));
    float *a[NSTATE], *qp[8];
    int rc = (mu && lam && u_allt && zero) ? 0 : -1;
    for (int k = 0; k < 8; ++k) {
        qp[k] = (float *)calloc((size_t)n, sizeof(float));
        if (!qp[k]) rc = -1;
    }
    for (int b = 0; b < g->B && rc == 0; ++b) {
        const long mo = g->vp_batched ? (long)b * n : 0;
        const float *vpb = vp + mo, *vsb = vs + mo, *rhob = rho + mo;
        const float *wb = wav + (long)b * nsrc * nt;
        const int *sl = sloc + (long)b * nsrc * 2, *rl = rloc + (long)b * nrec * 2;
        float *gv = gvp + (long)b * n, *gs = gvs + (long)b * n, *gr = grho + (long)b * n;
        lame(n, vpb, vsb, rhob, mu, lam);
        memset(u_allt, 0, (size_t)nt * 2 * n * sizeof(float));
        if (fwd_shot(g, rhob, lam, mu, wb, sl, nsrc, rl, nrec, rec ? rec + (long)b * nrec * nt : NULL, fstride,
                     u_allt)) { rc = -1; break; }
        if (alloc_state(a, n)) { free_state(a); rc = -1; break; }
        for (int it = nt - 1; it >= (g->bs ? 1 : 0); --it) {
            /* body-force rho correction at the velocity sources, before the residual (common.cu:40-75) */
            for (int k = 0; k < g->nsf; ++k) {
                const int sf = g->src_fields[k];
                if (sf > 1) continue;
                for (int s = 0; s < nsrc; ++s) {
                    const int ix = sl[2 * s], iz = sl[2 * s + 1];
                    if (ix < 0 || ix >= g->nx || iz < 0 || iz >= g->nz) continue;
                    const long i = (long)iz * g->nx + ix;
                    gr[i] += a[sf][i] * wb[(long)s * nt + it] / rhob[i];
                }
            }
            /* signed residuals: stress fields take -1 (elastic.h:696-713) */
            for (int k = 0; k < g->nrf; ++k) {
                const int rf = g->rec_fields[k];
                add_points(g, a[rf], res + (long)k * fstride + (long)b * nrec * nt, rl, nrec, it,
                           rf >= 2 ? -1.f : 1.f);
            }
            const float *vxn = u_allt + (long)it * 2 * n, *vzn = vxn + n;
            const float *vxp = it + 1 < nt ? u_allt + (long)(it + 1) * 2 * n : zero;
            const float *vzp = it + 1 < nt ? vxp + n : zero;
            image(g, a, vxn, vzn, vxp, vzp, vpb, vsb, rhob, gv, gs, gr);
            if (it > 0) adjoint_step(g, a, qp, rhob, lam, mu);
            /* receiver rho correction, velocity receivers inside [M, n-M) (common.cu:77-116) */
            for (int k = 0; k < g->nrf; ++k) {
                const int rf = g->rec_fields[k];
                if (rf > 1) continue;
                const float *fn = rf == 0 ? vxn : vzn, *fp = rf == 0 ? vxp : vzp;
                const float *rr = res + (long)k * fstride + (long)b * nrec * nt;
                for (int r = 0; r < nrec; ++r) {
                    const int ix = rl[2 * r], iz = rl[2 * r + 1];
                    if (ix < g->M || ix >= g->nx - g->M || iz < g->M || iz >= g->nz - g->M) continue;
                    const long i = (long)iz * g->nx + ix;
                    gr[i] += -rr[(long)r * nt + it] * (fn[i] - fp[i]) / rhob[i];
                }
            }
        }
        free_state(a);
    }
    for (int k = 0; k < 8; ++k) free(qp[k]);
    free(mu);
    free(lam);
    free(u_allt);
    free(zero);
    return rc;
}
