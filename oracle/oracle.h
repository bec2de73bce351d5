/* ORACLE -- test infrastructure only (see acoustic2d.c). */
#ifndef SWEEP_AMD_ORACLE_H
#define SWEEP_AMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One 2-D acoustic problem on the runtime grid (nz, nx) = physical + pad +
 * 2M.  Arrays are row-major (z slow, x fast); profiles have the runtime
 * length of their axis. */
typedef struct {
    int nz, nx, B, M, nt;
    int vp_batched;       /* 1: vp is (B, nz, nx); 0: one shared (nz, nx) */
    int abcn;             /* rim width of the reconstruction seed = abcn + M */
    int pad[4];           /* PML pad per face: z_lo, z_hi, x_lo, x_hi */
    float dt, dx, dz;
    const float *az, *bz, *dbz, *ax, *bx, *dbx;
    const float *lap;     /* (M+1): lap[0] = -c0, lap[1..M] (zero-sum float32) */
    const float *grad;    /* (M+1): grad[1..M] centred first-derivative weights */
    const float *slap;    /* fixed-order forward Laplacian weights c1..cM (M <= 4) */
    float slap_c0;        /* centre weight of the order-4 summed form */
    const float *sgrad;   /* fixed-order first-derivative weights (M <= 4) */
} orc_a2d;

void orc_a2d_forward(const orc_a2d *g, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, float *rec);

/* Forward + adjoint of every shot for the adjoint source res (B, nrec, nt).
 * mode 0: full storage; 1: boundary saving.  grad_vp (B, nz, nx) and
 * grad_wav (B, nsrc, nt) accumulate; rec/illum_* may be NULL. */
int orc_a2d_gradient(const orc_a2d *g, int mode, const float *vp, const float *wav, const int *sloc,
                     int nsrc, const int *rloc, int nrec, const float *res, float *rec,
                     float *grad_vp, float *grad_wav, float *illum_s, float *illum_r);

long orc_a2d_strip_cells(const orc_a2d *g);

/* Test hook: 1 = eager-path edge treatment in the forward (see acoustic2d.c). */
void orc_set_edge_mode(int mode);

/* 1: boundary-saving strips are stored as bfloat16, rounded to nearest with
 * ties to even and a NaN kept a quiet NaN -- the reference's
 * __float2bfloat16 (boundarysaver.cu:29; its host stand-in
 * csrc/host/include/cuda_bf16.h) -- both for the 2-D and the 3-D oracle. */
void orc_set_strip_bf16(int on);
/* x rounded to bfloat16 and widened back to float. */
float orc_bf16_round(float x);

void orc_a2d_fwd_step(const orc_a2d *g, const float *vp, const float *u_prev, const float *u_now,
                      float *u_next, const float *psix, const float *psiz, float *psixn,
                      float *psizn, float *zetax, float *zetaz, float *u_this);

void orc_a2d_adj_step(const orc_a2d *g, const float *vp, const float *l_prev, const float *l_now,
                      float *l_next, const float *psix, const float *psiz, const float *zetax,
                      const float *zetaz, float *psixo, float *psizo, float *zetaxo,
                      float *zetazo, const float *img, float *grad);

/* One 3-D acoustic problem on the runtime grid (nz, ny, nx); x fastest. */
typedef struct {
    int nz, ny, nx, B, M, nt;
    int vp_batched, abcn;
    int pad[6];           /* z_lo, z_hi, y_lo, y_hi, x_lo, x_hi */
    float dt, dx, dy, dz;
    const float *az, *bz, *dbz, *ay, *by, *dby, *ax, *bx, *dbx;
    const float *lap, *grad, *slap;
    float slap_c0;
    const float *sgrad;
} orc_a3d;

void orc_a3d_forward(const orc_a3d *g, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, float *rec);
int orc_a3d_gradient(const orc_a3d *g, int mode, const float *vp, const float *wav, const int *sloc, int nsrc,
                     const int *rloc, int nrec, const float *res, float *rec, float *grad_vp, float *grad_wav,
                     float *illum_s, float *illum_r);

/* One 2-D isotropic elastic problem (velocity-stress, cpmls) on the runtime
 * grid (nz, nx); fields 0 vx, 1 vz, 2 sxx, 3 szz, 4 sxz. */
typedef struct {
    int nz, nx, B, M, nt;
    int vp_batched;       /* 1: models are (B, nz, nx); 0: shared */
    int pad[4];           /* z_lo, z_hi, x_lo, x_hi */
    float dt, dx, dz;
    const float *az, *bz, *azh, *bzh, *ax, *bx, *axh, *bxh;
    const float *coef;    /* staggered weights c_1..c_M (index 0..M-1), used for M > 4 */
    int nsf, nrf;
    const int *src_fields, *rec_fields;
    int bs;               /* 1: the boundary-saving backward's reverse loop, it = nt-1 .. 1 */
} orc_e2d;

/* rec (nrf, B, nrec, nt) */
void orc_e2d_forward(const orc_e2d *g, const float *vp, const float *vs, const float *rho, const float *wav,
                     const int *sloc, int nsrc, const int *rloc, int nrec, float *rec);
/* full-storage gradient for the residual res (nrf, B, nrec, nt); gradients (B, nz, nx) accumulate */
int orc_e2d_gradient(const orc_e2d *g, const float *vp, const float *vs, const float *rho, const float *wav,
                     const int *sloc, int nsrc, const int *rloc, int nrec, const float *res, float *rec,
                     float *gvp, float *gvs, float *grho);

#ifdef __cplusplus
}
#endif
#endif
