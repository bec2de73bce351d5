// ORACLE -- test infrastructure only.
//
// A C entry point around the reference's own compiled path (its host core:
// the reference's CUDA sources built by g++ with its host shims, which is how
// the reference itself runs impl='c' on a CPU; backend/c/hostcore.py).  It
// fills the reference's ForwardInputCore / BackwardInputCore
// (csrc/core/input_core.h) from plain pointers and calls the equation's
// forward_core / backward_core (full-storage memory mode) directly -- the
// reference's capi.cu dispatcher is not compiled.  Built by oracle/build_ref.sh
// into oracle/_ref/ (never committed, never shipped); used by tests to pin
// the C oracle to the reference's compiled arithmetic.
#include <cstring>
#include <exception>
#include <vector>

#include "core/input_core.h"
#include "core/outputs.h"
#include "cuda/equations/acoustic2d/acoustic2d.h"
#include "cuda/equations/acoustic3d/acoustic3d.h"
#include "cuda/equations/elastic2d/elastic2d.h"

namespace {

Buf mk(const void *p, std::vector<int64_t> sizes, int elem = 4)
{
    Buf b;
    b.data_ = const_cast<void *>(p);
    b.ndim_ = int32_t(sizes.size());
    int64_t st = 1;
    for (int d = int(sizes.size()) - 1; d >= 0; --d) {
        b.sizes_[d] = sizes[d];
        b.strides_[d] = st;
        st *= sizes[d];
    }
    b.numel_ = st;
    b.elem_size_ = elem;
    b.defined_ = true;
    b.is_cuda_ = false;
    b.device_ = -1;
    return b;
}

BufList list(const std::vector<Buf> &v) { return BufList{v.data(), int64_t(v.size())}; }

void put_err(const char *m, char *err, int cap)
{
    if (err && cap > 0) {
        std::strncpy(err, m, size_t(cap) - 1);
        err[cap - 1] = 0;
    }
}

// grid: (nz, nx) in 2-D, (nz, ny, nx) in 3-D
template <int NDIM>
int run(const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zyx,
        const float *const *pml, const float *vp, const float *lap, const float *grad, const float *src,
        const int *sloc, int nsrc, const int *rloc, int nrec, const float *res, float *rec_out,
        float *grad_vp, float *grad_w, char *err, int cap, int bs = 0, float *illum_s = nullptr,
        float *illum_r = nullptr)
{
    try {
        const int nz = grid[0], ny = NDIM == 3 ? grid[1] : 1, nx = grid[NDIM - 1];
        const int64_t n = int64_t(nz) * ny * nx;
        std::vector<int64_t> gshape = NDIM == 3 ? std::vector<int64_t>{B, 1, nz, ny, nx}
                                                : std::vector<int64_t>{B, 1, nz, nx};
        const int nfwd = NDIM == 3 ? 12 : 9, nadj = NDIM == 3 ? 15 : 11;
        std::vector<std::vector<float>> store(nfwd + nadj, std::vector<float>(size_t(B * n), 0.f));
        std::vector<Buf> fw, aw;
        for (int k = 0; k < nfwd; ++k) fw.push_back(mk(store[k].data(), gshape));
        for (int k = 0; k < nadj; ++k) aw.push_back(mk(store[nfwd + k].data(), gshape));
        std::vector<Buf> models{mk(vp, gshape)};
        std::vector<Buf> pml_b;
        const int axes[3] = {nz, ny, nx};
        if (NDIM == 2) {
            for (int k = 0; k < 3; ++k) pml_b.push_back(mk(pml[k], {nz}));
            for (int k = 3; k < 6; ++k) pml_b.push_back(mk(pml[k], {nx}));
        } else {
            for (int k = 0; k < 9; ++k) pml_b.push_back(mk(pml[k], {axes[k / 3]}));
        }
        // spacing in Cartesian order (x[, y], z), propagator/_c.py:1129-1132
        std::vector<float> spacing;
        if (NDIM == 2) spacing = {spacing_zyx[1], spacing_zyx[0]};
        else spacing = {spacing_zyx[2], spacing_zyx[1], spacing_zyx[0]};
        const int fidx = 0;
        std::vector<float> u_allt(size_t(nt) * B * n, 0.f);
        std::vector<int64_t> ashape{nt};
        for (auto s : gshape) if (&s != &gshape[1]) ashape.push_back(s);
        ashape = NDIM == 3 ? std::vector<int64_t>{nt, B, nz, ny, nx} : std::vector<int64_t>{nt, B, nz, nx};

        ForwardInputCore f{};
        f.models = list(models);
        f.source = mk(src, {B, nsrc, nt});
        f.lap_coes = mk(lap, {M + 1});
        f.grad_coes = mk(grad, {M + 1});
        f.M = M;
        f.abcn = abcn;
        f.sources_loc = mk(sloc, {B, nsrc, NDIM});
        f.receivers_loc = mk(rloc, {B, nrec, NDIM});
        f.source_field_indices = IntSpan{&fidx, 1};
        f.receiver_field_indices = IntSpan{&fidx, 1};
        f.pml_vals = list(pml_b);
        f.wavefields = list(fw);
        // boundary saving: the saver stores u only, strips M + 1 wide at
        // offset 0.  2-D: top / bottom (1, nt, B, W, nx_phys), left / right
        // (1, nt, B, nz_phys, W).  3-D (time first, the per-step stride):
        // top / bottom (nt, B, W, ny_phys, nx_phys), front / back
        // (nt, B, nz_phys, W, nx_phys), left / right (nt, B, nz_phys, ny_phys, W).
        // Last state (1, 2, B, 1, grid).
        const int W = M + 1, nxp = nx - 2 * (abcn + M), nzp = nz - 2 * (abcn + M);
        const int nyp = NDIM == 3 ? ny - 2 * (abcn + M) : 1;
        const bool use_bs = bs && res;
        const int64_t ftb = int64_t(nt) * B * W * nyp * nxp, ffb = int64_t(nt) * B * nzp * W * nxp,
                      flr = int64_t(nt) * B * nzp * nyp * W;
        std::vector<float> s_tb(use_bs ? size_t(2 * ftb) : 0), s_fb(use_bs && NDIM == 3 ? size_t(2 * ffb) : 0),
            s_lr(use_bs ? size_t(2 * flr) : 0);
        std::vector<float> last(use_bs ? size_t(2) * B * n : 0, 0.f);
        std::vector<Buf> bgpu;
        if (use_bs && NDIM == 2) {
            bgpu = {mk(s_tb.data(), {1, nt, B, W, nxp}), mk(s_tb.data() + ftb, {1, nt, B, W, nxp}),
                    mk(s_lr.data(), {1, nt, B, nzp, W}), mk(s_lr.data() + flr, {1, nt, B, nzp, W})};
        } else if (use_bs) {
            bgpu = {mk(s_tb.data(), {nt, B, W, nyp, nxp}), mk(s_tb.data() + ftb, {nt, B, W, nyp, nxp}),
                    mk(s_fb.data(), {nt, B, nzp, W, nxp}), mk(s_fb.data() + ffb, {nt, B, nzp, W, nxp}),
                    mk(s_lr.data(), {nt, B, nzp, nyp, W}), mk(s_lr.data() + flr, {nt, B, nzp, nyp, W})};
        }
        std::vector<int64_t> lshape{1, 2, B, 1, nz};
        if (NDIM == 3) lshape.push_back(ny);
        lshape.push_back(nx);
        const Buf last_b = mk(last.data(), lshape);
        f.save_all_wavefields = res != nullptr && !use_bs;
        f.use_boundary_saving = use_bs;
        if (use_bs) {
            f.boundary_gpu = list(bgpu);
            f.last_two = last_b;
        }
        f.nt = unsigned(nt);
        f.dt = dt;
        f.spacing = FloatSpan{spacing.data(), int64_t(spacing.size())};
        f.record_out = mk(rec_out, {B, nrec, nt});
        if (res && !use_bs) f.u_allt_out = mk(u_allt.data(), ashape);
        if (NDIM == 2) acoustic2d::forward_core(f);
        else acoustic3d::forward_core(f);
        if (!res) return 0;

        std::vector<Buf> grads{mk(grad_w, {B, nsrc, nt}), mk(grad_vp, gshape)};
        BackwardInputCore bi{};
        std::vector<std::vector<float>> recon(use_bs ? 3 : 0, std::vector<float>(size_t(B * n), 0.f));
        std::vector<Buf> rw;
        for (auto &r : recon) rw.push_back(mk(r.data(), gshape));
        if (use_bs) {
            bi.u_last_two = last_b;
            bi.boundary_gpu = list(bgpu);
            bi.forward_wavefields = list(rw);
        } else {
            bi.u_forward = mk(u_allt.data(), ashape);
        }
        bi.adjoint_wavefields = list(aw);
        bi.models = list(models);
        bi.adjoint_source = mk(res, {B, nrec, nt});
        bi.forward_source = mk(src, {B, nsrc, nt});
        bi.lap_coes = f.lap_coes;
        bi.grad_coes = f.grad_coes;
        bi.M = M;
        bi.abcn = abcn;
        bi.adjoint_sources_loc = f.receivers_loc;
        bi.forward_sources_loc = f.sources_loc;
        bi.source_field_indices = f.source_field_indices;
        bi.receiver_field_indices = f.receiver_field_indices;
        bi.pml_vals = f.pml_vals;
        bi.nt = unsigned(nt);
        bi.dt = dt;
        bi.spacing = f.spacing;
        std::vector<Buf> illum;
        if (illum_s && illum_r) illum = {mk(illum_s, gshape), mk(illum_r, gshape)};
        bi.compute_illumination = !illum.empty();
        if (!illum.empty()) bi.illum_out = list(illum);
        bi.grads_out = list(grads);
        if (use_bs && NDIM == 2) acoustic2d::backward_bs_core(bi);
        else if (use_bs) acoustic3d::backward_bs_core(bi);
        else if (NDIM == 2) acoustic2d::backward_core(bi);
        else acoustic3d::backward_core(bi);
        return 0;
    } catch (const std::exception &e) {
        put_err(e.what(), err, cap);
    } catch (...) {
        put_err("unknown error", err, cap);
    }
    return 1;
}

// Elastic 2-D (cpmls, velocity-stress): vp, vs, rho (B, nz, nx); one vz
// source per shot (field 1), receivers recording vx and vz (fields 0, 1);
// pml = [az, bz, azh, bzh, ax, bx, axh, bxh]; res (2, B, nrec, nt).
int run_elastic(const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zx,
                const float *const *pml, const float *vp, const float *vs, const float *rho, const float *grad,
                const float *src, const int *sloc, int nsrc, const int *rloc, int nrec, const float *res,
                float *rec_out, float *g_vp, float *g_vs, float *g_rho, char *err, int cap, int bs = 0)
{
    try {
        const int nz = grid[0], nx = grid[1];
        const int64_t n = int64_t(nz) * nx;
        const std::vector<int64_t> gshape{B, 1, nz, nx};
        const int nstate = 15;
        std::vector<std::vector<float>> store(2 * nstate + 9 + 2, std::vector<float>(size_t(B * n), 0.f));
        std::vector<Buf> fw, aw, ws, dm;
        for (int k = 0; k < nstate; ++k) fw.push_back(mk(store[k].data(), gshape));
        for (int k = 0; k < nstate; ++k) aw.push_back(mk(store[nstate + k].data(), gshape));
        for (int k = 0; k < 9; ++k) ws.push_back(mk(store[2 * nstate + k].data(), gshape));
        for (int k = 0; k < 2; ++k) dm.push_back(mk(store[2 * nstate + 9 + k].data(), gshape));
        std::vector<Buf> models{mk(vp, gshape), mk(vs, gshape), mk(rho, gshape)};
        std::vector<Buf> pml_b;
        for (int k = 0; k < 4; ++k) pml_b.push_back(mk(pml[k], {nz}));
        for (int k = 4; k < 8; ++k) pml_b.push_back(mk(pml[k], {nx}));
        std::vector<float> spacing{spacing_zx[1], spacing_zx[0]};
        const int sfield = 1;
        const int rfields[2] = {0, 1};
        std::vector<float> u_allt(bs ? 0 : size_t(nt) * 2 * B * n, 0.f);
        const std::vector<int64_t> ashape{nt, 2, B, nz, nx};
        std::vector<float> lap_dummy(size_t(M + 1), 0.f);
        // boundary-saving storage: top / bottom (5, nt, B, W, nx_phys), left /
        // right (5, nt, B, nz_phys, W), last state (5, 2, B, 1, nz, nx)
        const int W = M + 1, nxp = nx - 2 * (abcn + M), nzp = nz - 2 * (abcn + M);
        const int64_t tb = int64_t(5) * nt * B * W * nxp, lr = int64_t(5) * nt * B * nzp * W;
        std::vector<float> s_top(bs ? tb : 0), s_bot(bs ? tb : 0), s_left(bs ? lr : 0), s_right(bs ? lr : 0);
        std::vector<float> last(bs ? size_t(5) * 2 * B * n : 0, 0.f);
        std::vector<Buf> bgpu;
        if (bs) {
            bgpu = {mk(s_top.data(), {5, nt, B, W, nxp}), mk(s_bot.data(), {5, nt, B, W, nxp}),
                    mk(s_left.data(), {5, nt, B, nzp, W}), mk(s_right.data(), {5, nt, B, nzp, W})};
        }
        const Buf last_b = mk(last.data(), {5, 2, B, 1, nz, nx});

        ForwardInputCore f{};
        f.models = list(models);
        f.derived_models = list(dm);
        f.source = mk(src, {B, nsrc, nt});
        f.lap_coes = mk(lap_dummy.data(), {M + 1});
        f.grad_coes = mk(grad, {M + 1});
        f.M = M;
        f.abcn = abcn;
        f.sources_loc = mk(sloc, {B, nsrc, 2});
        f.receivers_loc = mk(rloc, {B, nrec, 2});
        f.source_field_indices = IntSpan{&sfield, 1};
        f.receiver_field_indices = IntSpan{rfields, 2};
        f.pml_vals = list(pml_b);
        f.wavefields = list(fw);
        f.save_all_wavefields = res != nullptr && !bs;
        if (bs && res) {
            f.use_boundary_saving = true;
            f.boundary_gpu = list(bgpu);
            f.last_two = last_b;
        }
        f.nt = unsigned(nt);
        f.dt = dt;
        f.spacing = FloatSpan{spacing.data(), 2};
        f.record_out = mk(rec_out, {2, B, nrec, nt});
        if (res && !bs) f.u_allt_out = mk(u_allt.data(), ashape);
        elastic2d::forward_core(f);
        if (!res) return 0;

        std::vector<Buf> grads{mk(g_vp, gshape), mk(g_vs, gshape), mk(g_rho, gshape)};
        BackwardInputCore bi{};
        std::vector<std::vector<float>> recon(bs ? 7 : 0, std::vector<float>(size_t(B * n), 0.f));
        std::vector<Buf> rw;
        for (auto &r : recon) rw.push_back(mk(r.data(), gshape));
        if (bs) {
            bi.u_last_two = last_b;
            bi.boundary_gpu = list(bgpu);
            bi.forward_wavefields = list(rw);
        } else {
            bi.u_forward = mk(u_allt.data(), ashape);
        }
        bi.adjoint_wavefields = list(aw);
        bi.adjoint_workspace = BufList{ws.data(), int64_t(bs ? 8 : ws.size())};
        bi.derived_models = list(dm);
        bi.models = list(models);
        bi.adjoint_source = mk(res, {2, B, nrec, nt});
        bi.forward_source = mk(src, {B, nsrc, nt});
        bi.lap_coes = f.lap_coes;
        bi.grad_coes = f.grad_coes;
        bi.M = M;
        bi.abcn = abcn;
        bi.adjoint_sources_loc = f.receivers_loc;
        bi.forward_sources_loc = f.sources_loc;
        bi.source_field_indices = f.source_field_indices;
        bi.receiver_field_indices = f.receiver_field_indices;
        bi.pml_vals = f.pml_vals;
        bi.nt = unsigned(nt);
        bi.dt = dt;
        bi.spacing = f.spacing;
        bi.compute_illumination = false;
        bi.grads_out = list(grads);
        if (bs) elastic2d::backward_bs_core(bi);
        else elastic2d::backward_core(bi);
        return 0;
    } catch (const std::exception &e) {
        put_err(e.what(), err, cap);
    } catch (...) {
        put_err("unknown error", err, cap);
    }
    return 1;
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) int ref_elastic_run(
    const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zx, const float *const *pml,
    const float *vp, const float *vs, const float *rho, const float *grad, const float *src, const int *sloc,
    int nsrc, const int *rloc, int nrec, const float *res, float *rec_out, float *g_vp, float *g_vs, float *g_rho,
    char *err, int cap)
{
    return run_elastic(grid, B, M, nt, abcn, dt, spacing_zx, pml, vp, vs, rho, grad, src, sloc, nsrc, rloc, nrec,
                       res, rec_out, g_vp, g_vs, g_rho, err, cap);
}

// The same with boundary saving (backward_bs_core).
__attribute__((visibility("default"))) int ref_elastic_run_bs(
    const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zx, const float *const *pml,
    const float *vp, const float *vs, const float *rho, const float *grad, const float *src, const int *sloc,
    int nsrc, const int *rloc, int nrec, const float *res, float *rec_out, float *g_vp, float *g_vs, float *g_rho,
    char *err, int cap)
{
    return run_elastic(grid, B, M, nt, abcn, dt, spacing_zx, pml, vp, vs, rho, grad, src, sloc, nsrc, rloc, nrec,
                       res, rec_out, g_vp, g_vs, g_rho, err, cap, 1);
}

// 2-D / 3-D acoustic with the source and receiver illumination (sum u_tt^2,
// sum lambda^2, (B, grid) each), full storage (bs = 0) or boundary saving.
__attribute__((visibility("default"))) int ref_acoustic_run_illum(
    int ndim, int bs, const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zyx,
    const float *const *pml, const float *vp, const float *lap, const float *grad, const float *src,
    const int *sloc, int nsrc, const int *rloc, int nrec, const float *res, float *rec_out, float *grad_vp,
    float *grad_w, float *illum_s, float *illum_r, char *err, int cap)
{
    if (ndim == 2)
        return run<2>(grid, B, M, nt, abcn, dt, spacing_zyx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                      rec_out, grad_vp, grad_w, err, cap, bs, illum_s, illum_r);
    return run<3>(grid, B, M, nt, abcn, dt, spacing_zyx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                  rec_out, grad_vp, grad_w, err, cap, bs, illum_s, illum_r);
}

// 2-D acoustic with boundary saving (backward_bs_core).
__attribute__((visibility("default"))) int ref_acoustic2d_run_bs(
    const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zx, const float *const *pml,
    const float *vp, const float *lap, const float *grad, const float *src, const int *sloc, int nsrc,
    const int *rloc, int nrec, const float *res, float *rec_out, float *grad_vp, float *grad_w, char *err, int cap)
{
    return run<2>(grid, B, M, nt, abcn, dt, spacing_zx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                  rec_out, grad_vp, grad_w, err, cap, 1);
}

// 3-D acoustic with boundary saving (backward_bs_core).
__attribute__((visibility("default"))) int ref_acoustic3d_run_bs(
    const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zyx, const float *const *pml,
    const float *vp, const float *lap, const float *grad, const float *src, const int *sloc, int nsrc,
    const int *rloc, int nrec, const float *res, float *rec_out, float *grad_vp, float *grad_w, char *err, int cap)
{
    return run<3>(grid, B, M, nt, abcn, dt, spacing_zyx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                  rec_out, grad_vp, grad_w, err, cap, 1);
}

__attribute__((visibility("default"))) int ref_acoustic_run(
    int ndim, const int *grid, int B, int M, int nt, int abcn, float dt, const float *spacing_zyx,
    const float *const *pml, const float *vp, const float *lap, const float *grad, const float *src,
    const int *sloc, int nsrc, const int *rloc, int nrec, const float *res, float *rec_out, float *grad_vp,
    float *grad_w, char *err, int cap)
{
    if (ndim == 2)
        return run<2>(grid, B, M, nt, abcn, dt, spacing_zyx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                      rec_out, grad_vp, grad_w, err, cap);
    return run<3>(grid, B, M, nt, abcn, dt, spacing_zyx, pml, vp, lap, grad, src, sloc, nsrc, rloc, nrec, res,
                  rec_out, grad_vp, grad_w, err, cap);
}

}  // extern "C"
