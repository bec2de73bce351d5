// Receiver records on the wire: the core's raw layout <-> the canonical one,
// and the L2 misfit whose residual is the adjoint source.
//
// The forward writes a record with time innermost, (nfield, B, nrec, nt) --
// (B, nrec, nt) for one field -- so each step stores one value per receiver;
// callers and losses see (B, nt, nrec, nfield) (propagator/_c.py:95-186).
// The two transposes go through a 64 (t) x 32 (r) x nfield tile in LDS: the
// raw side is read or written along t, the canonical side along (r, field),
// both coalesced.
#include "records.h"

#include <map>
#include <mutex>

#include <algorithm>

namespace swp {
namespace rec {

// Partial sums of l2_misfit_noalloc: library-owned device memory, one copy
// per GPU, held until the process ends.  Host code never passes this name to
// a kernel (on the host it is only a stand-in for the symbol); it uses the
// device address hipGetSymbolAddress returns for the current GPU.
__device__ __align__(256) double scratch[kMisfitPartials];

namespace {

constexpr int kTT = 64;   // time samples per tile
constexpr int kTR = 32;   // receivers per tile

// tile[f][t][r], the r extent padded to 33 so that a column walk in t is
// bank-conflict free.
struct Tile {
    float v[kMaxFields][kTT][kTR + 1];
};

// TO_CANON: raw -> canonical; else canonical -> raw.  Grid (t tiles, r tiles, B).
template <bool TO_CANON>
__global__ void __launch_bounds__(256) transpose_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                        int nf, int B, int nrec, int nt)
{
    __shared__ Tile s;
    const int b = int(blockIdx.z);
    const int t0 = int(blockIdx.x) * kTT, r0 = int(blockIdx.y) * kTR;
    const int tid = int(threadIdx.x);
    const int tw = min(kTT, nt - t0), rw = min(kTR, nrec - r0);
    const int row = rw * nf;   // canonical floats of one time sample in this tile
    const int64_t fstride = int64_t(B) * nrec * nt;   // raw: one field

    // raw side: lane j walks t, 4 receivers per pass
    auto raw_at = [&](int f, int r, int t) -> int64_t {
        return int64_t(f) * fstride + (int64_t(b) * nrec + r0 + r) * nt + t0 + t;
    };
    // canonical side: time sample t, element q = r * nf + f of the tile's row
    auto canon_at = [&](int t, int q) -> int64_t {
        return ((int64_t(b) * nt + t0 + t) * nrec + r0) * nf + q;
    };

    if (TO_CANON) {
        for (int f = 0; f < nf; ++f)
            for (int k = tid; k < kTT * kTR; k += 256) {
                const int t = k % kTT, r = k / kTT;
                if (t < tw && r < rw) s.v[f][t][r] = src[raw_at(f, r, t)];
            }
        __syncthreads();
        for (int k = tid; k < tw * row; k += 256) {
            const int t = k / row, q = k % row;
            dst[canon_at(t, q)] = s.v[q % nf][t][q / nf];
        }
    } else {
        for (int k = tid; k < tw * row; k += 256) {
            const int t = k / row, q = k % row;
            s.v[q % nf][t][q / nf] = src[canon_at(t, q)];
        }
        __syncthreads();
        for (int f = 0; f < nf; ++f)
            for (int k = tid; k < kTT * kTR; k += 256) {
                const int t = k % kTT, r = k / kTT;
                if (t < tw && r < rw) dst[raw_at(f, r, t)] = s.v[f][t][r];
            }
    }
}

void check_shape(int nf, int B, int nrec, int nt)
{
    SWP_CHECK(nf >= 1 && nf <= kMaxFields, "records: nfield ", nf, " outside [1, ", kMaxFields, "]");
    SWP_CHECK(B >= 0 && nrec >= 0 && nt >= 0, "records: negative extent");
    SWP_CHECK(B <= 65535, "records: B ", B, " above 65535");
    SWP_CHECK((int64_t(nrec) + kTR - 1) / kTR <= 65535, "records: nrec ", nrec, " too large");
}

template <bool TO_CANON>
void transpose(const float *src, float *dst, int nf, int B, int nrec, int nt)
{
    check_shape(nf, B, nrec, nt);
    if (int64_t(B) * nrec * nt == 0) return;
    SWP_CHECK(src && dst, "records: null buffer");
    SWP_CHECK(src != dst, "records: the transpose cannot run in place");
    const dim3 grid(unsigned((nt + kTT - 1) / kTT), unsigned((nrec + kTR - 1) / kTR), unsigned(B));
    hipLaunchKernelGGL(transpose_kernel<TO_CANON>, grid, dim3(256), 0, stream(), src, dst, nf, B, nrec, nt);
    SWP_LAUNCHED();
}

// r = (pred - obs) * mask[b, r] over canonical (B, nt, nrec, nf); per block,
// sum (double) r^2 over a grid-stride range into partial[blockIdx.x].
__global__ void __launch_bounds__(256) misfit_kernel(const float *__restrict__ pred, const float *__restrict__ obs,
                                                     const float *__restrict__ mask, float *__restrict__ res,
                                                     double *__restrict__ partial, int64_t n, int nf, int nrec,
                                                     int nt)
{
    __shared__ double red[256];
    double acc = 0.0;
    const int64_t per_shot = int64_t(nt) * nrec * nf;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        float r = pred[i] - obs[i];
        if (mask) {
            const int64_t b = i / per_shot;
            const int rr = int((i / nf) % nrec);
            r = r * mask[b * nrec + rr];
        }
        if (res) res[i] = r;
        acc += double(r) * double(r);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// loss[0] = 0.5 * sum of partial[0..np), in a fixed order.
__global__ void __launch_bounds__(256) misfit_sum_kernel(const double *__restrict__ partial, int np,
                                                         double *__restrict__ loss)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = int(threadIdx.x); i < np; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = 0.5 * red[0];
}

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch;   // device ordinal -> that GPU's copy of `scratch`

// The current GPU's address of `scratch`, looked up once per device (with
// that device current) on the first call there.
double *scratch_on_current_device()
{
    int dev = -1;
    SWP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) return it->second;
    void *p = nullptr;
    SWP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SWP_CHECK(p != nullptr, "l2_misfit_noalloc: no device address for the partial-sum array");
    g_scratch[dev] = static_cast<double *>(p);
    return static_cast<double *>(p);
}

}  // namespace

void to_canonical(const float *raw, float *canon, int nf, int B, int nrec, int nt)
{
    transpose<true>(raw, canon, nf, B, nrec, nt);
}

void to_raw(const float *canon, float *raw, int nf, int B, int nrec, int nt)
{
    transpose<false>(canon, raw, nf, B, nrec, nt);
}

int misfit_blocks(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 1023) / 1024, kMisfitPartials)));
}

void l2_misfit(const float *pred, const float *obs, const float *mask, int nf, int B, int nrec, int nt,
               float *residual, double *loss, double *partials)
{
    check_shape(nf, B, nrec, nt);
    SWP_CHECK(loss && partials, "l2_misfit: loss and partials buffers are required");
    const int64_t n = int64_t(B) * nt * nrec * nf;
    SWP_CHECK(n == 0 || (pred && obs), "l2_misfit: null record");
    const int nb = misfit_blocks(n);
    hipLaunchKernelGGL(misfit_kernel, dim3(unsigned(nb)), dim3(256), 0, stream(), pred, obs, mask, residual,
                       partials, n, nf, nrec, nt);
    SWP_LAUNCHED();
    hipLaunchKernelGGL(misfit_sum_kernel, dim3(1), dim3(256), 0, stream(), partials, nb, loss);
    SWP_LAUNCHED();
}

void l2_misfit_noalloc(const float *pred, const float *obs, const float *mask, int nf, int B, int nrec, int nt,
                       float *residual, double *loss)
{
    check_shape(nf, B, nrec, nt);
    const int64_t n = int64_t(B) * nt * nrec * nf;
    // the partial-sum count is capped by construction (grid-stride blocks);
    // the check states the bound the static array is sized for
    SWP_CHECK(misfit_blocks(n) <= kMisfitPartials, "l2_misfit_noalloc: ", misfit_blocks(n),
              " partial sums exceed the ", int(kMisfitPartials), " the library's static array holds");
    l2_misfit(pred, obs, mask, nf, B, nrec, nt, residual, loss, scratch_on_current_device());
}

}  // namespace rec
}  // namespace swp
