// Internal plumbing of libswp_core: error type, the thread's stream, launch
// checking, workspace carving and the kernel-timing hook.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sweep_amd.h"

namespace swp {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class... A>
[[noreturn]] inline void fail(const A &...a)
{
    std::ostringstream s;
    (s << ... << a);
    throw Error(s.str());
}

#define SWP_CHECK(cond, ...)                        \
    do {                                            \
        if (!(cond)) ::swp::fail(__VA_ARGS__);      \
    } while (0)

#define SWP_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) ::swp::fail(#call, ": ", hipGetErrorString(e_));              \
    } while (0)

#define SWP_LAUNCHED() SWP_HIP(hipGetLastError())

hipStream_t stream();

// Launch-shape knob `key` (SWP_TUNE_*): the value set through
// swp_set_tuning, else the environment variable `env` if set, else `def`.
int tuning(int key, const char *env, int def);

// Bump allocator over the caller's workspace (or a size-only dry run when
// base == nullptr), 256-byte aligned slices.
struct Carver {
    char *base;
    int64_t cap;
    int64_t used = 0;
    Carver(void *b, int64_t c) : base(static_cast<char *>(b)), cap(c) {}
    template <class T>
    T *take(int64_t count)
    {
        int64_t off = (used + 255) & ~int64_t(255);
        used = off + count * int64_t(sizeof(T));
        if (!base) return nullptr;
        SWP_CHECK(used <= cap, "workspace/tape too small: need ", used, " bytes, have ", cap);
        return reinterpret_cast<T *>(base + off);
    }
};

// Kernel-class timing (bench.py's roofline leg): with timing on, the drivers
// bracket the launches of each class with events from a reused pool.
enum KClass { K_FWD = 0, K_ADJ = 1, K_RECON = 2, K_OTHER = 3, K_NCLASS = 4 };

// Timing samples one launch in every `stride` of a class (an unbiased mean
// of the launch duration at bounded event count); samples accumulate across
// calls until swp_set_timing turns timing on again.
struct KTimer {
    bool on = false;
    int stride = 1;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, int>> marks[K_NCLASS];   // pool index pairs
    int64_t seen[K_NCLASS] = {0, 0, 0, 0};
    size_t next = 0;
    void reset();
    int begin(int cls);
    void end(int cls, int token);
    void stats(int cls, double *ms, int64_t *n);
};
KTimer &timer();

struct TimedScope {
    int cls, tok;
    explicit TimedScope(int c) : cls(c), tok(timer().begin(c)) {}
    ~TimedScope() { timer().end(cls, tok); }
};

// Zero a float buffer with a fill kernel (graph-capture safe: no memset node).
void fill_zero(float *p, int64_t n);

// Launch census: while it is on (swp_census_enable), every z-march and
// region launch adds one launch and its workgroup count under its site name,
// so a test can assert which kernel variants a case actually ran.
bool census_on();
void census_add(const char *site, int64_t blocks);

}  // namespace swp
