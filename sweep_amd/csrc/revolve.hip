// Binomial checkpoint schedules (host code; see revolve.h).
#include "revolve.h"

#include "core.h"

#include <algorithm>

namespace swp {

namespace {

// C(n, k), saturating far above any step count.
int64_t binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) {
        r = r * double(n - k + i) / double(i);
        if (r > 1e15) return int64_t(1e15);
    }
    return int64_t(r + 0.5);
}

// Steps before the first snapshot when l >= 2 steps are reversed with s >= 2
// slots: the smallest repetition number r with C(s + r, s) >= l, then the
// left part as long as r - 1 repetitions allow.
int split(int l, int s)
{
    int r = 1;
    while (binom(s + r, s) < l) ++r;
    int64_t m = std::min<int64_t>(binom(s + r - 1, s), l - 1);
    m = std::max<int64_t>(m, l - binom(s - 1 + r, s - 1));
    return int(std::max<int64_t>(1, std::min<int64_t>(m, l - 1)));
}

struct Builder {
    std::vector<RevOp> ops;
    std::vector<int> have;   // step whose state each slot holds, -1: none
    int cur = -1;            // step of the working state, -1: unknown

    void emit(int32_t code, int32_t arg) { ops.push_back({code, arg}); }
    void at(int t0, int slot)
    {
        if (cur == t0) return;
        SWP_CHECK(have[slot] == t0, "revolve: slot ", slot, " holds step ", have[slot], ", not ", t0);
        emit(REV_LOAD, slot);
        cur = t0;
    }
    void advance(int to)
    {
        for (; cur < to; ++cur) emit(REV_STEP, cur);
    }
    // Reverse steps [t0, t0 + l) with slots slot .. slot + s - 1, slot holding t0.
    void rev(int t0, int l, int s, int slot)
    {
        if (l <= 0) return;
        if (l == 1 || s == 1) {
            for (int k = l - 1; k >= 0; --k) {
                at(t0, slot);
                advance(t0 + k);
                emit(REV_REVERSE, t0 + k);
                cur = t0 + k + 1;
            }
            return;
        }
        const int m = split(l, s);
        const int ns = slot + 1;
        if (have[ns] != t0 + m) {
            at(t0, slot);
            advance(t0 + m);
            emit(REV_SAVE, ns);
            have[ns] = t0 + m;
        }
        rev(t0 + m, l - m, s - 1, ns);
        rev(t0, m, s, slot);
    }
};

}  // namespace

std::vector<SpinePt> revolve_spine(int nt, int slots)
{
    SWP_CHECK(slots >= 1, "revolve: need at least one snapshot slot");
    std::vector<SpinePt> sp;
    if (nt <= 0) return sp;
    sp.push_back({0, 0});
    int t = 0, l = nt, s = slots, slot = 0;
    while (l >= 2 && s >= 2) {
        const int m = split(l, s);
        t += m;
        l -= m;
        --s;
        ++slot;
        sp.push_back({slot, t});
    }
    return sp;
}

std::vector<RevOp> revolve_schedule(int nt, int slots)
{
    SWP_CHECK(slots >= 1, "revolve: need at least one snapshot slot");
    Builder b;
    b.have.assign(size_t(slots) + 1, -1);
    for (const SpinePt &p : revolve_spine(nt, slots)) b.have[size_t(p.slot)] = p.step;
    b.rev(0, nt, slots, 0);
    return b.ops;
}

}  // namespace swp
