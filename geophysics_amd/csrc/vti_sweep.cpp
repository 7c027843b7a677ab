// vti_sweep.cpp — the reverse-sweep driver shared by RTM, the batched RTM
// and the adjoint / gradient path: binomial (revolve) checkpointing of the
// forward states, host logic only.  vti_sweep_check runs the schedule on
// integer positions instead of wavefields so its order and cost are
// testable without a device.
#include <algorithm>
#include <functional>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

namespace vti {

// ---------------------------------------------------------------------
// Binomial checkpointing (Griewank & Walther, "revolve", ACM TOMS 26, 2000).
// beta(c, r) = C(c+r, c): with c checkpoints and r forward repetitions at
// most beta(c, r) states can be reversed.  Here X_0 is free (the zero
// state) and each checkpointed state is consumed from its slot before the
// slot is freed, which makes the problem "nt+1 states, s+1 checkpoints" of
// the classical analysis: the minimum is r*(nt+1) - beta(s+2, r-1) forward
// steps, r the least integer with beta(s+1, r) >= nt+1.  The split below
// attains it (checked against an exhaustive search of the recursion
// cost(n, s) = min_j j + cost(n-j, s-1) + cost(j-1, s) for n < 220, s <= 10).
static double beta(int c, int r)
{
    if (r < 0) return 0.0;
    double b = 1.0;
    for (int i = 1; i <= c; ++i) b = b * (r + i) / i;   // exact while below 2^53
    return b;
}

// Steps to advance from X_l before checkpointing, for n = r-l states to
// consume with s >= 1 free slots: the least length of the optimal range.
static int split_len(int n, int s)
{
    int r = 0;
    while (beta(s + 1, r) < n + 1.0) ++r;
    const double j = std::max(beta(s + 1, r - 2), n - beta(s, r) + 1.0);
    return (int)std::min<double>(n, std::max(1.0, j));
}

int reverse_sweep(int nt, int nslot, const SweepOps &op, long long *fwd_steps)
{
    // States X_0..X_nt, X_k = state after k steps.  consume(k) receives X_k
    // in the working state for k = nt .. 1.  Steps past the furthest point
    // reached so far belong to the first sweep (first_step); the schedule
    // reaches X_nt by one uninterrupted sweep from X_0.
    if (nt < 0 || nslot < 0) return fail(VTI_ERR_ARG, "reverse_sweep: bad sizes");
    nslot = std::min(nslot, std::max(nt - 1, 0));   // more slots never help
    long long steps = 0;
    int work = -1, frontier = 0;
    bool first_done = false;
    std::vector<int> free_slots, pos_slot(nt + 1, -1);
    for (int i = nslot - 1; i >= 0; --i) free_slots.push_back(i);
    auto restore = [&](int pos) -> int {
        if (pos == work) return VTI_OK;
        int e;
        if (pos == 0) e = op.zero();
        else if (pos_slot[pos] < 0) return fail(VTI_ERR_STATE, "reverse_sweep: missing checkpoint");
        else e = op.restore(pos_slot[pos]);
        if (!e) work = pos;
        return e;
    };
    auto advance = [&](int to) -> int {
        while (work < to) {
            const int it = work;
            const int e = it >= frontier ? op.first_step(it) : op.step(it);
            if (e) return e;
            ++steps;
            if (++work > frontier) frontier = work;
        }
        return VTI_OK;
    };
    auto consume = [&](int k) -> int {
        if (!first_done) {
            first_done = true;
            if (k != nt || work != nt) return fail(VTI_ERR_STATE, "reverse_sweep: X_nt must come first");
            if (op.after_first) {
                const int e = op.after_first();
                if (e) return e;
            }
        }
        return op.consume(k);
    };
    // explicit task stack: kind 0 = consume X_r .. X_{l+1} (X_l restorable)
    // with s free slots; kind 1 = consume the checkpoint X_r, free its slot,
    // then continue with X_{r-1} .. X_{l+1}.
    struct Task {
        int kind, l, r, s;
    };
    std::vector<Task> todo{{0, 0, nt, nslot}};
    while (!todo.empty()) {
        const Task t = todo.back();
        todo.pop_back();
        int e;
        if (t.kind == 1) {
            if (op.consume_slot && first_done) {
                if ((e = op.consume_slot(t.r, pos_slot[t.r]))) return e;
            } else if ((e = restore(t.r)) || (e = consume(t.r))) {
                return e;
            }
            free_slots.push_back(pos_slot[t.r]);
            pos_slot[t.r] = -1;
            todo.push_back({0, t.l, t.r - 1, t.s});
            continue;
        }
        if (t.r <= t.l) continue;
        if (t.s <= 0) {
            for (int k = t.r; k > t.l; --k)
                if ((e = restore(t.l)) || (e = advance(k)) || (e = consume(k))) return e;
            continue;
        }
        const int m = t.l + split_len(t.r - t.l, t.s);
        if ((e = restore(t.l)) || (e = advance(m))) return e;
        if (m == t.r) {
            if ((e = consume(m))) return e;
            todo.push_back({0, t.l, m - 1, t.s});
            continue;
        }
        if (free_slots.empty()) return fail(VTI_ERR_STATE, "reverse_sweep: slot overflow");
        const int i = free_slots.back();
        free_slots.pop_back();
        if ((e = op.store(i))) return e;
        pos_slot[m] = i;
        todo.push_back({1, t.l, m, t.s});
        todo.push_back({0, m, t.r, t.s - 1});
    }
    if (fwd_steps) *fwd_steps = steps;
    return VTI_OK;
}

}  // namespace vti

extern "C" {

// Runs the schedule on integer positions: the working "state" is the number
// of steps taken from X_0, a slot holds a position.  Fails unless consume(k)
// sees X_k for k = nt .. 1 in that order with at most nslot slots in use,
// both with checkpointed states restored before they are consumed and
// consumed from their slots (the same forward steps either way).
int vti_sweep_check(int nt, int nslot, long long *fwd_steps, int *max_slots)
{
    if (nt < 1 || nslot < 0) return vti::fail(VTI_ERR_ARG, "vti_sweep_check: nt >= 1 and nslot >= 0");
    long long steps[2] = {0, 0};
    int peak = 0;
    for (int from_slot = 0; from_slot < 2; ++from_slot) {
        int work = -1, next = nt, used = 0;
        peak = 0;
        std::vector<int> slot(nslot, -1);
        bool first_done = false, bad = false;
        vti::SweepOps op;
        op.zero = [&]() { work = 0; return VTI_OK; };
        op.store = [&](int i) {
            if (i < 0 || i >= nslot) { bad = true; return VTI_ERR_STATE; }
            if (slot[i] < 0) peak = std::max(peak, ++used);
            slot[i] = work;
            return VTI_OK;
        };
        op.restore = [&](int i) {
            if (i < 0 || i >= nslot || slot[i] < 0) { bad = true; return VTI_ERR_STATE; }
            work = slot[i];
            return VTI_OK;
        };
        op.step = [&](int it) {
            if (it != work || !first_done) { bad = true; return VTI_ERR_STATE; }
            ++work;
            return VTI_OK;
        };
        op.first_step = [&](int it) {
            if (it != work || first_done) { bad = true; return VTI_ERR_STATE; }
            ++work;
            return VTI_OK;
        };
        op.after_first = [&]() {
            if (work != nt) { bad = true; return VTI_ERR_STATE; }
            first_done = true;
            return VTI_OK;
        };
        op.consume = [&](int k) {
            if (k != next || work != k) { bad = true; return VTI_ERR_STATE; }
            --next;
            return VTI_OK;
        };
        if (from_slot)
            op.consume_slot = [&](int k, int i) {
                // the slot holds X_k; the working state is left as it is
                if (k != next || i < 0 || i >= nslot || slot[i] != k) { bad = true; return VTI_ERR_STATE; }
                --next;
                return VTI_OK;
            };
        const int st = vti::reverse_sweep(nt, nslot, op, &steps[from_slot]);
        if (st || bad || next != 0) return vti::fail(VTI_ERR_STATE, "vti_sweep_check: schedule out of order");
    }
    if (steps[0] != steps[1]) return vti::fail(VTI_ERR_STATE, "vti_sweep_check: slot consumes change the steps");
    if (fwd_steps) *fwd_steps = steps[0];
    if (max_slots) *max_slots = peak;
    return VTI_OK;
}

}  // extern "C"
