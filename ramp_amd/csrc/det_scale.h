/* det_scale.h - the fixed-point scale and int64 budget of deterministic
 * histograms (host arithmetic only; no HIP).
 *
 * A deterministic histogram adds q = rint(w * 2^sh) per event into int64
 * sums, so the sums are exact and independent of the order of the events.
 * The scale is chosen from the first run: its histories times a bound on
 * one history's weight, with 16x headroom, must stay below 2^61.  Histograms
 * accumulate across runs (as the reference's do), so the plan also keeps a
 * budget: the sum over runs of (weight bound x histories run x 2^sh), an
 * upper bound of any bin's sum.  A run that would take the budget past 2^62
 * is preceded by a fold: the device sums are added into FP64 host
 * accumulators and cleared, the budget restarts at zero and the scale is
 * chosen again from that run.
 */
#ifndef RAMP_DET_SCALE_H
#define RAMP_DET_SCALE_H

#include <algorithm>
#include <cmath>
#include <cstdint>

struct DetScale {
    bool set = false;
    int sh = 0, sh2 = 0;          /* q = rint(w * 2^sh), q2 = rint(w^2 * 2^sh2) */
    double used = 0, used2 = 0;   /* budgets of the w and w^2 sums (units of q) */
};

static const double DET_BUDGET = 4611686018427387904.0;  /* 2^62 */

/* the scale for runs of n_scale histories of weight at most wb */
static inline void det_scale_choose(DetScale &s, double wb, double n_scale)
{
    const double tot = wb * n_scale * 16.0;
    const double tot2 = wb * wb * n_scale * 16.0;
    int sh = (int)std::floor(std::log2(std::ldexp(1.0, 61) / tot));
    int sh2 = (int)std::floor(std::log2(std::ldexp(1.0, 61) / tot2));
    s.sh = std::max(-120, std::min(120, sh));
    s.sh2 = std::max(-120, std::min(120, sh2));
    s.set = true;
}

/* Account a run of n_run histories (weight at most wb) whose scale is
 * chosen for n_scale histories (a shard's whole chunk).  Returns 1 when the
 * device sums must be folded into the host accumulators and cleared before
 * the run; the scale and the budget have then been chosen anew for it. */
static inline int det_scale_run(DetScale &s, double wb, double n_scale, double n_run)
{
    if (n_run > n_scale) n_scale = n_run;
    if (!s.set) det_scale_choose(s, wb, n_scale);
    double add = wb * n_run * std::ldexp(1.0, s.sh);
    double add2 = wb * wb * n_run * std::ldexp(1.0, s.sh2);
    int fold = 0;
    if (s.used + add > DET_BUDGET || s.used2 + add2 > DET_BUDGET) {
        fold = s.used > 0 || s.used2 > 0;
        s.used = s.used2 = 0;
        det_scale_choose(s, wb, n_scale);
        add = wb * n_run * std::ldexp(1.0, s.sh);
        add2 = wb * wb * n_run * std::ldexp(1.0, s.sh2);
    }
    s.used += add;
    s.used2 += add2;
    return fold;
}

#endif
