// shape_tuner.h - the online choice of launch shapes (host code, no HIP
// dependency).  Used by the film and batch hosts; not part of the C ABI.
#pragma once
#include <stdint.h>

#include <mutex>

namespace czi {

// Online choice between launch shapes of one kernel family whose results are
// bit-identical (so exploring costs nothing but time): per size bucket
// (log2 of the work items) each shape is timed with HIP events until it has
// WARM samples, then the one with the least time per work item (its best
// sample: the first launches of a kernel carry one-off costs, and a shape's
// best sample is what it can do) is used and one of the others re-timed once
// every RECHECK uses.  A sample's time is
// known only once its frame has run, so exploration counts the samples
// issued, not the ones folded: the first frame of a bucket times every shape
// in turn (WARM substeps each), and the frame after it uses the fastest.
// Until a bucket has samples, its substeps take the fastest shape of the
// nearest bucket that has some (the ranking moves slowly with size), not a
// fixed default.  fixed >= 0 pins a shape.
struct ShapeTuner {
    static constexpr int NB = 24, NS = 7, WARM = 3, RECHECK = 64;
    int ns = 2;  // shapes in use (at most NS)
    double best_t[NB][NS] = {};   // least time per item over the samples
    uint32_t cnt[NB][NS] = {};
    uint32_t issued[NB][NS] = {};
    uint32_t uses[NB] = {};
    int fixed = -1;
    // a tuner may be shared by films stepped from several host threads
    mutable std::mutex mu;
    static int bucket(uint64_t items) {
        int b = 0;
        while (b + 1 < NB && (1ull << (b + 1)) <= items) b++;
        return b;
    }
    // shape to use; timed = record events around this launch (only when
    // can_time: not inside a graph capture)
    int choose(uint64_t items, bool can_time, int& bk, bool& timed) {
        bk = bucket(items);
        timed = false;
        if (fixed >= 0) return fixed;
        std::lock_guard<std::mutex> g(mu);
        if (!can_time) return best_of(bk);
        // explore: the first shape still short of WARM samples issued, so
        // that a shape's samples come in consecutive substeps (the fused
        // substep runs in pairs)
        for (int k = 0; k < ns; k++) {
            if (issued[bk][k] < (uint32_t)WARM) {
                issued[bk][k]++;
                timed = true;
                return k;
            }
        }
        const int best = best_of(bk);
        if (++uses[bk] % RECHECK == 0) {
            // re-time one of the others, in turn
            timed = true;
            return (best + 1 + (int)((uses[bk] / RECHECK) % (uint32_t)(ns - 1))) % ns;
        }
        return best;
    }
    // a substep issued as shape `from` ran as shape `to` (the fused substep
    // runs in pairs; a lone one runs as separate passes).  The issue still
    // counts for `from`: were it taken back, a shape that can only be issued
    // alone would be issued again on every frame and its bucket never leave
    // the learning state
    void reissue(int bk, int from, int to) {
        (void)from;
        std::lock_guard<std::mutex> g(mu);
        issued[bk][to]++;
    }
    // true while a shape of this size class still has warm-up samples to
    // issue (a graph replay cannot time them: such frames run uncaptured)
    bool learning(uint64_t items) const {
        if (fixed >= 0) return false;
        const int b = bucket(items);
        std::lock_guard<std::mutex> g(mu);
        for (int k = 0; k < ns; k++)
            if (issued[b][k] < (uint32_t)WARM) return true;
        return false;
    }
    // one timed launch: us of device time over `items` work items
    void record(int bk, int shape, double us, double items) {
        const double t = us / (items > 1.0 ? items : 1.0);
        std::lock_guard<std::mutex> g(mu);
        if (cnt[bk][shape] == 0 || t < best_t[bk][shape]) best_t[bk][shape] = t;
        cnt[bk][shape] += 1;
    }

  private:
    // the shape with the least time per item among the bucket's sampled
    // shapes, -1 if it has none
    int best_in(int bk) const {
        int best = -1;
        double bt = 0.0;
        for (int k = 0; k < ns; k++) {
            if (cnt[bk][k] == 0) continue;
            const double t = best_t[bk][k];
            if (best < 0 || t < bt) {
                best = k;
                bt = t;
            }
        }
        return best;
    }
    // the bucket's fastest shape, else the nearest sampled bucket's (the
    // smaller size first at equal distance), else shape 0
    int best_of(int bk) const {
        for (int d = 0; d < NB; d++) {
            if (bk - d >= 0) {
                const int k = best_in(bk - d);
                if (k >= 0) return k;
            }
            if (d > 0 && bk + d < NB) {
                const int k = best_in(bk + d);
                if (k >= 0) return k;
            }
        }
        return 0;
    }
};

// The substep tuner of one device, shared by the single films on it in the
// process: a film created after others starts from their measurements
// instead of timing every shape again, and films stepped in turn share the
// exploration.  (A batch keeps a tuner of its own: its menu differs.)
ShapeTuner* device_substep_tuner(int device);

// A film's handle on its device's tuner; the pinned shape is the film's own.
struct FilmTuner {
    ShapeTuner* t = nullptr;
    int fixed = -1;
    int choose(uint64_t items, bool can_time, int& bk, bool& timed) {
        if (fixed >= 0) {
            bk = ShapeTuner::bucket(items);
            timed = false;
            return fixed;
        }
        return t->choose(items, can_time, bk, timed);
    }
    bool learning(uint64_t items) const { return fixed < 0 && t->learning(items); }
    void reissue(int bk, int from, int to) { t->reissue(bk, from, to); }
};

}  // namespace czi
