// Shape autotuning of the run launches (node samples, leaf-depth centres, particle normals).
//
// A run launch evaluates points grouped by hash cell, one workgroup of T threads per run of kx
// cells of a grid row. The best (T, kx) depends on how many points a cell holds, on the precision
// and on the launch. The first launch of a (precision, launch kind, density bucket) times every
// candidate shape on its own data with HIP events and keeps the fastest; later launches of that
// bucket take the kept shape at once. Every shape gives the same sums (a point's candidates, and
// their order, do not depend on the run that holds it), so tuning changes only the time. Timed
// trial launches count their events into a separate counter block, so the statistics equal those
// of one launch. PISURF_RUN_TUNING=0 switches tuning off (the shape then
// comes from the occupancy heuristic of run_shape_guess).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

namespace pisurf {

struct RunShape {
    int T;   // threads per workgroup (128 or 256)
    int kx;  // cells per run
};

// kx p is about 90 % of T (a run may hold more points than the average) and the block size that
// keeps more of its threads busy wins; a cell of more than T points is taken in several passes
inline RunShape run_shape_guess(double p) {
    RunShape best{128, 1};
    double util_best = -1.0;
    for (int t : {128, 256}) {
        const int kx = std::max(1, std::min(128, (int)(0.9 * t / std::max(p, 1e-9))));
        const double pts = kx * p;
        const double util = pts / (t * std::ceil(pts / t));
        if (util > util_best + 0.02) {
            util_best = util;
            best = RunShape{t, kx};
        }
    }
    return best;
}

// the heuristic shape, then for each block size a ladder of cells per run from a quarter to twice
// its heuristic count (the mean p counts empty cells too, so occupied cells can hold several times
// p points)
inline std::vector<RunShape> run_shape_candidates(double p) {
    std::vector<RunShape> c;
    auto add = [&](RunShape s) {
        s.kx = std::max(1, std::min(128, s.kx));
        for (const auto& o : c)
            if (o.T == s.T && o.kx == s.kx) return;
        c.push_back(s);
    };
    add(run_shape_guess(p));
    for (int t : {128, 256}) {
        const double kx = 0.9 * t / std::max(p, 1e-9);
        for (double f : {0.25, 0.375, 0.5, 0.625, 0.75, 1.0, 1.5, 2.0}) add(RunShape{t, (int)std::lround(kx * f)});
    }
    return c;
}

enum RunKind { kRunSamples = 0, kRunCenters = 1, kRunNormals = 2 };

class RunTuner {
  public:
    static RunTuner& get() {
        static RunTuner t;
        return t;
    }
    bool enabled() const { return enabled_; }
    void set_enabled(bool on) { enabled_ = on; }
    // density bucket: a quarter octave of points per cell
    static int bucket(double p) { return (int)std::lround(4.0 * std::log2(std::max(p, 1.0 / 64))); }
    static long key(int prec, int kind, int bucket) { return ((long)bucket * 8 + kind) * 4 + prec; }
    bool lookup(long k, RunShape* s) {
        std::lock_guard<std::mutex> lk(m_);
        auto it = best_.find(k);
        if (it == best_.end()) return false;
        *s = it->second;
        return true;
    }
    void store(long k, RunShape s) {
        std::lock_guard<std::mutex> lk(m_);
        best_[k] = s;
    }
    size_t size() {
        std::lock_guard<std::mutex> lk(m_);
        return best_.size();
    }
    void clear() {
        std::lock_guard<std::mutex> lk(m_);
        best_.clear();
    }

  private:
    RunTuner() {
        const char* e = std::getenv("PISURF_RUN_TUNING");
        enabled_ = !(e && e[0] == '0');
    }
    std::mutex m_;
    std::map<long, RunShape> best_;
    std::atomic<bool> enabled_{true};
};

// Runs `launch(shape, trial)` for a bucket: with a kept shape once, otherwise once per candidate,
// timed (trial = true: count into the scratch block); returns the shape whose launch's results
// stand (the last one run). `prepare` builds the runs of a shape (it may read counts back) and
// returns false to skip the launch.
template <class Prepare, class Launch>
RunShape tuned_run_launch(int prec, int kind, double p, void* scratch_ctr, hipStream_t s, Prepare&& prepare,
                          Launch&& launch) {
    RunTuner& tu = RunTuner::get();
    const long k = RunTuner::key(prec, kind, RunTuner::bucket(p));
    RunShape sh = run_shape_guess(p);
    if (!tu.enabled() || tu.lookup(k, &sh) || !scratch_ctr) {
        if (prepare(sh)) launch(sh, false);
        return sh;
    }
    const std::vector<RunShape> cand = run_shape_candidates(p);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        if (prepare(sh)) launch(sh, false);
        return sh;
    }
    // the heuristic shape (tried first) stays unless a candidate is clearly faster, so that shapes
    // within timing noise of each other do not change from one process to the next
    RunShape best = sh, last = sh;
    float best_ms = INFINITY;
    bool first = true, any = false;
    for (const RunShape& c : cand) {
        if (!prepare(c)) continue;
        // the first shape also runs once untimed (its counts are the ones kept), so that loading
        // the code and warming the caches is not charged to it
        if (first) launch(c, false);
        (void)hipEventRecord(e0, s);
        launch(c, true);
        (void)hipEventRecord(e1, s);
        float ms = INFINITY;
        if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
        if (first ? ms < best_ms : ms < 0.97f * best_ms) {
            best_ms = ms;
            best = c;
        }
        first = false;
        any = true;
        last = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (any) tu.store(k, best);
    return last;
}

}  // namespace pisurf
