// Batch of independent c2c 3D transforms of different sizes, run as phases over device work
// lists (one launch per register class of tile), optionally replayed from a captured hipGraph:
// the fused x-y phase for the transforms whose planes and tiles allow it, the x and y phases for
// the others, then the z phase for all.  There is no counterpart in heFFTe, whose
// batches share one geometry (include/heffte_fft3d.h:392-402).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hmi_kernels.h"

namespace hmi {

// How the batched kernel runs one axis length: the compile-time tile of that length, or the
// runtime-plan tile with `lines` lines per tile and the given radix order.
struct batch_choice {
    bool table = false;
    int lines = 0;
    std::vector<int> radices;
};

// Picks the batch_choice of each (length, precision) by timing the candidates on the current
// GPU (a probe batch of n x n x m transforms, fastest of five runs) and keeps the result in
// memory and in a text cache file: $HMI_TUNE_CACHE, else $HOME/.cache/heffte_mi355/
// batch_tuning.txt (one line per choice: "n dbl table lines r1,r2,...").
class batch_tuner {
  public:
    static batch_choice choose(int n, bool dbl);
    // the untimed default: the compile-time tile if there is one, else the runtime plan with
    // as many lines as fit 40 KiB of LDS
    static batch_choice fallback(int n, bool dbl);
    static std::vector<batch_choice> candidates(int n, bool dbl);
    // milliseconds of the probe batch run with choice c for length n
    static double time_choice(int n, bool dbl, const batch_choice &c);
};

class batch_plan {
  public:
    // sizes: 3 per transform (x fastest); offsets in complex elements (nullptr = packed)
    batch_plan(int count, const int *sizes, const long long *in_off, const long long *out_off);
    ~batch_plan();

    void set_stream(hipStream_t s) { stream_ = s; }
    void set_graph(bool on) { use_graph_ = on; }
    // off: every length takes batch_tuner::fallback (no timing runs on the first execute)
    void set_autotune(bool on) { autotune_ = on; }
    // off: every transform takes the separate x and y phases (the fused x-y phase is on by default)
    void set_fused(bool on) { fused_ = on; }
    // transforms the fused x-y phase takes (the work lists are built on the first execute of each
    // precision; -1 before that)
    int fused_count(bool dbl) const { return dev_[dbl ? 1 : 0].ready ? dev_[dbl ? 1 : 0].nfused : -1; }
    int count() const { return (int)xf_.size(); }
    long long input_elems() const { return in_extent_; }
    long long output_elems() const { return out_extent_; }
    // dir: -1 forward, +1 backward.  Aliasing rule: in == out only when every transform's input
    // and output offsets are equal (each transform in place on its own range); otherwise the
    // input and output element ranges must not overlap.  Violations throw std::invalid_argument.
    void execute(int dir, bool dbl, const void *in, void *out, int scaling);
    // number of hipGraph captures so far (a replayed call captures nothing)
    long long graph_captures() const { return captures_; }

  private:
    struct device_state {
        void *tw = nullptr;
        batch_xform *xf = nullptr;
        // work lists per axis and register class of tile (batch_kernel_class), each run by its
        // own build of the batched kernel
        batch_item *items[3][kBatchClasses] = {};
        int nitems[3][kBatchClasses] = {};
        // the x and y work lists restricted to the transforms the fused phase does not take
        batch_item *rest[2][kBatchClasses] = {};
        int nrest[2][kBatchClasses] = {};
        // fused x-y work lists: one per small build (register class), the last for the
        // 1024-thread build; planes per work item set by xy_budget, LDS bytes per block
        batch_item *xy[kBatchClasses + 1] = {};
        int nxy[kBatchClasses + 1] = {};
        int xy_budget[kBatchClasses + 1] = {};
        size_t xy_lds[kBatchClasses + 1] = {};
        // column-block z phase: one list per register class (xy_budget = kXySmallBudget), and the
        // z work list of the transforms it does not take
        batch_item *zc[kBatchClasses] = {};
        int nzc[kBatchClasses] = {};
        batch_item *zrest[kBatchClasses] = {};
        int nzrest[kBatchClasses] = {};
        int nfused = 0;
        generic_plan *plans = nullptr;
        int *twoff = nullptr;
        int *kid = nullptr;
        size_t lds[kBatchClasses] = {};
        bool ready = false;
    };
    struct graph_key {
        int dir, scaling;
        const void *in;
        void *out;
        hipStream_t stream;
        bool fused;
        bool operator<(const graph_key &o) const {
            if (dir != o.dir) return dir < o.dir;
            if (scaling != o.scaling) return scaling < o.scaling;
            if (fused != o.fused) return fused < o.fused;
            if (in != o.in) return std::less<const void *>()(in, o.in);
            if (out != o.out) return std::less<void *>()(out, o.out);
            return std::less<hipStream_t>()(stream, o.stream);
        }
    };
    // instantiated graphs kept per precision; beyond this many the stream is synchronised and
    // the cache is emptied before a new capture
    static constexpr size_t kMaxGraphs = 16;
    void prepare(bool dbl);
    void launch_phases(int dir, bool dbl, const void *in, void *out, int scaling, hipStream_t s);

    std::vector<batch_xform> xf_;
    std::vector<int> lengths_;  // distinct axis lengths, index = size_id
    long long in_extent_ = 0, out_extent_ = 0;
    long long in_first_ = 0, out_first_ = 0;  // smallest input / output offset
    bool same_offsets_ = true;                // in_off == out_off for every transform
    hipStream_t stream_ = nullptr;
    bool use_graph_ = true;
    bool autotune_ = true;
    bool fused_ = true;
    std::map<int, batch_choice> forced_;  // probe plans of the tuner
    friend class batch_tuner;
    device_state dev_[2];
    std::map<graph_key, hipGraphExec_t> graphs_[2];
    long long captures_ = 0;
    hipStream_t capture_stream_ = nullptr;
    std::mutex mtx_;
};

}  // namespace hmi
