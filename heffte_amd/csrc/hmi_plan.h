// Host-side plan of a single-rank 3D transform on one MI355X.
//
// A plan is what heFFTe builds in fft3d::setup (include/heffte_fft3d.h:599-651)
// for a communicator of size one: the reshapes and 1D executors of
// compute_transform (src/heffte_compute_transform.cpp:15-258), re-expressed as
// a short list of fused passes.  Each pass transforms one axis and writes the
// layout the next pass reads best; the reshapes are folded into the load or
// the store of a pass and the scale into the store of the last pass.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hmi_kernels.h"

namespace hmi {

enum class kind { c2c = 0, r2c = 1, cos = 2, sin = 3, cos1 = 4 };
enum class buf { in = 0, out = 1, ws = 2 };

struct box3 {
    long long low[3], high[3];
    int order[3];
    long long size(int d) const { return high[d] - low[d] + 1; }
    long long count() const {
        for (int d = 0; d < 3; d++)
            if (size(d) <= 0) return 0;
        return size(0) * size(1) * size(2);
    }
    bool empty() const { return count() == 0; }
    bool same_indexes(const box3 &o) const {
        for (int d = 0; d < 3; d++)
            if (low[d] != o.low[d] || high[d] != o.high[d]) return false;
        return true;
    }
};

// Memory layout of a 3D array: element (i0, i1, i2) at sum_d i_d * stride[d].
struct layout {
    long long n[3];
    long long stride[3];
    static layout ordered(const long long n[3], const int order[3]);
    int fastest() const;  // axis with the smallest stride among axes with n > 1 (-1 if none)
};

struct pass {
    int axis = 0, n = 1;
    buf src = buf::in, dst = buf::out;
    layout lin{}, lout{};
    int col = 0, b1 = 0;  // column axis and third axis
    int load_colfast = 0, store_colfast = 0;
    int in_real = 0, out_real = 0, herm_load = 0, store_len = 0;
    bool last = false;
    bool identity = false;  // copy pass (no FFT stages), used when no in-place schedule exists
    bool real_axis = false; // the r2c / c2r pass of a fft3d_r2c plan
    int mode = 0;           // kModeC2C, or kModeR2C / kModeC2R when the half-length kernel runs it
    // launch configuration
    const fast_kernel_info *fast = nullptr;
    generic_plan gen{};
    int lines = 1;  // lines (columns) per block
    // A length whose line does not fit one LDS tile runs as two passes (four-step), n = n1 n2,
    // position j = n2 j1 + j2, bin k = k1 + n1 k2: the first step takes the n1-point lines of
    // stride n2 (one per j2) into a buffer of the pass's input layout at position n2 k1 + j2,
    // each value twiddled by w_n^(j2 k1); the second takes the n2-point lines (one per k1) of that
    // buffer to the destination at k1 + n1 k2.  split holds the kernel configuration of the two
    // steps; the buffer lies in the workspace at split_ws (elements per transform).
    int split_n1 = 0, split_n2 = 0;
    std::shared_ptr<const std::array<pass, 2>> split;
    long long split_ws = 0;
};

struct exec_plan {
    std::vector<pass> passes;
    long long ws_elems = 0;  // workspace needed per transform, in complex elements
};

// Device twiddle tables for one length and precision.
struct twiddles {
    void *dev = nullptr;
    size_t bytes = 0;
};

class fft_plan {
  public:
    fft_plan(kind k, const box3 &inbox, const box3 &outbox, int r2c_dir);
    ~fft_plan();

    kind get_kind() const { return kind_; }
    long long size_inbox() const { return inbox_.count(); }
    long long size_outbox() const { return outbox_.count(); }
    long long size_workspace() const { return ws_elems_; }
    int r2c_dir() const { return r2c_dir_; }
    double scale_factor(int scaling) const;
    void set_stream(hipStream_t s) { stream_ = s; }
    hipStream_t stream() const { return stream_; }
    // kPrecPlain, kSplitF16 or kSplitBF16: the form of the 512-point complex<float> c2c passes
    void set_precision(int p) { precision_ = p; }
    int precision() const { return precision_; }
    // passes of the complex<float> transform in direction dir that run on the matrix cores
    // under the current precision form; *total = all passes of that transform
    int matrix_core_passes(int dir, bool in_place, int *total);

    // dir: -1 forward, +1 backward.  in_real/out_real select the real-valued
    // overloads of fft3d (real input on forward, real output on backward).  mul (c2c only):
    // the first pass loads in[i] * mul[i] instead of in[i] (mul laid out like in).
    // no_alloc: the call allocates nothing -- ws must be null (the workspace is the static
    // scratch of the current device, kScratchBytes at most) and the device tables must be
    // resident (prepare); otherwise the call throws before it launches anything.  A call on a
    // stream under capture is held to the same rule for the tables (their upload is a
    // synchronous copy, which a capture cannot record).
    void execute(int dir, bool dbl, bool in_real, bool out_real, int batch, const void *in, void *out, void *ws,
                 int scaling, const void *mul = nullptr, bool no_alloc = false);
    // Uploads the device tables (twiddles, split-precision operands of the current precision)
    // of every variant of the plan in one precision, so that later calls upload nothing.
    void prepare(bool dbl);
    // out = backward(forward(x) * forward(y)) with the backward scaling (scale::full gives the
    // circular convolution of x and y); the product is fused into the first backward pass,
    // and x == y runs one forward transform.
    // c2c plans only; ws holds batch * convolve_workspace() elements or is null (allocated).
    void convolve(bool dbl, int batch, const void *x, const void *y, void *out, void *ws, int scaling);
    long long convolve_workspace() const { return ws_elems_ + outbox_.count(); }
    // out = backward(forward(x) * yhat) with yhat a spectrum the caller computed once (forward
    // of the second operand, outbox layout, no scale): one forward and one backward per call.
    // ws holds batch * size_workspace() elements or is null; out must not alias yhat.
    void convolve_spectrum(bool dbl, int batch, const void *x, const void *yhat, void *out, void *ws, int scaling);

    // Plan (without running it): exposed for tests of the planner.
    const exec_plan &get_exec_plan(int dir, bool dbl, bool in_real, bool out_real, bool in_place);

    // Per-pass timing: while on, every call records a hipEvent on the plan's stream before its
    // first pass and after each pass, so the time between two consecutive events is the
    // duration of one pass launch.  timing_report() waits for the recorded events and returns
    // one line per pass kernel, "name<TAB>launches<TAB>total ms", aggregated over every call
    // since timing was switched on; switching it off (or on again) drops the records.
    void set_timing(bool on);
    std::string timing_report();

  private:
    exec_plan build(int dir, bool dbl, bool in_real, bool out_real, bool in_place) const;
    void configure(pass &p, bool dbl) const;
    const twiddles &get_twiddles(bool dbl, const pass &p, bool may_alloc);
    const twiddles &get_split_table(bool dbl, int n, bool may_alloc);
    void launch_split(const pass &p, const pass_args &base, bool dbl, int batch, void *tmp);
    const twiddles &get_mfma_table(int prec, bool conj, bool may_alloc);
    template <typename F> void for_each_variant(F &&f) const;

    kind kind_;
    box3 inbox_, outbox_;
    int r2c_dir_;
    long long ws_elems_ = 0;
    hipStream_t stream_ = nullptr;
    int precision_ = kPrecPlain;
    std::mutex mtx_;
    twiddles mfma_tw_[2][2] = {};  // [split FP16, split BF16][forward, backward]
    std::map<int, exec_plan> plans_;
    std::map<long long, twiddles> tw_;
    // per-pass timing records: launch i ran between events marks_[i].first and .second
    bool timing_ = false;
    std::vector<std::string> timing_names_;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timing_launches_;
    std::vector<hipEvent_t> timing_events_;
    void clear_timing();
};

// Runtime radix plan of length n for the generic tile (stage/root/post offsets
// filled in; lines = lines per tile) and its twiddle table in either precision.
generic_plan make_generic_plan(int n, int lines);
// The same with an explicit radix order (their product must be n).
generic_plan make_generic_plan(int n, int lines, const std::vector<int> &radices);
// The radix order the planner uses by default: 8s, then a 4 and a 2, then odd primes.
std::vector<int> default_radices(int n);
std::vector<float> generic_table_f32(const generic_plan &g);
std::vector<double> generic_table_f64(const generic_plan &g);
// Stage twiddles of a compile-time radix plan (no root or real-data tables).
std::vector<float> stage_table_f32(const int *radix, int nstages);
std::vector<double> stage_table_f64(const int *radix, int nstages);

// Launch arguments of pass p of plan ep (the split steps derive theirs from these).
pass_args pass_launch_args(kind k, const exec_plan &ep, const pass &p, int dir, bool dbl, int batch, const void *in,
                           void *out, void *ws, const void *tw, double scale, const void *mul);
// Launch arguments of the two steps of a split pass for transform bt of a batch (tables not
// set): base = the arguments the pass would have unsplit, tmp = the step buffer.
void split_steps(const pass &p, const pass_args &base, bool dbl, int bt, void *tmp, pass_args step[2]);

// Permutation of a box between two orders (reshape3d_transpose semantics).
void permute(int elem_bytes, const box3 &inbox, const int out_order[3], int batch, const void *in, void *out,
             hipStream_t s);

}  // namespace hmi
