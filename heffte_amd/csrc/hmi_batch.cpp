// Batched mixed-size c2c transforms (see hmi_batch.h).
#include "hmi_batch.h"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "hmi_plan.h"

namespace hmi {

namespace {
void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <typename T> T *upload(const std::vector<T> &v) {
    T *d = nullptr;
    check_hip(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)), "hipMalloc(batch tables)");
    if (!v.empty()) check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}
}  // namespace

batch_plan::batch_plan(int count, const int *sizes, const long long *in_off, const long long *out_off) {
    if (count < 1) throw std::invalid_argument("a batch needs at least one transform");
    if (!sizes) throw std::invalid_argument("null size array");
    xf_.resize(count);
    long long in_pack = 0, out_pack = 0;
    for (int i = 0; i < count; i++) {
        batch_xform &x = xf_[i];
        long long cnt = 1;
        for (int d = 0; d < 3; d++) {
            const int n = sizes[3 * i + d];
            if (n < 1) throw std::invalid_argument("transform " + std::to_string(i) + " has an empty side");
            x.n[d] = n;
            cnt *= n;
            auto it = std::find(lengths_.begin(), lengths_.end(), n);
            if (it == lengths_.end()) {
                lengths_.push_back(n);
                x.size_id[d] = (int)lengths_.size() - 1;
            } else {
                x.size_id[d] = (int)(it - lengths_.begin());
            }
        }
        if ((long long)x.n[0] * x.n[1] > 0x7fffffffLL || (long long)x.n[1] * x.n[2] > 0x7fffffffLL ||
            (long long)x.n[0] * x.n[2] > 0x7fffffffLL)
            throw std::invalid_argument("transform too large for a batch entry");
        x.in_off = in_off ? in_off[i] : in_pack;
        x.out_off = out_off ? out_off[i] : (in_off ? in_off[i] : out_pack);
        if (x.in_off < 0 || x.out_off < 0) throw std::invalid_argument("negative offset");
        in_pack += cnt;
        out_pack += cnt;
        in_extent_ = std::max(in_extent_, x.in_off + cnt);
        out_extent_ = std::max(out_extent_, x.out_off + cnt);
    }
    for (int n : lengths_)
        if (2LL * n * 16 > kGenericMaxLds)
            throw std::invalid_argument("batch entry side " + std::to_string(n) + " exceeds the single-pass LDS capacity");
}

batch_plan::~batch_plan() {
    for (auto &g : graph_)
        if (g) (void)hipGraphExecDestroy(g);
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
    for (auto &d : dev_) {
        (void)hipFree(d.tw);
        (void)hipFree(d.xf);
        for (auto *p : d.items) (void)hipFree(p);
        (void)hipFree(d.plans);
        (void)hipFree(d.twoff);
        (void)hipFree(d.kid);
    }
}

void batch_plan::prepare(bool dbl) {
    device_state &d = dev_[dbl ? 1 : 0];
    if (d.ready) return;
    const size_t elem = dbl ? 16 : 8;
    // one radix plan and twiddle table per distinct length; lines per tile bounded by 64 KiB of LDS
    std::vector<generic_plan> plans;
    std::vector<int> twoff, kids;
    std::vector<float> tf;
    std::vector<double> td;
    size_t lds = 0;
    for (int n : lengths_) {
        batch_kernel_info info{};
        const int kid = find_batch_kernel(n, &info);
        generic_plan g{};
        if (kid >= 0) {
            // compile-time tile: only its line count is read from the plan
            g.n = n;
            g.lines = dbl ? info.lines_f64 : info.lines_f32;
            lds = std::max(lds, dbl ? info.lds_f64 : info.lds_f32);
        } else {
            const size_t per_line = 2 * (size_t)n * elem;
            // keep every tile of the batch within 40 KiB of LDS: several resident blocks per CU
            const int lines = (int)std::min<size_t>(64, std::max<size_t>(1, (40 * 1024) / per_line));
            g = make_generic_plan(n, lines);
            lds = std::max(lds, per_line * lines);
        }
        if (dbl) {
            twoff.push_back((int)(td.size() / 2));
            auto t = kid >= 0 ? stage_table_f64(info.radix, info.nstages) : generic_table_f64(g);
            td.insert(td.end(), t.begin(), t.end());
        } else {
            twoff.push_back((int)(tf.size() / 2));
            auto t = kid >= 0 ? stage_table_f32(info.radix, info.nstages) : generic_table_f32(g);
            tf.insert(tf.end(), t.begin(), t.end());
        }
        plans.push_back(g);
        kids.push_back(kid);
    }
    // work lists: every tile of every transform, one list per axis
    for (int ax = 0; ax < 3; ax++) {
        std::vector<batch_item> items;
        for (int i = 0; i < (int)xf_.size(); i++) {
            const batch_xform &x = xf_[i];
            const int col = (ax == 0) ? 1 : 0, b1 = 3 - ax - col;
            const long long nlines = (long long)x.n[col] * x.n[b1];
            const int lines = plans[x.size_id[ax]].lines;
            const long long ntiles = (nlines + lines - 1) / lines;
            for (long long t = 0; t < ntiles; t++) items.push_back({i, (int)t});
        }
        if (items.size() > 0x7fffffffULL) throw std::runtime_error("batch work list too long");
        d.nitems[ax] = (int)items.size();
        d.items[ax] = upload(items);
    }
    d.plans = upload(plans);
    d.twoff = upload(twoff);
    d.kid = upload(kids);
    d.xf = upload(xf_);
    if (dbl) d.tw = upload(td);
    else d.tw = upload(tf);
    d.lds = lds;
    d.ready = true;
}

void batch_plan::launch_phases(int dir, bool dbl, const void *in, void *out, int scaling, hipStream_t s) {
    device_state &d = dev_[dbl ? 1 : 0];
    for (int ax = 0; ax < 3; ax++) {
        batch_args b{};
        b.in = in;
        b.out = out;
        b.tw = d.tw;
        b.xforms = d.xf;
        b.items = d.items[ax];
        b.plans = d.plans;
        b.twoff = d.twoff;
        b.kid = d.kid;
        b.nitems = d.nitems[ax];
        b.axis = ax;
        b.from_in = (ax == 0) ? 1 : 0;
        b.conj = dir > 0 ? 1 : 0;
        b.scaling = (ax == 2) ? scaling : 0;
        // persistent grid: a few resident blocks per CU, never more blocks than work items
        const int nblocks = std::min(b.nitems, 2048);
        if (nblocks > 0) check_hip(launch_batched(dbl, b, nblocks, d.lds, s), "batched FFT phase");
    }
}

void batch_plan::execute(int dir, bool dbl, const void *in, void *out, int scaling) {
    if (!in || !out) throw std::invalid_argument("null input or output array");
    std::lock_guard<std::mutex> lock(mtx_);
    prepare(dbl);
    if (!use_graph_) {
        launch_phases(dir, dbl, in, out, scaling, stream_);
        return;
    }
    const int g = dbl ? 1 : 0;
    graph_key key{dir, scaling, in, const_cast<void *>(out), stream_};
    if (!graph_[g] || !(gkey_[g] == key)) {
        // capture the three phase launches once per (direction, scale, arrays) and replay them
        if (graph_[g]) {
            check_hip(hipGraphExecDestroy(graph_[g]), "hipGraphExecDestroy");
            graph_[g] = nullptr;
        }
        if (!capture_stream_) check_hip(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "hipStreamCreate");
        hipGraph_t graph = nullptr;
        check_hip(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch_phases(dir, dbl, in, out, scaling, capture_stream_);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(capture_stream_, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        check_hip(hipStreamEndCapture(capture_stream_, &graph), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check_hip(e, "hipGraphInstantiate");
        graph_[g] = exec;
        gkey_[g] = key;
    }
    check_hip(hipGraphLaunch(graph_[g], stream_), "hipGraphLaunch");
}

}  // namespace hmi
