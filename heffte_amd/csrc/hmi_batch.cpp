// Batched mixed-size c2c transforms (see hmi_batch.h).
#include "hmi_batch.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
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
        in_first_ = i == 0 ? x.in_off : std::min(in_first_, x.in_off);
        out_first_ = i == 0 ? x.out_off : std::min(out_first_, x.out_off);
        same_offsets_ = same_offsets_ && x.in_off == x.out_off;
    }
    // every phase writes each transform's output range while the others are read and written:
    // output ranges must be disjoint
    {
        std::vector<std::pair<long long, long long>> r(count);
        for (int i = 0; i < count; i++)
            r[i] = {xf_[i].out_off, xf_[i].out_off + (long long)xf_[i].n[0] * xf_[i].n[1] * xf_[i].n[2]};
        std::sort(r.begin(), r.end());
        for (int i = 1; i < count; i++)
            if (r[i].first < r[i - 1].second)
                throw std::invalid_argument("the output ranges of two transforms of the batch overlap");
    }
    for (int n : lengths_)
        if (2LL * n * 16 > kGenericMaxLds)
            throw std::invalid_argument("batch entry side " + std::to_string(n) + " exceeds the single-pass LDS capacity");
}

batch_plan::~batch_plan() {
    bool any = false;
    for (auto &m : graphs_) any = any || !m.empty();
    if (any) (void)hipStreamSynchronize(stream_);  // a replay may still run
    for (auto &m : graphs_)
        for (auto &kv : m) (void)hipGraphExecDestroy(kv.second);
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
    for (auto &d : dev_) {
        (void)hipFree(d.tw);
        (void)hipFree(d.xf);
        for (auto &ax : d.items)
            for (auto *p : ax) (void)hipFree(p);
        for (auto &ax : d.rest)
            for (auto *p : ax) (void)hipFree(p);
        for (auto *p : d.xy) (void)hipFree(p);
        for (auto *p : d.zc) (void)hipFree(p);
        for (auto *p : d.zrest) (void)hipFree(p);
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
    size_t lds[kBatchClasses] = {};
    for (int n : lengths_) {
        auto fc = forced_.find(n);
        const batch_choice c = fc != forced_.end() ? fc->second
                                                   : (autotune_ ? batch_tuner::choose(n, dbl) : batch_tuner::fallback(n, dbl));
        batch_kernel_info info{};
        const int kid = c.table ? find_batch_kernel(n, &info) : -1;
        generic_plan g{};
        if (kid >= 0) {
            // compile-time tile: only its line count is read from the plan
            g.n = n;
            g.lines = dbl ? info.lines_f64 : info.lines_f32;
            size_t &l = lds[batch_kernel_class(kid)];
            l = std::max(l, dbl ? info.lds_f64 : info.lds_f32);
        } else {
            const size_t per_line = 2 * (size_t)n * elem;
            g = make_generic_plan(n, std::max(1, c.lines), c.radices.empty() ? default_radices(n) : c.radices);
            size_t &l = lds[batch_kernel_class(-1)];
            l = std::max(l, per_line * g.lines);
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
    // Fused x-y phase: a transform whose x and y lengths both have compile-time tiles of register
    // classes up to kXyMaxClass takes the 256-thread build when one plane fits its budget, else
    // the 512-thread build when a plane fits that one; work item = as many consecutive z-planes
    // as the budget holds (at least one).
    const int pr = dbl ? 1 : 0;
    std::vector<int> xy_build(xf_.size(), -1);
    {
        std::vector<batch_item> xy[kBatchClasses + 1];
        long long held[kBatchClasses + 1] = {};
        for (int i = 0; i < (int)xf_.size(); i++) {
            const batch_xform &x = xf_[i];
            const int k0 = kids[x.size_id[0]], k1 = kids[x.size_id[1]];
            if (!batch_xy_supported(k0) || !batch_xy_supported(k1)) continue;
            const int cls = std::max(batch_kernel_class(k0), batch_kernel_class(k1));
            const long long plane = (long long)xy_row_stride(x.n[0]) * x.n[1];  // LDS elements
            int b = -1;
            // the 256-thread build is one build carrying every tile (kXyMaxClass): one launch for
            // the items of all classes (at 2 waves per SIMD and two blocks per CU the larger
            // register budget costs nothing; separate launches per class added their tails)
            if (cls <= kXyMaxClass[pr] && plane <= kXySmallBudget[pr]) b = kXyMaxClass[pr];
            else if (cls <= kXyLargeMaxClass[pr] && plane <= kXyLargeBudget[pr]) b = kBatchClasses;
            if (b < 0) continue;
            xy_build[i] = b;
            const long long budget = b == kBatchClasses ? kXyLargeBudget[pr] : kXySmallBudget[pr];
            const long long g = std::max(1LL, budget / plane);
            for (long long t = 0; t * g < x.n[2]; t++) xy[b].push_back({i, (int)t});
            held[b] = std::max(held[b], std::min<long long>(g, x.n[2]) * plane);
        }
        d.nfused = 0;
        for (int b : xy_build) d.nfused += b >= 0 ? 1 : 0;
        // largest items first (the blocks take them in list order, so the launch ends on small
        // ones: 3.925 -> 3.910 ms per pair on the 1000-transform batch)
        for (int b = 0; b <= kBatchClasses; b++) {
            const long long budget = b == kBatchClasses ? kXyLargeBudget[pr] : kXySmallBudget[pr];
            auto elems = [&](const batch_item &it) {
                const batch_xform &x = xf_[it.xform];
                const long long plane = (long long)x.n[0] * x.n[1];
                const long long g = std::max(1LL, budget / ((long long)xy_row_stride(x.n[0]) * x.n[1]));
                return std::min<long long>(g, x.n[2] - (long long)it.tile * g) * plane;
            };
            std::stable_sort(xy[b].begin(), xy[b].end(),
                             [&](const batch_item &a, const batch_item &c) { return elems(a) > elems(c); });
            if (xy[b].size() > 0x7fffffffULL) throw std::runtime_error("batch work list too long");
            d.nxy[b] = (int)xy[b].size();
            d.xy[b] = xy[b].empty() ? nullptr : upload(xy[b]);
            d.xy_budget[b] = b == kBatchClasses ? kXyLargeBudget[pr] : kXySmallBudget[pr];
            d.xy_lds[b] = (size_t)held[b] * elem;
        }
    }
    // Column-block z phase: a transform whose z length has a compile-time tile of a class the
    // 256-thread builds run; work item = zcol_width(n2, budget, n0 n1) consecutive (x, y)
    // positions.  The others keep the tiled z phase.
    std::vector<bool> zcol(xf_.size(), false);
    {
        std::vector<batch_item> zc[kBatchClasses];
        for (int i = 0; i < (int)xf_.size(); i++) {
            const batch_xform &x = xf_[i];
            const int kz = kids[x.size_id[2]];
            if (!batch_xy_supported(kz)) continue;
            const int cls = batch_kernel_class(kz);
            if (cls > kXyMaxClass[pr] || (long long)x.n[2] > kXySmallBudget[pr]) continue;
            const long long plane = (long long)x.n[0] * x.n[1];
            if (plane > 0x7fffffffLL / 2) continue;
            const long long w = zcol_width(x.n[2], kXySmallBudget[pr], (int)plane);
            zcol[i] = true;
            for (long long t = 0; t * w < plane; t++) zc[kXyMaxClass[pr]].push_back({i, (int)t});
        }
        for (int c = 0; c < kBatchClasses; c++) {
            auto elems = [&](const batch_item &it) {
                const batch_xform &x = xf_[it.xform];
                const long long plane = (long long)x.n[0] * x.n[1];
                const long long w = zcol_width(x.n[2], kXySmallBudget[pr], (int)plane);
                return std::min<long long>(w, plane - (long long)it.tile * w) * x.n[2];
            };
            std::stable_sort(zc[c].begin(), zc[c].end(),
                             [&](const batch_item &a, const batch_item &b) { return elems(a) > elems(b); });
            if (zc[c].size() > 0x7fffffffULL) throw std::runtime_error("batch work list too long");
            d.nzc[c] = (int)zc[c].size();
            d.zc[c] = zc[c].empty() ? nullptr : upload(zc[c]);
        }
    }
    // work lists: every tile of every transform, one list per axis and register class; for the
    // x and y axes also the lists without the transforms of the fused phase, for z the list
    // without those of the column-block phase
    for (int ax = 0; ax < 3; ax++) {
        std::vector<batch_item> items[kBatchClasses], rest[kBatchClasses];  // rest: z phase without zcol
        for (int i = 0; i < (int)xf_.size(); i++) {
            const batch_xform &x = xf_[i];
            const int col = (ax == 0) ? 1 : 0, b1 = 3 - ax - col;
            const long long nlines = (long long)x.n[col] * x.n[b1];
            const int lines = plans[x.size_id[ax]].lines;
            const long long ntiles = (nlines + lines - 1) / lines;
            const int c = batch_kernel_class(kids[x.size_id[ax]]);
            for (long long t = 0; t < ntiles; t++) items[c].push_back({i, (int)t});
            if (ax < 2 ? xy_build[i] < 0 : !zcol[i])
                for (long long t = 0; t < ntiles; t++) rest[c].push_back({i, (int)t});
        }
        for (int c = 0; c < kBatchClasses; c++) {
            if (items[c].size() > 0x7fffffffULL) throw std::runtime_error("batch work list too long");
            d.nitems[ax][c] = (int)items[c].size();
            d.items[ax][c] = items[c].empty() ? nullptr : upload(items[c]);
            if (ax < 2) {
                d.nrest[ax][c] = (int)rest[c].size();
                d.rest[ax][c] = rest[c].empty() ? nullptr : upload(rest[c]);
            } else {
                d.nzrest[c] = (int)rest[c].size();
                d.zrest[c] = rest[c].empty() ? nullptr : upload(rest[c]);
            }
        }
    }
    d.plans = upload(plans);
    d.twoff = upload(twoff);
    d.kid = upload(kids);
    d.xf = upload(xf_);
    if (dbl) d.tw = upload(td);
    else d.tw = upload(tf);
    for (int c = 0; c < kBatchClasses; c++) d.lds[c] = lds[c];
    d.ready = true;
}

void batch_plan::launch_phases(int dir, bool dbl, const void *in, void *out, int scaling, hipStream_t s) {
    device_state &d = dev_[dbl ? 1 : 0];
    batch_args b{};
    b.in = in;
    b.out = out;
    b.tw = d.tw;
    b.xforms = d.xf;
    b.plans = d.plans;
    b.twoff = d.twoff;
    b.kid = d.kid;
    b.conj = dir > 0 ? 1 : 0;
    // at most 2^32 - 1 work-items in a grid: beyond that the grid stride covers the rest
    const int max_blocks = (int)(0xffffffffLL / kXyLargeThreads);
    const bool fused = fused_ && d.nfused > 0;
    if (fused) {
        // fused x-y phase: reads the input arrays, writes the output arrays
        b.axis = 0;
        b.from_in = 1;
        b.scaling = 0;
        for (int k = 0; k <= kBatchClasses; k++) {
            b.items = d.xy[k];
            b.nitems = d.nxy[k];
            b.xy_budget = d.xy_budget[k];
            const int nblocks = std::min(b.nitems, max_blocks);
            if (nblocks > 0)
                check_hip(launch_batched_xy(dbl, k == kBatchClasses, k == kBatchClasses ? kXyLargeMaxClass[0] : k, b,
                                            nblocks, d.xy_lds[k], s),
                          "batched fused x-y phase");
        }
    }
    for (int ax = 0; ax < 3; ax++) {
        b.axis = ax;
        b.from_in = (ax == 0) ? 1 : 0;
        b.scaling = (ax == 2) ? scaling : 0;
        if (fused && ax == 2) {
            // column-block z phase: reads and writes the output arrays, applies the scale
            b.xy_budget = kXySmallBudget[dbl ? 1 : 0];
            for (int c = 0; c < kBatchClasses; c++) {
                b.items = d.zc[c];
                b.nitems = d.nzc[c];
                const int nblocks = std::min(b.nitems, max_blocks);
                if (nblocks > 0)
                    check_hip(launch_batched_zcol(dbl, c, b, nblocks, (size_t)b.xy_budget * (dbl ? 16 : 8), s),
                              "batched column-block z phase");
            }
        }
        b.xy_budget = 0;
        for (int c = 0; c < kBatchClasses; c++) {
            const bool rest = fused;
            b.items = !rest ? d.items[ax][c] : (ax < 2 ? d.rest[ax][c] : d.zrest[c]);
            b.nitems = !rest ? d.nitems[ax][c] : (ax < 2 ? d.nrest[ax][c] : d.nzrest[c]);
            // one work item per block: the hardware schedules the blocks as others retire (a
            // persistent grid of 2048 blocks walking the list was 13 % slower on the
            // 1000-transform batch); the kernel walks the list with a grid stride, so any grid
            // size is valid
            const int nblocks = std::min(b.nitems, max_blocks);
            if (nblocks > 0) check_hip(launch_batched(dbl, c, b, nblocks, d.lds[c], s), "batched FFT phase");
        }
    }
}

void batch_plan::execute(int dir, bool dbl, const void *in, void *out, int scaling) {
    if (!in || !out) throw std::invalid_argument("null input or output array");
    {
        // aliasing: phase 0 reads every input range while it writes every output range
        const size_t elem = dbl ? 16 : 8;
        if (in == out) {
            if (!same_offsets_)
                throw std::invalid_argument("in-place batch calls need equal input and output offsets per transform");
        } else {
            const uintptr_t i0 = reinterpret_cast<uintptr_t>(in) + (uintptr_t)in_first_ * elem;
            const uintptr_t i1 = reinterpret_cast<uintptr_t>(in) + (uintptr_t)in_extent_ * elem;
            const uintptr_t o0 = reinterpret_cast<uintptr_t>(out) + (uintptr_t)out_first_ * elem;
            const uintptr_t o1 = reinterpret_cast<uintptr_t>(out) + (uintptr_t)out_extent_ * elem;
            if (i0 < o1 && o0 < i1)
                throw std::invalid_argument("the input and output ranges of the batch overlap (and in != out)");
        }
    }
    std::lock_guard<std::mutex> lock(mtx_);
    prepare(dbl);
    if (!use_graph_) {
        launch_phases(dir, dbl, in, out, scaling, stream_);
        return;
    }
    const int g = dbl ? 1 : 0;
    graph_key key{dir, scaling, in, const_cast<void *>(out), stream_, fused_};
    auto found = graphs_[g].find(key);
    if (found == graphs_[g].end()) {
        // capture the three phase launches once per (direction, scale, arrays, stream) and replay
        // them; a full cache is emptied first, after the replays in flight have finished
        if (graphs_[g].size() >= kMaxGraphs) {
            check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            for (auto &kv : graphs_[g]) (void)hipGraphExecDestroy(kv.second);
            graphs_[g].clear();
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
        found = graphs_[g].emplace(key, exec).first;
        captures_++;
    }
    check_hip(hipGraphLaunch(found->second, stream_), "hipGraphLaunch");
}


// ------------------------------------------------------------------ tuner

namespace {

std::mutex g_tune_mtx;
std::map<std::string, std::map<long long, batch_choice>> g_tuned;  // cache path -> (n, dbl) -> choice

long long tune_key(int n, bool dbl) { return 2LL * n + (dbl ? 1 : 0); }

std::string cache_path() {
    if (const char *p = std::getenv("HMI_TUNE_CACHE")) return p;
    if (const char *h = std::getenv("HOME")) return std::string(h) + "/.cache/heffte_mi355/batch_tuning.txt";
    return "";
}

// Choices are timed with the kernels of one build of the library: the cache file starts with
// this line, which names the build, and a file written by another build is ignored (and
// replaced by the next save).
const std::string &cache_header() {
    static const std::string h =
        std::string("# heffte_mi355 batch tuning, build ") + __DATE__ " " __TIME__ + " / " + kernels_build_stamp();
    return h;
}

std::map<long long, batch_choice> load_cache(const std::string &path) {
    std::map<long long, batch_choice> m;
    if (path.empty()) return m;
    std::ifstream f(path);
    std::string line;
    if (!std::getline(f, line) || line != cache_header()) return m;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        int n = 0, dbl = 0, table = 0, lines = 0;
        std::string rads;
        if (!(in >> n >> dbl >> table >> lines >> rads) || n < 1) continue;
        batch_choice c;
        c.table = table != 0;
        c.lines = lines;
        long long prod = 1;
        std::istringstream rs(rads);
        std::string tok;
        while (std::getline(rs, tok, ','))
            if (!tok.empty() && tok != "-") {
                c.radices.push_back(std::atoi(tok.c_str()));
                prod *= c.radices.back();
            }
        // entries that no longer describe a valid plan are ignored (and tuned again)
        if (c.table ? find_batch_kernel(n, nullptr) < 0 : (lines < 1 || prod != n)) continue;
        m[tune_key(n, dbl != 0)] = c;
    }
    return m;
}

void save_cache(const std::string &path, const std::map<long long, batch_choice> &m) {
    if (path.empty()) return;
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) {
        // create the directory chain; failures show up as a failed open below
        for (size_t i = 1; i <= slash; i++)
            if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
    }
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream f(tmp);
        if (!f) return;  // a read-only location only loses the cache
        f << cache_header() << '\n';
        for (const auto &kv : m) {
            const batch_choice &c = kv.second;
            f << (kv.first / 2) << ' ' << (kv.first & 1) << ' ' << (c.table ? 1 : 0) << ' ' << c.lines << ' ';
            if (c.radices.empty()) f << '-';
            for (size_t i = 0; i < c.radices.size(); i++) f << (i ? "," : "") << c.radices[i];
            f << '\n';
        }
    }
    (void)std::rename(tmp.c_str(), path.c_str());
}

}  // namespace

// Fastest of five runs of a probe batch, forward, in place, on a stream of its own: one
// n x n2 x m transform of at most about 2^20 elements (n2 = n and m a power of two from 16 to 512
// for n <= 256; n2 = max(1, 2^16 / n) and m = 16 for longer sides), so the probe stays at most
// 16 MiB whatever the side.
double batch_tuner::time_choice(int n, bool dbl, const batch_choice &c) {
    long long m = 16, n2 = n;
    if (n <= 256) {
        while (m < 512 && (long long)n * n * m * 2 <= (1LL << 20)) m *= 2;
    } else {
        n2 = std::max(1, (1 << 16) / n);
    }
    const int sizes[3] = {n, (int)n2, (int)m};
    batch_plan probe(1, sizes, nullptr, nullptr);
    probe.set_graph(false);
    probe.set_autotune(false);
    probe.forced_[n] = c;
    const size_t bytes = (size_t)n * n2 * m * (dbl ? 16 : 8);
    void *buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = std::numeric_limits<double>::infinity();
    try {
        check_hip(hipMalloc(&buf, bytes), "hipMalloc(tuning probe)");
        check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate(tuning)");
        check_hip(hipMemsetAsync(buf, 0, bytes, s), "hipMemsetAsync(tuning probe)");
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        check_hip(hipEventCreate(&e1), "hipEventCreate");
        probe.set_stream(s);
        probe.execute(-1, dbl, buf, buf, 0);
        for (int r = 0; r < 5; r++) {
            check_hip(hipEventRecord(e0, s), "hipEventRecord");
            probe.execute(-1, dbl, buf, buf, 0);
            check_hip(hipEventRecord(e1, s), "hipEventRecord");
            check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            best = std::min(best, (double)ms);
        }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s);
        if (buf) (void)hipFree(buf);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipFree(buf);
    return best;
}

batch_choice batch_tuner::fallback(int n, bool dbl) {
    batch_choice c;
    if (find_batch_kernel(n, nullptr) >= 0) {
        c.table = true;
        return c;
    }
    const size_t per_line = 2 * (size_t)n * (dbl ? 16 : 8);
    // every tile of the batch within 40 KiB of LDS: several resident blocks per CU
    c.lines = (int)std::min<size_t>(64, std::max<size_t>(1, (40 * 1024) / per_line));
    c.radices = default_radices(n);
    return c;
}

std::vector<batch_choice> batch_tuner::candidates(int n, bool dbl) {
    std::vector<batch_choice> v;
    if (find_batch_kernel(n, nullptr) >= 0) {
        batch_choice t;
        t.table = true;
        v.push_back(t);
    }
    const size_t per_line = 2 * (size_t)n * (dbl ? 16 : 8);
    const int lmax = (int)std::min<size_t>(64, std::max<size_t>(1, (40 * 1024) / per_line));
    const std::vector<int> rad = default_radices(n);
    for (int lines : {lmax, lmax / 2, lmax / 4}) {
        if (lines < 1) continue;
        bool seen = false;
        for (const batch_choice &c : v) seen = seen || (!c.table && c.lines == lines);
        if (seen) continue;
        batch_choice c;
        c.lines = lines;
        c.radices = rad;
        v.push_back(c);
    }
    std::vector<int> rev(rad.rbegin(), rad.rend());
    if (rev != rad) {
        batch_choice c;
        c.lines = lmax;
        c.radices = rev;
        v.push_back(c);
    }
    return v;
}

batch_choice batch_tuner::choose(int n, bool dbl) {
    const std::string path = cache_path();
    std::lock_guard<std::mutex> lock(g_tune_mtx);
    auto pit = g_tuned.find(path);
    if (pit == g_tuned.end()) pit = g_tuned.emplace(path, load_cache(path)).first;
    auto it = pit->second.find(tune_key(n, dbl));
    if (it != pit->second.end()) return it->second;
    const std::vector<batch_choice> cands = candidates(n, dbl);
    batch_choice best = fallback(n, dbl);
    if (cands.size() > 1) {
        // A runtime plan replaces the compile-time tile only when it is clearly faster: the
        // probes are short, and a noisy timing (another process on the GPU, a profiler) must not
        // lock a slower plan into the cache.
        constexpr double kRuntimeMargin = 0.85;
        const double inf = std::numeric_limits<double>::infinity();
        double ttable = inf, trun = inf;
        batch_choice run;
        for (const batch_choice &c : cands) {
            double t = inf;
            try {
                t = time_choice(n, dbl, c);
            } catch (const std::runtime_error &) {
                // the probe could not run (device memory): keep the untimed default
                (void)hipGetLastError();
                return best;
            }
            if (c.table) ttable = t;
            else if (t < trun) {
                trun = t;
                run = c;
            }
        }
        if (trun < inf && (ttable == inf || trun < kRuntimeMargin * ttable)) best = run;
    }
    pit->second[tune_key(n, dbl)] = best;
    save_cache(path, pit->second);
    return best;
}

}  // namespace hmi
