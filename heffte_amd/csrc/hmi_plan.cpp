// Planner and executor of the single-rank 3D transform (see hmi_plan.h).
#include "hmi_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>

namespace hmi {

namespace {

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool is_permutation(const int o[3]) {
    bool seen[3] = {false, false, false};
    for (int i = 0; i < 3; i++) {
        if (o[i] < 0 || o[i] > 2 || seen[o[i]]) return false;
        seen[o[i]] = true;
    }
    return true;
}

const int kOrders[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Radix decomposition of the generic pass: 8s, then 4, 2, then odd primes.
std::vector<int> factorize(int n) {
    std::vector<int> r;
    int m = n;
    while (m % 8 == 0) { r.push_back(8); m /= 8; }
    if (m % 4 == 0) { r.push_back(4); m /= 4; }
    if (m % 2 == 0) { r.push_back(2); m /= 2; }
    for (int p = 3; m > 1; p += 2) {
        while (m % p == 0) { r.push_back(p); m /= p; }
        if ((long long)p * p > m && m > 1) { r.push_back(m); m = 1; }
    }
    return r;
}

bool has_butterfly(int r) {
    return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 11 || r == 13 || r == 16;
}

struct candidate_pass {
    int axis;
    bool identity;
};

// Extra cost of a pass whose load or store walks the column axis, by the width of its column
// chunks: the tile of a long line fills the LDS with few lines (2048 complex<float> points: 8
// lines, 64-byte chunks; 4096: 4 lines, 32 bytes), and a chunk narrower than a 128-byte cache
// line costs HBM efficiency.  Measured on MI355X per pass against the same axis with contiguous
// lines on both sides: +40 % at 64-byte chunks (2048 points, either precision), +150-175 % at 32
// bytes (4096 points).  Passes without a compile-time tile are not priced here.
double narrow_chunk_cost(const pass &p, bool dbl) {
    int n = p.n;
    if (p.mode < kModeDCT2 && p.real_axis && n % 2 == 0) n /= 2;
    const fast_kernel_info *f = find_fast_kernel(n, dbl, true);
    if (!f) return 0.0;
    const int chunk = f->T * (dbl ? 16 : 8);
    return chunk >= 128 ? 0.0 : chunk >= 64 ? 0.4 : chunk >= 32 ? 2.5 : 4.0;
}

// Factor pair n1 * n2 = n of a length beyond one LDS tile, both at most maxn: pairs whose two
// lengths have compile-time kernels first, then the most balanced pair.
bool split_length(int n, int maxn, bool dbl, int &n1, int &n2) {
    bool found = false;
    double best = 0.0;
    for (int d = 2; d <= maxn && d < n; d++) {
        if (n % d) continue;
        const int e = n / d;
        if (e > maxn || e < 2) continue;
        const int fast = (find_fast_kernel(d, dbl, true) ? 1 : 0) + (find_fast_kernel(e, dbl, true) ? 1 : 0);
        const double score = fast * 1000.0 - std::fabs(std::log((double)d) - std::log((double)e));
        if (!found || score > best) {
            found = true;
            best = score;
            n1 = d;
            n2 = e;
        }
    }
    return found;
}

}  // namespace

layout layout::ordered(const long long n[3], const int order[3]) {
    layout L;
    for (int d = 0; d < 3; d++) L.n[d] = n[d];
    L.stride[order[0]] = 1;
    L.stride[order[1]] = n[order[0]];
    L.stride[order[2]] = n[order[0]] * n[order[1]];
    return L;
}

int layout::fastest() const {
    int best = -1;
    for (int d = 0; d < 3; d++)
        if (n[d] > 1 && (best < 0 || stride[d] < stride[best])) best = d;
    return best;
}

fft_plan::fft_plan(kind k, const box3 &inbox, const box3 &outbox, int r2c_dir)
    : kind_(k), inbox_(inbox), outbox_(outbox), r2c_dir_(r2c_dir) {
    if (!is_permutation(inbox.order) || !is_permutation(outbox.order))
        throw std::invalid_argument("box order must be a permutation of (0, 1, 2)");
    if (inbox.empty() || outbox.empty()) throw std::invalid_argument("empty box: a single rank must own the whole world");
    if (k == kind::r2c) {
        if (r2c_dir < 0 || r2c_dir > 2) throw std::invalid_argument("r2c direction must be 0, 1 or 2");
        // box3d::r2c (include/heffte_geometry.h:96-104)
        for (int d = 0; d < 3; d++) {
            long long hi = (d == r2c_dir) ? inbox.low[d] + inbox.size(d) / 2 : inbox.high[d];
            if (outbox.low[d] != inbox.low[d] || outbox.high[d] != hi)
                throw std::invalid_argument("outbox must equal inbox.r2c(r2c_direction) on a single rank");
        }
    } else {
        if (!inbox.same_indexes(outbox))
            throw std::invalid_argument("inbox and outbox must hold the same indexes on a single rank");
        if (k == kind::cos1)
            for (int d = 0; d < 3; d++)
                if (inbox.size(d) < 2) throw std::invalid_argument("DCT-I needs at least 2 points along every axis");
    }
    // workspace: the largest over every variant the plan may run
    long long ws = 0;
    // (a length may be split in one precision only: lines of complex<double> fill LDS sooner)
    for (bool dbl : {false, true})
        for_each_variant([&](int dir, bool ir, bool orl, bool ip) { ws = std::max(ws, build(dir, dbl, ir, orl, ip).ws_elems); });
    ws_elems_ = ws;
}

// Calls f(dir, in_real, out_real, in_place) for every overload of fft3d the plan's kind has.
template <typename F> void fft_plan::for_each_variant(F &&f) const {
    const kind k = kind_;
    for (int dir : {-1, 1})
        for (int ir = 0; ir < 2; ir++)
            for (int orl = 0; orl < 2; orl++)
                for (int ip = 0; ip < 2; ip++) {
                    if (k == kind::c2c && ir && dir == 1) continue;
                    if (k == kind::c2c && orl && dir == -1) continue;
                    if (k == kind::r2c && (ir != (dir == -1) || orl != (dir == 1))) continue;
                    if ((k == kind::cos || k == kind::sin || k == kind::cos1) && (!ir || !orl)) continue;
                    f(dir, ir != 0, orl != 0, ip != 0);
                }
}

void fft_plan::prepare(bool dbl) {
    for_each_variant([&](int dir, bool ir, bool orl, bool ip) {
        const exec_plan &ep = get_exec_plan(dir, dbl, ir, orl, ip);
        for (const pass &p : ep.passes) {
            if (p.split) {
                for (const pass &q : *p.split) (void)get_twiddles(dbl, q, true);
                (void)get_split_table(dbl, p.n, true);
            } else {
                (void)get_twiddles(dbl, p, true);
            }
        }
    });
    if (!dbl && precision_ != kPrecPlain && kind_ == kind::c2c) {
        (void)get_mfma_table(precision_, false, true);
        (void)get_mfma_table(precision_, true, true);
    }
}

fft_plan::~fft_plan() {
    clear_timing();
    for (auto &kv : tw_)
        if (kv.second.dev) (void)hipFree(kv.second.dev);
    for (auto &tp : mfma_tw_)
        for (auto &t : tp)
            if (t.dev) (void)hipFree(t.dev);
}

double fft_plan::scale_factor(int scaling) const {
    // include/heffte_fft3d.h:577 (1 / N of the real-space world) and :545-547
    const double n = (double)inbox_.count();
    double s = 1.0 / n;
    // rtransform of a GPU/stock backend: round trip 4 n per axis, 64 N in 3D
    // (include/heffte_fft3d.h:635-650, the branch every non-FFTW r2r tag takes);
    // DCT-I: round trip 4 (n - 1) per axis, 1 / (64 prod(n - 1)) (:643-646)
    if (kind_ == kind::cos || kind_ == kind::sin) s /= 64.0;
    if (kind_ == kind::cos1)
        s = 1.0 / (64.0 * (double)(inbox_.size(0) - 1) * (double)(inbox_.size(1) - 1) * (double)(inbox_.size(2) - 1));
    if (scaling == 0) return 1.0;
    if (scaling == 2) return std::sqrt(s);
    return s;
}

const exec_plan &fft_plan::get_exec_plan(int dir, bool dbl, bool in_real, bool out_real, bool in_place) {
    const int key = ((dir > 0) << 4) | (dbl << 3) | (in_real << 2) | (out_real << 1) | (int)in_place;
    std::lock_guard<std::mutex> lock(mtx_);
    auto it = plans_.find(key);
    if (it == plans_.end()) it = plans_.emplace(key, build(dir, dbl, in_real, out_real, in_place)).first;
    return it->second;
}

exec_plan fft_plan::build(int dir, bool dbl, bool in_real, bool out_real, bool in_place) const {
    const bool fwd = dir < 0;
    const box3 &bin = fwd ? inbox_ : outbox_;
    const box3 &bout = fwd ? outbox_ : inbox_;
    long long nin[3], nout[3];
    for (int d = 0; d < 3; d++) { nin[d] = bin.size(d); nout[d] = bout.size(d); }
    const layout Lin = layout::ordered(nin, bin.order);
    const layout Lout = layout::ordered(nout, bout.order);
    const bool r2c = (kind_ == kind::r2c);
    // full (real-space) extents; the half-spectrum extents differ along r2c_dir
    long long nfull[3];
    for (int d = 0; d < 3; d++) nfull[d] = inbox_.size(d);

    // axes to transform
    const bool r2r = (kind_ == kind::cos || kind_ == kind::sin || kind_ == kind::cos1);
    std::vector<int> axes;
    if (r2c && fwd) axes.push_back(r2c_dir_);
    for (int d = 0; d < 3; d++)
        // a length-1 cosine/sine transform still multiplies by 2, so r2r keeps every axis
        if ((nfull[d] > 1 || r2r) && !(r2c && d == r2c_dir_)) axes.push_back(d);
    if (r2c && !fwd) axes.push_back(r2c_dir_);
    if (axes.empty()) axes.push_back(0);

    // enumerate axis orders x intermediate layouts x intermediate buffers
    std::vector<int> perm = axes;
    std::sort(perm.begin(), perm.end());
    double best_cost = std::numeric_limits<double>::infinity();
    exec_plan best;

    auto try_plan = [&](const std::vector<candidate_pass> &seq) {
        const int k = (int)seq.size();
        int nlay = 1, nbuf = 1;
        for (int i = 0; i + 1 < k; i++) { nlay *= 6; nbuf *= 2; }
        for (int li = 0; li < nlay; li++) {
            for (int bi = 0; bi < nbuf; bi++) {
                exec_plan ep;
                double cost = 0.0;
                bool ok = true;
                bool uses_ws = false;
                long long dims[3] = {nin[0], nin[1], nin[2]};
                layout cur = Lin;
                buf cur_buf = buf::in;
                int lcode = li, bcode = bi;
                for (int i = 0; i < k && ok; i++) {
                    const candidate_pass &cp = seq[i];
                    pass p;
                    p.axis = cp.axis;
                    p.last = (i + 1 == k);
                    long long odims[3] = {dims[0], dims[1], dims[2]};
                    if (r2c && cp.axis == r2c_dir_ && !cp.identity) {
                        if (fwd) odims[r2c_dir_] = nfull[r2c_dir_] / 2 + 1;
                        else odims[r2c_dir_] = nfull[r2c_dir_];
                    }
                    p.n = cp.identity ? (int)dims[cp.axis] : (int)nfull[cp.axis];
                    if (r2c && cp.axis == r2c_dir_ && !cp.identity) {
                        p.real_axis = true;
                        if (fwd) p.store_len = (int)odims[r2c_dir_];
                        else p.herm_load = 1;
                    }
                    if (!p.store_len) p.store_len = p.n;
                    if (kind_ == kind::cos1 && !cp.identity) {
                        // the kernel runs the 2(N-1)-point even extension and stores N values
                        p.store_len = p.n;
                        p.n = 2 * (p.n - 1);
                    }
                    p.identity = cp.identity;
                    p.lin = cur;
                    p.src = cur_buf;
                    if (p.last) {
                        p.lout = Lout;
                        p.dst = buf::out;
                    } else {
                        p.lout = layout::ordered(odims, kOrders[lcode % 6]);
                        lcode /= 6;
                        p.dst = (bcode % 2) ? buf::ws : buf::out;
                        bcode /= 2;
                    }
                    p.in_real = (r2r || (i == 0 && in_real)) ? 1 : 0;
                    p.out_real = (r2r || (p.last && out_real)) ? 1 : 0;
                    if (r2r && !cp.identity) {
                        if (kind_ == kind::cos1) p.mode = fwd ? kModeDCT1 : kModeDCT1B;
                        else p.mode = (kind_ == kind::cos) ? (fwd ? kModeDCT2 : kModeDCT3) : (fwd ? kModeDST2 : kModeDST3);
                    }
                    // an intermediate in OUT must fit the output buffer as complex data
                    if (!p.last && p.dst == buf::out) {
                        const long long cnt = odims[0] * odims[1] * odims[2];
                        if ((out_real && !r2r) || cnt > bout.count()) { ok = false; break; }
                    }
                    if (p.dst == buf::ws) uses_ws = true;
                    // coalescing: pick the column axis
                    const int fin = p.lin.fastest(), fout = p.lout.fastest();
                    p.load_colfast = (fin >= 0 && fin != p.axis);
                    p.store_colfast = (fout >= 0 && fout != p.axis);
                    if (p.load_colfast && p.store_colfast) {
                        if (fin != fout) { ok = false; break; }
                        p.col = fin;
                    } else if (p.load_colfast) {
                        p.col = fin;
                    } else if (p.store_colfast) {
                        p.col = fout;
                    } else {
                        int c = -1;
                        for (int d = 0; d < 3; d++) {
                            if (d == p.axis) continue;
                            if (c < 0 || (dims[d] > 1 && (dims[c] <= 1 || p.lin.stride[d] < p.lin.stride[c]))) c = d;
                        }
                        p.col = c;
                    }
                    p.b1 = 3 - p.axis - p.col;
                    // in-place safety: same buffer => identical layout and element type
                    const bool same_mem = (p.src == p.dst) ||
                                          (in_place && ((p.src == buf::in && p.dst == buf::out) ||
                                                        (p.src == buf::out && p.dst == buf::in)));
                    if (same_mem) {
                        // r2r passes read and write reals: same element type on both sides
                        bool preserve = (r2r || (!p.in_real && !p.out_real)) && !p.herm_load &&
                                        (p.store_len == p.n || p.mode == kModeDCT1 || p.mode == kModeDCT1B);
                        for (int d = 0; d < 3 && preserve; d++)
                            if (dims[d] != odims[d] || (dims[d] > 1 && p.lin.stride[d] != p.lout.stride[d]))
                                preserve = false;
                        if (!preserve) { ok = false; break; }
                    }
                    cost += 1.0;
                    if (p.load_colfast && p.store_colfast) cost += 0.15;
                    else if (p.load_colfast || p.store_colfast) cost += 0.05;
                    if (!cp.identity && (p.load_colfast || p.store_colfast)) cost += narrow_chunk_cost(p, dbl);
                    // a copy pass runs one LDS tile per line: not along an axis too long for one
                    if (cp.identity) cost += 0.5 + (2 * dims[cp.axis] * (dbl ? 16 : 8) > kGenericMaxLds ? 100.0 : 0.0);
                    ep.passes.push_back(p);
                    cur = p.lout;
                    cur_buf = p.dst;
                    for (int d = 0; d < 3; d++) dims[d] = odims[d];
                }
                if (!ok) continue;
                if (uses_ws) cost += 0.01;
                if (cost < best_cost) {
                    long long ws = 0;
                    for (const pass &p : ep.passes)
                        if (p.dst == buf::ws) ws = std::max(ws, p.lout.n[0] * p.lout.n[1] * p.lout.n[2]);
                    ep.ws_elems = ws;
                    best_cost = cost;
                    best = ep;
                }
            }
        }
    };

    do {
        if (r2c && fwd && perm[0] != r2c_dir_) continue;
        if (r2c && !fwd && perm.back() != r2c_dir_) continue;
        std::vector<candidate_pass> seq;
        for (int a : perm) seq.push_back({a, false});
        try_plan(seq);
        if (!std::isfinite(best_cost)) {
            // no valid in-place schedule with one pass per axis: append a copy pass
            for (int d = 0; d < 3; d++) {
                std::vector<candidate_pass> s2 = seq;
                s2.push_back({d, true});
                try_plan(s2);
            }
        }
    } while (std::next_permutation(perm.begin(), perm.end()));
    if (!std::isfinite(best_cost)) throw std::runtime_error("no valid pass schedule for these boxes");

    for (pass &p : best.passes) configure(p, dbl);
    // the step buffer of split axes follows the intermediate arrays in the workspace
    long long split_elems = 0;
    for (pass &p : best.passes)
        if (p.split) {
            p.split_ws = best.ws_elems;
            split_elems = std::max(split_elems, p.lin.n[0] * p.lin.n[1] * p.lin.n[2]);
        }
    best.ws_elems += split_elems;
    return best;
}

void fft_plan::configure(pass &p, bool dbl) const {
    const size_t elem = dbl ? 16 : 8;
    const bool wide = p.load_colfast || p.store_colfast;
    if (p.mode >= kModeDCT2) {
        // r2r axis: an n-point complex FFT between the Makhoul pack and unpack
        p.fast = find_fast_kernel(p.n, dbl, wide);
        if (p.fast) {
            p.lines = p.fast->T;
            return;
        }
    } else {
        p.mode = kModeC2C;
    }
    if (p.mode == kModeC2C && p.real_axis && p.n % 2 == 0) {
        // real line of n values through a complex FFT of n / 2 (pack / unpack in the same kernel)
        p.fast = find_fast_kernel(p.n / 2, dbl, wide);
        if (p.fast) {
            p.mode = p.herm_load ? kModeC2R : kModeR2C;
            p.lines = p.fast->T;
            return;
        }
    }
    p.fast = p.identity ? nullptr : find_fast_kernel(p.n, dbl, wide);
    if (p.fast) {
        p.lines = p.fast->T;
        return;
    }
    generic_plan &g = p.gen;
    g = generic_plan{};
    g.n = p.n;
    std::vector<int> rad = p.identity ? std::vector<int>{} : factorize(p.n);
    if ((int)rad.size() > kMaxStages) throw std::runtime_error("too many radix stages");
    g.nstages = (int)rad.size();
    int off = 0, ls = 1;
    for (int s = 0; s < g.nstages; s++) {
        g.radix[s] = rad[s];
        g.twoff[s] = off;
        off += ls * (rad[s] - 1);
        ls *= rad[s];
    }
    for (int s = 0; s < g.nstages; s++) {
        g.rootoff[s] = off;
        if (!has_butterfly(rad[s])) off += rad[s];
    }
    g.postoff = off;
    const size_t per_line = 2 * (size_t)p.n * elem;
    if (per_line > (size_t)kGenericMaxLds) {
        // beyond one LDS tile: two passes of lengths n1 * n2 = n (plain c2c lines only)
        int n1 = 0, n2 = 0;
        if (p.mode != kModeC2C || p.identity || p.store_len != p.n || p.herm_load ||
            !split_length(p.n, (int)(kGenericMaxLds / (2 * elem)), dbl, n1, n2))
            throw std::runtime_error("transform length " + std::to_string(p.n) +
                                     " exceeds the single-pass LDS capacity and has no two-pass split");
        auto steps = std::make_shared<std::array<pass, 2>>();
        for (int i = 0; i < 2; i++) {
            pass &q = (*steps)[i];
            q.n = q.store_len = (i == 0) ? n1 : n2;
            q.mode = kModeC2C;
            q.load_colfast = (i == 0) ? 1 : 0;
            q.store_colfast = 1;
            const long long big = 1LL << 20;  // many columns: no cap on the lines per block
            q.lin = q.lout = layout{{big, big, big}, {1, big, big * big}};
            q.col = 0;
            configure(q, dbl);
        }
        p.split_n1 = n1;
        p.split_n2 = n2;
        p.split = steps;
        p.fast = nullptr;
        p.lines = 1;
        return;
    }
    int lines = (int)std::min<size_t>(64, std::max<size_t>(1, (64 * 1024) / per_line));
    // no more lines per block than columns
    const long long ncol = p.lin.n[p.col] > 0 ? std::max(p.lin.n[p.col], p.lout.n[p.col]) : 1;
    while (lines > 1 && lines / 2 >= ncol) lines /= 2;
    g.lines = lines;
    p.lines = lines;
}

namespace {

// exp(-2 pi i e / m) rounded from long double
template <typename R> void root(long long e, long long m, R &re, R &im) {
    const long double pi = 3.141592653589793238462643383279502884L;
    e %= m;
    const long double a = 2.0L * pi * (long double)e / (long double)m;
    re = (R)cosl(a);
    im = (R)(-sinl(a));
}

template <typename R>
std::vector<R> make_table(const int *radix, int nstages, bool with_roots, int tail_count, long long tail_mod) {
    std::vector<R> t;
    long long ls = 1;
    for (int s = 0; s < nstages; s++) {
        const int r = radix[s];
        for (long long k = 0; k < ls; k++)
            for (int p = 1; p < r; p++) {
                R re, im;
                root<R>((long long)p * k, ls * r, re, im);
                t.push_back(re);
                t.push_back(im);
            }
        ls *= r;
    }
    if (with_roots)
        for (int s = 0; s < nstages; s++) {
            if (has_butterfly(radix[s])) continue;
            for (int j = 0; j < radix[s]; j++) {
                R re, im;
                root<R>(j, radix[s], re, im);
                t.push_back(re);
                t.push_back(im);
            }
        }
    // w_{tail_mod}^k, k < tail_count: the pack / unpack twiddles of the real-data modes
    for (int k = 0; k < tail_count; k++) {
        R re, im;
        root<R>(k, tail_mod, re, im);
        t.push_back(re);
        t.push_back(im);
    }
    if (t.empty()) { t.push_back(R(1)); t.push_back(R(0)); }
    return t;
}

}  // namespace

std::vector<int> default_radices(int n) { return factorize(n); }

generic_plan make_generic_plan(int n, int lines) { return make_generic_plan(n, lines, factorize(n)); }

generic_plan make_generic_plan(int n, int lines, const std::vector<int> &rad) {
    generic_plan g{};
    g.n = n;
    long long prod = 1;
    for (int r : rad) prod *= r;
    if (prod != n || (rad.empty() && n != 1)) throw std::invalid_argument("radices do not multiply to the length");
    if ((int)rad.size() > kMaxStages) throw std::runtime_error("too many radix stages");
    g.nstages = (int)rad.size();
    int off = 0, ls = 1;
    for (int s = 0; s < g.nstages; s++) {
        g.radix[s] = rad[s];
        g.twoff[s] = off;
        off += ls * (rad[s] - 1);
        ls *= rad[s];
    }
    for (int s = 0; s < g.nstages; s++) {
        g.rootoff[s] = off;
        if (!has_butterfly(rad[s])) off += rad[s];
    }
    g.postoff = off;
    g.lines = lines;
    return g;
}

std::vector<float> generic_table_f32(const generic_plan &g) {
    return make_table<float>(g.radix, g.nstages, true, 0, 1);
}
std::vector<double> generic_table_f64(const generic_plan &g) {
    return make_table<double>(g.radix, g.nstages, true, 0, 1);
}
std::vector<float> stage_table_f32(const int *radix, int nstages) {
    return make_table<float>(radix, nstages, false, 0, 1);
}
std::vector<double> stage_table_f64(const int *radix, int nstages) {
    return make_table<double>(radix, nstages, false, 0, 1);
}

namespace {
// device copy of a host table (freed again if the copy fails)
void *upload_bytes(const void *src, size_t bytes) {
    void *d = nullptr;
    check_hip(hipMalloc(&d, bytes), "hipMalloc(twiddles)");
    const hipError_t e = hipMemcpy(d, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) (void)hipFree(d);
    check_hip(e, "hipMemcpy(twiddles)");
    return d;
}

const char *kNotResident =
    "device tables are not resident: a call that may not allocate (stream capture, or the no-allocation entry "
    "point) needs heffte_mi355_plan_prepare on this plan and precision first";
}

namespace {
// Name of the kernel a pass launches, in the form the profilers print the template.
std::string pass_kernel_name(const pass &p, bool dbl, bool mfma, int precision, const pass_args &a) {
    const char *R = dbl ? "double" : "float";
    std::string s;
    if (p.split) {
        s = std::string("split ") + R + " " + std::to_string(p.split_n1) + " x " + std::to_string(p.split_n2) +
            " (two launches per transform)";
    } else if (mfma) {
        s = std::string("fft512_mfma<") + (precision == kSplitF16 ? "kSplitF16" : "kSplitBF16") + ", " +
            (a.load_colfast ? "true" : "false") + ", " + (a.store_colfast ? "true" : "false") + ">";
    } else if (p.fast) {
        s = std::string("fft_uniform<") + R + ", radices<";
        for (int i = 0; i < p.fast->nstages; i++) s += (i ? ", " : "") + std::to_string(p.fast->radix[i]);
        s += ">, " + std::to_string(p.fast->E) + ", " + std::to_string(p.fast->T) + ", " +
             std::to_string(p.fast->PS) + ", " + std::to_string(a.stream) + ">";
    } else {
        s = std::string("fft_generic<") + R + "> n=" + std::to_string(p.n);
    }
    return s;
}
}  // namespace

const twiddles &fft_plan::get_twiddles(bool dbl, const pass &p, bool may_alloc) {
    const int n = p.n;
    const long long key = ((long long)n << 8) | (dbl ? 1 : 0) | (p.fast ? 2 : 0) | (p.identity ? 4 : 0) |
                          ((long long)p.mode << 3) | (p.fast ? (long long)p.fast->T << 40 : 0);
    std::lock_guard<std::mutex> lock(mtx_);
    auto it = tw_.find(key);
    if (it != tw_.end()) return it->second;
    if (!may_alloc) throw std::runtime_error(kNotResident);
    const int *radix = p.fast ? p.fast->radix : p.gen.radix;
    const int ns = p.fast ? p.fast->nstages : p.gen.nstages;
    twiddles t;
    // r2c / c2r: w_n^k for k < n/2 (n = real line length); r2r: w_{4n}^k for k < n
    int tail_count = 0;
    long long tail_mod = 1;
    if (p.mode == kModeR2C || p.mode == kModeC2R) { tail_count = n / 2; tail_mod = n; }
    if (p.mode >= kModeDCT2 && p.mode <= kModeDST3) { tail_count = n; tail_mod = 4LL * n; }
    if (dbl) {
        std::vector<double> v = make_table<double>(radix, ns, !p.fast, tail_count, tail_mod);
        t.bytes = v.size() * sizeof(double);
        t.dev = upload_bytes(v.data(), t.bytes);
    } else {
        std::vector<float> v = make_table<float>(radix, ns, !p.fast, tail_count, tail_mod);
        t.bytes = v.size() * sizeof(float);
        t.dev = upload_bytes(v.data(), t.bytes);
    }
    return tw_.emplace(key, t).first->second;
}

const twiddles &fft_plan::get_mfma_table(int prec, bool conj, bool may_alloc) {
    // W512 and the split W32 / W16 operand tables of one precision form and direction
    std::lock_guard<std::mutex> lock(mtx_);
    twiddles &mt = mfma_tw_[prec == kSplitF16 ? 0 : 1][conj ? 1 : 0];
    if (!mt.dev) {
        if (!may_alloc) throw std::runtime_error(kNotResident);
        std::vector<unsigned char> t = mfma512_operand_table(prec, conj);
        mt.dev = upload_bytes(t.data(), t.size());
        mt.bytes = t.size();
    }
    return mt;
}

const twiddles &fft_plan::get_split_table(bool dbl, int n, bool may_alloc) {
    // w_n^m, m < n, of a split axis (key: bit 7 set, out of the range of the pass keys' mode bits)
    const long long key = ((long long)n << 8) | (dbl ? 1 : 0) | 0x80;
    std::lock_guard<std::mutex> lock(mtx_);
    auto it = tw_.find(key);
    if (it != tw_.end()) return it->second;
    if (!may_alloc) throw std::runtime_error(kNotResident);
    twiddles t;
    if (dbl) {
        std::vector<double> v = make_table<double>(nullptr, 0, false, n, n);
        t.bytes = v.size() * sizeof(double);
        t.dev = upload_bytes(v.data(), t.bytes);
    } else {
        std::vector<float> v = make_table<float>(nullptr, 0, false, n, n);
        t.bytes = v.size() * sizeof(float);
        t.dev = upload_bytes(v.data(), t.bytes);
    }
    return tw_.emplace(key, t).first->second;
}

pass_args pass_launch_args(kind k, const exec_plan &ep, const pass &p, int dir, bool dbl, int batch, const void *in,
                           void *out, void *ws, const void *tw, double scale, const void *mul) {
    pass_args a{};
    auto bufptr = [&](buf b) -> void * {
        return b == buf::in ? const_cast<void *>(in) : (b == buf::out ? out : ws);
    };
    a.in = bufptr(p.src);
    a.out = bufptr(p.dst);
    a.tw = tw;
    a.in_sa = p.lin.stride[p.axis];
    a.in_sc = p.lin.stride[p.col];
    a.in_sb1 = p.lin.stride[p.b1];
    a.in_sb2 = p.lin.n[0] * p.lin.n[1] * p.lin.n[2];
    a.out_sa = p.lout.stride[p.axis];
    a.out_sc = p.lout.stride[p.col];
    a.out_sb1 = p.lout.stride[p.b1];
    a.out_sb2 = p.lout.n[0] * p.lout.n[1] * p.lout.n[2];
    a.ncol = (int)p.lout.n[p.col];
    a.nb1 = (int)p.lout.n[p.b1];
    a.nbatch = batch;
    if ((long long)a.ncol * a.nb1 * batch > 0x7fffffffLL) throw std::runtime_error("too many lines in one pass");
    a.nlines = a.ncol * a.nb1 * batch;
    a.load_colfast = p.load_colfast;
    a.store_colfast = p.store_colfast;
    // r2r modes run their inverse through explicit conjugation inside the kernel
    a.conj = (dir > 0 && p.mode < kModeDCT2) ? 1 : 0;
    a.in_real = p.in_real;
    a.out_real = p.out_real;
    a.store_len = p.store_len;
    a.herm_load = p.herm_load;
    a.mode = p.mode;
    // the output of a pass that fits the 256 MB Infinity Cache can serve the next pass from
    // there; larger ones stream past the caches: non-temporal stores, and non-temporal loads
    // when a side walks the column axis (measured on the 512^3 complex<float> passes: the
    // contiguous pass 365 -> 353 us with plain loads, the column-chunk pass 388 -> 413 us)
    const long long out_bytes = (long long)batch * p.lout.n[0] * p.lout.n[1] * p.lout.n[2] *
                                (p.out_real ? (dbl ? 8 : 4) : (dbl ? 16 : 8));
    a.stream = out_bytes <= kStreamBytes ? 0
               : (p.load_colfast || p.store_colfast) ? kStreamLoadsStores : kStreamStores;
    a.scale = p.last ? scale : 1.0;
    a.mul = nullptr;
    if (mul && &p == &ep.passes.front()) {
        // the fused product needs the plain complex load of the first pass
        if (k != kind::c2c || p.in_real || p.herm_load || p.mode != kModeC2C || p.src != buf::in)
            throw std::runtime_error("internal error: the product cannot be fused into this pass");
        a.mul = mul;
    }
    return a;
}

void split_steps(const pass &p, const pass_args &base, bool dbl, int bt, void *tmp, pass_args step[2]) {
    const int ax = p.axis, n1 = p.split_n1, n2 = p.split_n2;
    const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
    const size_t elem = dbl ? 16 : 8;
    const long long cnt_in = p.lin.n[0] * p.lin.n[1] * p.lin.n[2];
    const long long cnt_out = p.lout.n[0] * p.lout.n[1] * p.lout.n[2];
    // the line-index dimensions of a step: extent, stride on the load side, stride on the store side
    struct dim {
        long long n, si, so;
        int id;  // 0: the interleaved index (j2 of step 1, k1 of step 2), 1 and 2: the other axes
    };
    auto order = [](std::array<dim, 3> d, bool by_load) {
        // the fastest dimension (on the side that walks the column axis) is the column
        std::stable_sort(d.begin(), d.end(), [&](const dim &x, const dim &y) {
            const long long sx = x.n > 1 ? (by_load ? x.si : x.so) : (1LL << 62);
            const long long sy = y.n > 1 ? (by_load ? y.si : y.so) : (1LL << 62);
            return sx < sy;
        });
        return d;
    };
    const long long sa_in = p.lin.stride[ax], sa_out = p.lout.stride[ax];
    const char *src = static_cast<const char *>(base.in) + (size_t)bt * cnt_in * (p.in_real ? elem / 2 : elem);
    char *buf = static_cast<char *>(tmp) + (size_t)bt * cnt_in * elem;
    char *dst = static_cast<char *>(base.out) + (size_t)bt * cnt_out * (p.out_real ? elem / 2 : elem);
    auto fill = [](pass_args &a, const std::array<dim, 3> &d) {
        a.in_sc = d[0].si;
        a.out_sc = d[0].so;
        a.in_sb1 = d[1].si;
        a.out_sb1 = d[1].so;
        a.in_sb2 = d[2].si;
        a.out_sb2 = d[2].so;
        a.ncol = (int)d[0].n;
        a.nb1 = (int)d[1].n;
        a.nbatch = (int)d[2].n;
        a.nlines = (int)(d[0].n * d[1].n * d[2].n);
    };
    // step 1: n1-point lines of stride n2 sa_in, one per (j2, o1, o2), into buf (layout lin)
    {
        const std::array<dim, 3> d = order({dim{n2, sa_in, sa_in, 0}, dim{p.lin.n[o1], p.lin.stride[o1], p.lin.stride[o1], 1},
                                            dim{p.lin.n[o2], p.lin.stride[o2], p.lin.stride[o2], 2}},
                                           true);
        pass_args &a = step[0];
        a = base;
        a.in = src;
        a.out = buf;
        a.mul = base.mul ? static_cast<const char *>(base.mul) + (size_t)bt * cnt_in * elem : nullptr;
        a.in_sa = a.out_sa = (long long)n2 * sa_in;
        fill(a, d);
        a.load_colfast = a.store_colfast = 1;
        a.out_real = 0;
        a.store_len = n1;
        a.mode = kModeC2C;
        a.scale = 1.0;
        a.tw4_dim = d[0].id == 0 ? 0 : (d[1].id == 0 ? 1 : 2);
    }
    // step 2: n2-point lines of buf (stride sa_in), one per (k1, o1, o2), to k1 + n1 k2 of dst
    {
        const bool lcf = p.lin.fastest() != ax;
        const std::array<dim, 3> d = order({dim{n1, (long long)n2 * sa_in, sa_out, 0},
                                            dim{p.lin.n[o1], p.lin.stride[o1], p.lout.stride[o1], 1},
                                            dim{p.lin.n[o2], p.lin.stride[o2], p.lout.stride[o2], 2}},
                                           lcf);
        pass_args &a = step[1];
        a = base;
        a.in = buf;
        a.out = dst;
        a.mul = nullptr;
        a.in_real = 0;
        a.in_sa = sa_in;
        a.out_sa = (long long)n1 * sa_out;
        fill(a, d);
        a.load_colfast = lcf ? 1 : 0;
        a.store_colfast = 1;
        a.store_len = n2;
        a.mode = kModeC2C;
        a.tw4 = nullptr;
        a.tw4_dim = 0;
    }
}

void fft_plan::launch_split(const pass &p, const pass_args &base, bool dbl, int batch, void *tmp) {
    const size_t elem = dbl ? 16 : 8;
    for (int bt = 0; bt < batch; bt++) {
        pass_args st[2];
        split_steps(p, base, dbl, bt, tmp, st);
        st[0].tw4 = get_split_table(dbl, p.n, false).dev;
        for (int i = 0; i < 2; i++) {
            const pass &q = (*p.split)[i];
            st[i].tw = get_twiddles(dbl, q, false).dev;
            const long long nblocks = ((long long)st[i].nlines + q.lines - 1) / q.lines;
            if (q.fast) check_hip(q.fast->launch(st[i], nblocks, stream_), "split FFT pass");
            else check_hip(launch_generic(dbl, st[i], q.gen, nblocks, 2 * (size_t)q.gen.lines * q.n * elem, stream_),
                           "split FFT pass");
        }
    }
}

void fft_plan::execute(int dir, bool dbl, bool in_real, bool out_real, int batch, const void *in, void *out, void *ws,
                       int scaling, const void *mul, bool no_alloc) {
    if (batch < 1) throw std::invalid_argument("batch size must be positive");
    if (!in || !out) throw std::invalid_argument("null input or output array");
    const bool in_place = (in == out);
    const exec_plan &ep = get_exec_plan(dir, dbl, in_real, out_real, in_place);
    if (no_alloc) {
        // the workspace is the static scratch of this device (one per GPU, shared by every call)
        if (ws) throw std::invalid_argument("the no-allocation call takes no workspace (it uses the static scratch)");
        const unsigned long long need = (unsigned long long)ep.ws_elems * (unsigned long long)batch * (dbl ? 16 : 8);
        if (need > kScratchBytes)
            throw std::invalid_argument("the workspace of this call (" + std::to_string(need) +
                                        " bytes) exceeds the static scratch of the no-allocation entry point (" +
                                        std::to_string(kScratchBytes) + " bytes)");
        if (need > 0) ws = scratch_address();
    }
    // the device tables must be resident before anything is launched when the call may not
    // allocate: look them all up first (a missing one throws here, with nothing enqueued)
    bool may_alloc = !no_alloc;
    if (may_alloc) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream_, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) may_alloc = false;
    }
    const bool split = precision_ != kPrecPlain && !dbl;
    constexpr size_t kMaxPasses = 8;
    if (ep.passes.size() > kMaxPasses) throw std::runtime_error("internal error: too many passes");
    std::array<const void *, kMaxPasses> tables{};
    for (size_t i = 0; i < ep.passes.size(); i++) {
        const pass &p = ep.passes[i];
        const bool mfma = split && p.n == 512 && p.mode == kModeC2C && !p.identity && !p.in_real && !p.out_real &&
                          !p.herm_load && p.store_len == p.n && !(mul && i == 0);
        if (p.split) {
            // both steps' tables and the step twiddles, looked up now (launch_split uses them)
            for (const pass &q : *p.split) (void)get_twiddles(dbl, q, may_alloc);
            tables[i] = get_split_table(dbl, p.n, may_alloc).dev;
        } else {
            tables[i] = mfma ? get_mfma_table(precision_, dir > 0, may_alloc).dev : get_twiddles(dbl, p, may_alloc).dev;
        }
    }
    void *tmp = nullptr;
    if (ep.ws_elems > 0 && !ws) {
        // fft3d::forward(in, out) allocates the workspace per call (include/heffte_fft3d.h:353-360)
        check_hip(hipMallocAsync(&tmp, (size_t)ep.ws_elems * batch * (dbl ? 16 : 8), stream_), "hipMallocAsync(workspace)");
        ws = tmp;
    }
    const double scale = scale_factor(scaling);
    // capacity of each buffer (elements of its own type) under the fft3d contract
    const bool fwd = dir < 0;
    const long long cap_in = batch * (fwd ? inbox_.count() : outbox_.count());
    const long long cap_out = batch * (fwd ? outbox_.count() : inbox_.count());
    const long long cap_ws = batch * ep.ws_elems;
    auto capacity = [&](buf b) { return b == buf::in ? cap_in : (b == buf::out ? cap_out : cap_ws); };
    auto extent = [&](const layout &L) {
        long long e = 0;
        for (int d = 0; d < 3; d++) e += (L.n[d] - 1) * L.stride[d];
        return e + (long long)(batch - 1) * (L.n[0] * L.n[1] * L.n[2]) + 1;
    };
    const bool debug = std::getenv("HMI_DEBUG_PLAN") != nullptr;
    // per-pass timing: one event before the first pass and one after each pass
    std::unique_lock<std::mutex> tlock(mtx_, std::defer_lock);
    hipEvent_t tprev = nullptr;
    if (timing_) {
        tlock.lock();
        check_hip(hipEventCreate(&tprev), "hipEventCreate");
        timing_events_.push_back(tprev);
        check_hip(hipEventRecord(tprev, stream_), "hipEventRecord");
    }
    for (const pass &p : ep.passes) {
        // every element a pass touches lies inside the buffer it names
        if (extent(p.lin) > capacity(p.src) || extent(p.lout) > capacity(p.dst))
            throw std::runtime_error("internal error: pass exceeds its buffer");
        if (p.split && batch * (p.split_ws + p.lin.n[0] * p.lin.n[1] * p.lin.n[2]) > cap_ws)
            throw std::runtime_error("internal error: split pass exceeds the workspace");
        if (debug)
            std::fprintf(stderr,
                         "[hmi] pass axis=%d n=%d src=%d dst=%d col=%d b1=%d lcf=%d scf=%d real=%d/%d herm=%d "
                         "store_len=%d fast=%d lines=%d in_stride=(%lld,%lld,%lld) out_stride=(%lld,%lld,%lld)\n",
                         p.axis, p.n, (int)p.src, (int)p.dst, p.col, p.b1, p.load_colfast, p.store_colfast,
                         p.in_real, p.out_real, p.herm_load, p.store_len, p.fast ? 1 : 0, p.lines, p.lin.stride[0],
                         p.lin.stride[1], p.lin.stride[2], p.lout.stride[0], p.lout.stride[1], p.lout.stride[2]);
        pass_args a = pass_launch_args(kind_, ep, p, dir, dbl, batch, in, out, ws, tables[&p - &ep.passes.front()],
                                       scale, mul);
        const long long nblocks = ((long long)a.nlines + p.lines - 1) / p.lines;
        const bool mfma = split && p.n == 512 && p.mode == kModeC2C && !p.identity && !p.in_real && !p.out_real &&
                          !p.herm_load && p.store_len == p.n && !a.mul;
        if (p.split) {
            launch_split(p, a, dbl, batch, static_cast<char *>(ws) + (size_t)batch * p.split_ws * (dbl ? 16 : 8));
        } else if (mfma) {
            // the split-precision form: a.tw holds W512 and the split W32 / W16 operand tables
            // of the direction, 16-line tiles; the backward pass multiplies by the conjugate
            // matrices directly (no conjugation of the data on load and store)
            a.conj = 0;
            check_hip(launch_mfma512(precision_, a, ((long long)a.nlines + 15) / 16, stream_), "MFMA FFT pass");
        } else if (p.fast) {
            check_hip(p.fast->launch(a, nblocks, stream_), "fast FFT pass");
        } else {
            const size_t lds = 2 * (size_t)p.gen.lines * p.n * (dbl ? 16 : 8);
            check_hip(launch_generic(dbl, a, p.gen, nblocks, lds, stream_), "generic FFT pass");
        }
        if (tprev) {
            const std::string name = pass_kernel_name(p, dbl, mfma, precision_, a);
            int id = (int)(std::find(timing_names_.begin(), timing_names_.end(), name) - timing_names_.begin());
            if (id == (int)timing_names_.size()) timing_names_.push_back(name);
            hipEvent_t e = nullptr;
            check_hip(hipEventCreate(&e), "hipEventCreate");
            timing_events_.push_back(e);
            check_hip(hipEventRecord(e, stream_), "hipEventRecord");
            timing_launches_.push_back({id, {tprev, e}});
            tprev = e;
        }
    }
    if (tmp) check_hip(hipFreeAsync(tmp, stream_), "hipFreeAsync(workspace)");
}

int fft_plan::matrix_core_passes(int dir, bool in_place, int *total) {
    const bool r2c = kind_ == kind::r2c, r2r = kind_ != kind::c2c && !r2c;
    const bool in_real = r2r || (r2c && dir < 0), out_real = r2r || (r2c && dir > 0);
    const exec_plan &ep = get_exec_plan(dir, false, in_real, out_real, in_place && !r2c);
    int n = 0;
    for (const pass &p : ep.passes)
        // the same condition as the launch in execute (a convolution's fused product aside)
        n += (precision_ != kPrecPlain && p.n == 512 && p.mode == kModeC2C && !p.identity && !p.in_real &&
              !p.out_real && !p.herm_load && p.store_len == p.n && !p.split)
                 ? 1
                 : 0;
    if (total) *total = (int)ep.passes.size();
    return n;
}

void fft_plan::clear_timing() {
    for (hipEvent_t e : timing_events_) (void)hipEventDestroy(e);
    timing_events_.clear();
    timing_launches_.clear();
}

void fft_plan::set_timing(bool on) {
    std::lock_guard<std::mutex> lock(mtx_);
    clear_timing();
    timing_ = on;
}

std::string fft_plan::timing_report() {
    std::lock_guard<std::mutex> lock(mtx_);
    std::vector<long long> count(timing_names_.size(), 0);
    std::vector<double> total(timing_names_.size(), 0.0);
    for (const auto &rec : timing_launches_) {
        check_hip(hipEventSynchronize(rec.second.second), "hipEventSynchronize");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, rec.second.first, rec.second.second), "hipEventElapsedTime");
        count[rec.first] += 1;
        total[rec.first] += ms;
    }
    std::string out;
    char line[64];
    for (size_t i = 0; i < timing_names_.size(); i++) {
        if (!count[i]) continue;
        std::snprintf(line, sizeof(line), "\t%lld\t%.6f\n", count[i], total[i]);
        out += timing_names_[i] + line;
    }
    return out;
}


void fft_plan::convolve(bool dbl, int batch, const void *x, const void *y, void *out, void *ws, int scaling) {
    if (kind_ != kind::c2c) throw std::invalid_argument("convolution needs a c2c plan");
    if (batch < 1) throw std::invalid_argument("batch size must be positive");
    if (!x || !y || !out) throw std::invalid_argument("null input or output array");
    const size_t elem = dbl ? 16 : 8;
    void *tmp = nullptr;
    if (!ws) {
        check_hip(hipMallocAsync(&tmp, (size_t)convolve_workspace() * batch * elem, stream_), "hipMallocAsync(workspace)");
        ws = tmp;
    }
    // layout of ws: the transforms' own workspace, then forward(y) in the outbox layout
    void *yhat = static_cast<char *>(ws) + (size_t)ws_elems_ * batch * elem;
    try {
        execute(-1, dbl, false, false, batch, y, yhat, ws, 0);
        if (x == y) {
            // one spectrum: the first backward pass squares it on load
            execute(+1, dbl, false, false, batch, yhat, out, ws, scaling, yhat);
        } else {
            execute(-1, dbl, false, false, batch, x, out, ws, 0);
            execute(+1, dbl, false, false, batch, out, out, ws, scaling, yhat);
        }
    } catch (...) {
        if (tmp) (void)hipFreeAsync(tmp, stream_);
        throw;
    }
    if (tmp) check_hip(hipFreeAsync(tmp, stream_), "hipFreeAsync(workspace)");
}

void fft_plan::convolve_spectrum(bool dbl, int batch, const void *x, const void *yhat, void *out, void *ws,
                                 int scaling) {
    if (kind_ != kind::c2c) throw std::invalid_argument("convolution needs a c2c plan");
    if (batch < 1) throw std::invalid_argument("batch size must be positive");
    if (!x || !yhat || !out) throw std::invalid_argument("null input or output array");
    if (out == yhat) throw std::invalid_argument("the output must not alias the spectrum");
    execute(-1, dbl, false, false, batch, x, out, ws, 0);
    execute(+1, dbl, false, false, batch, out, out, ws, scaling, yhat);
}

void permute(int elem_bytes, const box3 &inbox, const int out_order[3], int batch, const void *in, void *out,
             hipStream_t s) {
    if (!is_permutation(inbox.order) || !is_permutation(out_order))
        throw std::invalid_argument("box order must be a permutation of (0, 1, 2)");
    if (elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16)
        throw std::invalid_argument("element size must be 4, 8 or 16 bytes");
    if (batch < 1) throw std::invalid_argument("batch size must be positive");
    if (inbox.empty()) return;
    if (!in || !out) throw std::invalid_argument("null input or output array");
    if (in == out) throw std::invalid_argument("the permutation kernel is out-of-place");
    long long n[3];
    for (int d = 0; d < 3; d++) n[d] = inbox.size(d);
    const layout Li = layout::ordered(n, inbox.order), Lo = layout::ordered(n, out_order);
    int a = inbox.order[0], b = out_order[0];
    permute_args p{};
    p.same_fast = (a == b);
    if (a == b) b = inbox.order[1];
    const int c = 3 - a - b;
    p.n_a = n[a]; p.n_b = n[b]; p.n_c = n[c]; p.nbatch = batch;
    p.si_a = Li.stride[a]; p.si_b = Li.stride[b]; p.si_c = Li.stride[c]; p.bs_in = inbox.count();
    p.so_a = Lo.stride[a]; p.so_b = Lo.stride[b]; p.so_c = Lo.stride[c]; p.bs_out = inbox.count();
    p.ntile_a = (n[a] + 31) / 32;
    p.ntile_b = (n[b] + 31) / 32;
    if (p.ntile_a * p.ntile_b * p.n_c * p.nbatch > 0x7fffffffLL) throw std::runtime_error("grid too large");
    check_hip(launch_permute(elem_bytes, in, out, p, s), "permute kernel");
}

}  // namespace hmi
