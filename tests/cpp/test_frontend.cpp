// C++ front end test (include/heffte_mi355.hpp), in the style of heFFTe's
// test/test_fft3d.h:397-572 (test_fft3d_arrays / vectors / batches) and
// test/test_cos.cpp (r2r known answers).  Checks each result against a
// brute-force long-double DFT of the same box, the reference's known answers,
// or the input (round trips).  Exit status 0 = every check passed.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../include/heffte_mi355.hpp"

static int failures = 0;
#define CHECK(cond, what)                                                     \
    do {                                                                      \
        if (!(cond)) { std::printf("FAIL %s (line %d)\n", what, __LINE__); failures++; } \
        else std::printf("pass %s\n", what);                                  \
    } while (0)

template <typename T> double err_vs(std::vector<T> const &a, std::vector<std::complex<long double>> const &b) {
    long double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        std::complex<long double> x((long double)std::real(std::complex<double>(a[i])), (long double)std::imag(std::complex<double>(a[i])));
        num += std::norm(x - b[i]);
        den += std::norm(b[i]);
    }
    return (double)std::sqrt(num / (den > 0 ? den : 1));
}

// brute-force 3D DFT over a box stored in `in` order, output in `out` order
template <typename T>
std::vector<std::complex<long double>> dft3(heffte::box3d<> const &in, heffte::box3d<> const &out,
                                            std::vector<T> const &x, int sign) {
    const long double pi = 3.141592653589793238462643383279502884L;
    auto stride = [](heffte::box3d<> const &b) {
        std::array<long long, 3> s;
        s[b.order[0]] = 1;
        s[b.order[1]] = b.size[b.order[0]];
        s[b.order[2]] = (long long)b.size[b.order[0]] * b.size[b.order[1]];
        return s;
    };
    auto si = stride(in), so = stride(out);
    std::vector<std::complex<long double>> y((size_t)out.count());
    for (int k2 = 0; k2 < in.size[2]; k2++)
        for (int k1 = 0; k1 < in.size[1]; k1++)
            for (int k0 = 0; k0 < in.size[0]; k0++) {
                std::complex<long double> acc = 0;
                for (int j2 = 0; j2 < in.size[2]; j2++)
                    for (int j1 = 0; j1 < in.size[1]; j1++)
                        for (int j0 = 0; j0 < in.size[0]; j0++) {
                            long double ph = 2 * pi * sign *
                                ((long double)((long long)j0 * k0 % in.size[0]) / in.size[0] +
                                 (long double)((long long)j1 * k1 % in.size[1]) / in.size[1] +
                                 (long double)((long long)j2 * k2 % in.size[2]) / in.size[2]);
                            std::complex<double> v(x[(size_t)(j0 * si[0] + j1 * si[1] + j2 * si[2])]);
                            acc += std::complex<long double>(v.real(), v.imag()) * std::polar(1.0L, ph);
                        }
                y[(size_t)(k0 * so[0] + k1 * so[1] + k2 * so[2])] = acc;
            }
    return y;
}

template <typename R> void test_c2c(double tol) {
    using C = std::complex<R>;
    const char *pn = sizeof(R) == 4 ? "float" : "double";
    heffte::box3d<> inbox({0, 0, 0}, {5, 6, 7}, {0, 1, 2});
    heffte::box3d<> outbox({0, 0, 0}, {5, 6, 7}, {2, 0, 1});
    heffte::fft3d<heffte::backend::mi355> fft(inbox, outbox);
    std::mt19937 gen(4242);
    std::uniform_real_distribution<R> u(0, 1);
    std::vector<C> x((size_t)fft.size_inbox());
    for (auto &v : x) v = C(u(gen), u(gen));
    auto ref = dft3(inbox, outbox, x, -1);
    char name[128];

    // vector API
    auto gx = heffte::gpu::transfer::load(x);
    auto gy = fft.forward(gx);
    std::snprintf(name, 128, "%s c2c vector forward", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(gy), ref) < tol, name);
    auto gz = fft.backward(gy, heffte::scale::full);
    std::vector<std::complex<long double>> xl(x.begin(), x.end());
    std::snprintf(name, 128, "%s c2c vector round trip", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(gz), xl) < tol, name);

    // arrays with a caller workspace, in place
    heffte::gpu::vector<C> ws(std::max<size_t>(1, fft.size_workspace()));
    heffte::gpu::vector<C> buf = heffte::gpu::transfer::load(x);
    fft.forward(buf.data(), buf.data(), ws.data());
    std::snprintf(name, 128, "%s c2c in-place forward", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(buf), ref) < tol, name);
    fft.backward(buf.data(), buf.data(), ws.data(), heffte::scale::full);
    std::snprintf(name, 128, "%s c2c in-place round trip", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(buf), xl) < tol, name);

    // batch of 3 with symmetric scaling
    const int batch = 3;
    std::vector<C> xb(batch * x.size());
    for (auto &v : xb) v = C(u(gen), u(gen));
    heffte::gpu::vector<C> gb = heffte::gpu::transfer::load(xb), gob((size_t)batch * fft.size_outbox());
    heffte::gpu::vector<C> wsb(std::max<size_t>(1, batch * fft.size_workspace()));
    fft.forward(batch, gb.data(), gob.data(), wsb.data(), heffte::scale::symmetric);
    auto hb = heffte::gpu::transfer::unload(gob);
    double worst = 0;
    for (int b = 0; b < batch; b++) {
        std::vector<C> xi(xb.begin() + b * x.size(), xb.begin() + (b + 1) * x.size());
        auto r = dft3(inbox, outbox, xi, -1);
        const long double s = 1.0L / std::sqrt((long double)x.size());
        for (auto &v : r) v *= s;
        std::vector<C> yi(hb.begin() + b * x.size(), hb.begin() + (b + 1) * x.size());
        worst = std::max(worst, err_vs(yi, r));
    }
    std::snprintf(name, 128, "%s c2c batch symmetric", pn);
    CHECK(worst < tol, name);

    // real input (full spectrum) and real output (truncation)
    std::vector<R> xr((size_t)fft.size_inbox());
    for (auto &v : xr) v = u(gen);
    auto gr = heffte::gpu::transfer::load(xr);
    auto gc = fft.forward(gr);
    std::snprintf(name, 128, "%s real-input forward", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(gc), dft3(inbox, outbox, xr, -1)) < tol, name);
    auto gback = fft.backward_real(gc, heffte::scale::full);
    std::vector<std::complex<long double>> xrl(xr.begin(), xr.end());
    std::snprintf(name, 128, "%s complex-to-real backward", pn);
    CHECK(err_vs(heffte::gpu::transfer::unload(gback), xrl) < tol, name);

    // undersized vector throws std::invalid_argument (include/heffte_fft3d.h:442-443)
    bool thrown = false;
    try {
        heffte::gpu::vector<C> small(10);
        (void)fft.forward(small);
    } catch (std::invalid_argument const &) { thrown = true; }
    std::snprintf(name, 128, "%s undersized input throws", pn);
    CHECK(thrown, name);
}

template <typename R> void test_r2c(double tol) {
    using C = std::complex<R>;
    const char *pn = sizeof(R) == 4 ? "float" : "double";
    for (int dir = 0; dir < 3; dir++) {
        heffte::box3d<> inbox({0, 0, 0}, {7, 9, 5}, {1, 2, 0});
        heffte::box3d<> outbox = inbox.r2c(dir);
        heffte::fft3d_r2c<heffte::backend::mi355> fft(inbox, outbox, dir);
        std::mt19937 gen(17 + dir);
        std::uniform_real_distribution<R> u(-1, 1);
        std::vector<R> x((size_t)fft.size_inbox());
        for (auto &v : x) v = u(gen);
        auto full = dft3(inbox, inbox, x, -1);
        // keep the half spectrum in outbox order
        std::vector<std::complex<long double>> ref((size_t)outbox.count());
        auto so = [&](int a, int b, int c) {
            std::array<long long, 3> s;
            s[outbox.order[0]] = 1;
            s[outbox.order[1]] = outbox.size[outbox.order[0]];
            s[outbox.order[2]] = (long long)outbox.size[outbox.order[0]] * outbox.size[outbox.order[1]];
            return a * s[0] + b * s[1] + c * s[2];
        };
        auto si = [&](int a, int b, int c) {
            std::array<long long, 3> s;
            s[inbox.order[0]] = 1;
            s[inbox.order[1]] = inbox.size[inbox.order[0]];
            s[inbox.order[2]] = (long long)inbox.size[inbox.order[0]] * inbox.size[inbox.order[1]];
            return a * s[0] + b * s[1] + c * s[2];
        };
        for (int c = 0; c < outbox.size[2]; c++)
            for (int b = 0; b < outbox.size[1]; b++)
                for (int a = 0; a < outbox.size[0]; a++) ref[(size_t)so(a, b, c)] = full[(size_t)si(a, b, c)];
        auto gx = heffte::gpu::transfer::load(x);
        auto gy = fft.forward(gx);
        char name[128];
        std::snprintf(name, 128, "%s r2c dir %d forward", pn, dir);
        CHECK(err_vs(heffte::gpu::transfer::unload(gy), ref) < tol, name);
        auto gz = fft.backward(gy, heffte::scale::full);
        std::vector<std::complex<long double>> xl(x.begin(), x.end());
        std::snprintf(name, 128, "%s r2c dir %d round trip", pn, dir);
        CHECK(err_vs(heffte::gpu::transfer::unload(gz), xl) < tol, name);
        (void)sizeof(C);
    }
}

template <typename R> void test_r2r(double tol) {
    const char *pn = sizeof(R) == 4 ? "float" : "double";
    heffte::box3d<> world({0, 0, 0}, {1, 2, 3});
    std::vector<R> x(24);
    for (int i = 0; i < 24; i++) x[i] = (R)(i + 1);
    // test/test_cos.cpp:40-43
    std::vector<double> cosref = {2.4e+03, -6.7882250993908571e+01, -2.2170250336881628e+02, 0.0, 0.0, 0.0,
                                  -9.0844474461089760e+02, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                                  -6.4561180200187039e+01, 0.0, 0.0, 0.0, 0.0, 0.0};
    heffte::rtransform<heffte::backend::mi355_cos> tcos(world, world);
    auto gx = heffte::gpu::transfer::load(x);
    auto gy = tcos.forward(gx);
    auto hy = heffte::gpu::transfer::unload(gy);
    double worst = 0;
    for (int i = 0; i < 24; i++) worst = std::max(worst, std::abs((double)hy[i] - cosref[i]));
    char name[128];
    std::snprintf(name, 128, "%s rtransform cos known answer", pn);
    CHECK(worst < (sizeof(R) == 4 ? 1e-2 : 1e-9), name);
    auto gz = tcos.backward(gy, heffte::scale::full);
    auto hz = heffte::gpu::transfer::unload(gz);
    worst = 0;
    for (int i = 0; i < 24; i++) worst = std::max(worst, std::abs((double)hz[i] - (double)x[i]));
    std::snprintf(name, 128, "%s rtransform cos round trip", pn);
    CHECK(worst < (sizeof(R) == 4 ? 1e-3 : 1e-11), name);
    heffte::rtransform<heffte::backend::mi355_sin> tsin(world, world);
    auto gs = tsin.forward(gx);
    auto gsb = tsin.backward(gs, heffte::scale::full);
    auto hs = heffte::gpu::transfer::unload(gsb);
    worst = 0;
    for (int i = 0; i < 24; i++) worst = std::max(worst, std::abs((double)hs[i] - (double)x[i]));
    std::snprintf(name, 128, "%s rtransform sin round trip", pn);
    CHECK(worst < (sizeof(R) == 4 ? 1e-3 : 1e-11), name);
    // DCT-I, test/test_cos.cpp:48-50; scale::full = 1 / (64 (n0-1)(n1-1)(n2-1))
    std::vector<double> cos1ref(24, 0.0);
    cos1ref[0] = 600.0; cos1ref[1] = -24.0; cos1ref[2] = -48.0; cos1ref[6] = -192.0; cos1ref[18] = -48.0;
    heffte::rtransform<heffte::backend::mi355_cos1> tcos1(world, world);
    auto g1 = tcos1.forward(gx);
    auto h1 = heffte::gpu::transfer::unload(g1);
    worst = 0;
    for (int i = 0; i < 24; i++) worst = std::max(worst, std::abs((double)h1[i] - cos1ref[i]));
    std::snprintf(name, 128, "%s rtransform cos1 known answer", pn);
    CHECK(worst < (sizeof(R) == 4 ? 1e-3 : 1e-10), name);
    CHECK(std::abs(tcos1.get_scale_factor(heffte::scale::full) - 1.0 / 384.0) < 1e-15, "cos1 scale factor");
    auto h1b = heffte::gpu::transfer::unload(tcos1.backward(g1, heffte::scale::full));
    worst = 0;
    for (int i = 0; i < 24; i++) worst = std::max(worst, std::abs((double)h1b[i] - (double)x[i]));
    std::snprintf(name, 128, "%s rtransform cos1 round trip", pn);
    CHECK(worst < (sizeof(R) == 4 ? 1e-3 : 1e-11), name);
    (void)tol;
}

int main() {
    test_c2c<float>(5e-6);
    test_c2c<double>(1e-13);
    test_r2c<float>(5e-6);
    test_r2c<double>(1e-13);
    test_r2r<float>(5e-6);
    test_r2r<double>(1e-13);
    // plan errors are std::invalid_argument
    bool thrown = false;
    try {
        heffte::fft3d<heffte::backend::mi355> bad(heffte::box3d<>({0, 0, 0}, {3, 3, 3}), heffte::box3d<>({0, 0, 0}, {3, 3, 2}));
    } catch (std::invalid_argument const &) { thrown = true; }
    CHECK(thrown, "mismatched boxes throw");
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
