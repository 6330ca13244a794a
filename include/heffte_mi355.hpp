/*
 * heffte_mi355.hpp -- C++ front end of the MI355X single-rank 3D FFT backend.
 *
 * It reproduces, for one rank, the interface callers of heFFTe use
 * (include/heffte_fft3d.h, include/heffte_fft3d_r2c.h, include/heffte_geometry.h):
 *
 *     heffte::box3d<> inbox({0,0,0}, {511,511,511}, {0,1,2});
 *     heffte::box3d<> outbox({0,0,0}, {511,511,511}, {1,2,0});
 *     heffte::fft3d<heffte::backend::mi355> fft(inbox, outbox, MPI_COMM_SELF);
 *     fft.forward(input, output, workspace, heffte::scale::symmetric);
 *
 * so that existing code switches by changing the backend tag.  Every transform
 * runs in the library's HIP kernels through the C ABI of heffte_mi355.h; this
 * header holds no arithmetic.  Errors are thrown as std::runtime_error (device
 * or internal failures) and std::invalid_argument (bad boxes, sizes, types),
 * the exceptions heFFTe throws (include/heffte_backend_rocm.h:51-62,
 * include/heffte_fft3d.h:442-443).
 *
 * This header is the standalone form of the backend (it defines the heffte
 * namespace types itself).  To plug the library into an existing heFFTe
 * build instead, see INTEGRATION.md.
 */
#ifndef HEFFTE_MI355_HPP
#define HEFFTE_MI355_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "heffte_mi355.h"

#if !defined(HEFFTE_MI355_NO_MPI) && defined(__has_include)
#if __has_include(<mpi.h>)
#include <mpi.h>
#define HEFFTE_MI355_HAVE_MPI 1
#endif
#endif

namespace heffte {

//! heffte::scale (include/heffte_fft3d.h:140-147)
enum class scale { none, full, symmetric };

//! heffte::reshape_algorithm (include/heffte_plan_logic.h); no effect on one rank.
enum class reshape_algorithm { alltoallv = 0, alltoall = 3, p2p_plined = 1, p2p = 2 };

//! heffte::plan_options (include/heffte_plan_logic.h:131-176); accepted, no effect on one rank.
struct plan_options {
    plan_options(bool reorder = false, reshape_algorithm alg = reshape_algorithm::alltoallv, bool pencils = true)
        : use_reorder(reorder), algorithm(alg), use_pencils(pencils), use_gpu_aware(true) {}
    bool use_reorder;
    reshape_algorithm algorithm;
    bool use_pencils;
    bool use_gpu_aware;
};

/*!
 * \brief A box of indexes, inclusive low and high, with an order of the fast, mid and slow
 *        dimensions (the semantics of heffte::box3d, include/heffte_geometry.h:67-133).
 */
template <typename index = int> struct box3d {
    box3d(std::array<index, 3> clow, std::array<index, 3> chigh, std::array<int, 3> corder = {0, 1, 2})
        : low(clow), high(chigh),
          size({chigh[0] - clow[0] + 1, chigh[1] - clow[1] + 1, chigh[2] - clow[2] + 1}), order(corder) {}
    box3d(std::array<index, 2> clow, std::array<index, 2> chigh)
        : box3d(std::array<index, 3>{clow[0], clow[1], 0}, std::array<index, 3>{chigh[0], chigh[1], 0}) {}
    bool empty() const { return size[0] <= 0 || size[1] <= 0 || size[2] <= 0; }
    long long count() const {
        return empty() ? 0 : (long long)size[0] * (long long)size[1] * (long long)size[2];
    }
    //! The box of the half spectrum of an r2c transform along dimension (size / 2 + 1 entries).
    box3d r2c(int dimension) const {
        std::array<index, 3> h = high;
        h[dimension] = low[dimension] + size[dimension] / 2;
        return box3d(low, h, order);
    }
    bool is2d() const { return size[0] == 1 || size[1] == 1 || size[2] == 1; }
    bool operator==(box3d const &o) const { return low == o.low && high == o.high; }
    bool operator!=(box3d const &o) const { return !(*this == o); }
    bool ordered_same_as(box3d const &o) const { return order == o.order; }
    int find_order(int dir) const { return dir == order[0] ? 0 : (dir == order[1] ? 1 : 2); }
    index osize(int dimension) const { return size[order[dimension]]; }
    std::array<index, 3> low, high, size;
    std::array<int, 3> order;
};
template <typename index = int> using box2d = box3d<index>;

namespace backend {
//! Complex transforms (fft3d, fft3d_r2c) on an MI355X.
struct mi355 {};
//! rtransform: DCT-II forward / DCT-III backward (the convention of backend::rocfft_cos).
struct mi355_cos {};
//! rtransform: DST-II forward / DST-III backward (the convention of backend::rocfft_sin).
struct mi355_sin {};
//! rtransform: DCT-I forward and backward (the convention of backend::rocfft_cos1).
struct mi355_cos1 {};

template <typename tag> struct is_enabled : std::false_type {};
template <> struct is_enabled<mi355> : std::true_type {};
template <> struct is_enabled<mi355_cos> : std::true_type {};
template <> struct is_enabled<mi355_sin> : std::true_type {};
template <> struct is_enabled<mi355_cos1> : std::true_type {};

template <typename tag> inline std::string name() { return "unknown"; }
template <> inline std::string name<mi355>() { return "mi355"; }
template <> inline std::string name<mi355_cos>() { return "mi355-cos"; }
template <> inline std::string name<mi355_sin>() { return "mi355-sin"; }
template <> inline std::string name<mi355_cos1>() { return "mi355-cos1"; }
}  // namespace backend

//! default_options<tag>: the fused executor absorbs the layout permutation, no reorder pass.
template <typename tag> inline plan_options default_options() { return plan_options(false); }

namespace detail {
inline void check(int rc) {
    if (rc == HMI_SUCCESS) return;
    const std::string msg = heffte_mi355_last_error();
    if (rc == HMI_ERROR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error("heffte mi355 error " + std::to_string(rc) + ": " + msg);
}
inline void hip_check(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}
template <typename index> void unpack(box3d<index> const &b, long long low[3], long long high[3], int order[3]) {
    for (int d = 0; d < 3; d++) {
        low[d] = b.low[d];
        high[d] = b.high[d];
        order[d] = b.order[d];
    }
}
struct plan_deleter {
    void operator()(heffte_mi355_plan_s *p) const { heffte_mi355_plan_destroy(p); }
};
using plan_ptr = std::unique_ptr<heffte_mi355_plan_s, plan_deleter>;
}  // namespace detail

namespace gpu {
/*!
 * \brief RAII device array (the role of heffte::gpu::device_vector,
 *        include/heffte_backend_vector.h:52-157).
 */
template <typename T> class vector {
  public:
    using value_type = T;
    vector() = default;
    explicit vector(size_t n) : n_(n) {
        if (n) detail::hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
    }
    vector(T const *host_begin, T const *host_end) : vector((size_t)(host_end - host_begin)) {
        if (n_) detail::hip_check(hipMemcpy(p_, host_begin, n_ * sizeof(T), hipMemcpyHostToDevice));
    }
    vector(vector const &o) : vector(o.n_) {
        if (n_) detail::hip_check(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice));
    }
    vector(vector &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    vector &operator=(vector o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~vector() {
        if (p_) (void)hipFree(p_);
    }
    T *data() { return p_; }
    T const *data() const { return p_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

//! gpu::transfer::load / unload (include/heffte_backend_data_transfer.h:29-177).
struct transfer {
    template <typename T> static vector<T> load(std::vector<T> const &h) { return vector<T>(h.data(), h.data() + h.size()); }
    template <typename T> static std::vector<T> unload(vector<T> const &d) {
        std::vector<T> h(d.size());
        if (!h.empty()) detail::hip_check(hipMemcpy(h.data(), d.data(), d.size() * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
};
}  // namespace gpu

namespace detail {
// forward / backward dispatch on the element types of heFFTe's compatible-type table
// (include/heffte_fft3d.h:173-187)
inline int fwd(heffte_mi355_plan p, int b, std::complex<float> const *i, std::complex<float> *o, void *w, int s) {
    return heffte_mi355_forward_c2c(p, b, i, o, w, s);
}
inline int fwd(heffte_mi355_plan p, int b, std::complex<double> const *i, std::complex<double> *o, void *w, int s) {
    return heffte_mi355_forward_z2z(p, b, i, o, w, s);
}
inline int fwd(heffte_mi355_plan p, int b, float const *i, std::complex<float> *o, void *w, int s) {
    return heffte_mi355_forward_s2c(p, b, i, o, w, s);
}
inline int fwd(heffte_mi355_plan p, int b, double const *i, std::complex<double> *o, void *w, int s) {
    return heffte_mi355_forward_d2z(p, b, i, o, w, s);
}
inline int fwd(heffte_mi355_plan p, int b, float const *i, float *o, void *w, int s) {
    return heffte_mi355_forward_s2s(p, b, i, o, w, s);
}
inline int fwd(heffte_mi355_plan p, int b, double const *i, double *o, void *w, int s) {
    return heffte_mi355_forward_d2d(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, std::complex<float> const *i, std::complex<float> *o, void *w, int s) {
    return heffte_mi355_backward_c2c(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, std::complex<double> const *i, std::complex<double> *o, void *w, int s) {
    return heffte_mi355_backward_z2z(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, std::complex<float> const *i, float *o, void *w, int s) {
    return heffte_mi355_backward_c2s(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, std::complex<double> const *i, double *o, void *w, int s) {
    return heffte_mi355_backward_z2d(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, float const *i, float *o, void *w, int s) {
    return heffte_mi355_backward_s2s(p, b, i, o, w, s);
}
inline int bwd(heffte_mi355_plan p, int b, double const *i, double *o, void *w, int s) {
    return heffte_mi355_backward_d2d(p, b, i, o, w, s);
}

template <typename T> struct complex_of { using type = std::complex<T>; };
template <typename T> struct complex_of<std::complex<T>> { using type = std::complex<T>; };
template <typename T> struct real_of { using type = T; };
template <typename T> struct real_of<std::complex<T>> { using type = T; };

template <typename tag> struct plan_kind { static constexpr int value = HMI_KIND_C2C; };
template <> struct plan_kind<backend::mi355_cos> { static constexpr int value = HMI_KIND_COS; };
template <> struct plan_kind<backend::mi355_sin> { static constexpr int value = HMI_KIND_SIN; };
template <> struct plan_kind<backend::mi355_cos1> { static constexpr int value = HMI_KIND_COS1; };

inline int to_c(scale s) { return s == scale::none ? HMI_SCALE_NONE : (s == scale::full ? HMI_SCALE_FULL : HMI_SCALE_SYMMETRIC); }
}  // namespace detail

/*!
 * \brief Single-rank 3D FFT (heffte::fft3d, include/heffte_fft3d.h:239-692), with
 *        backend::mi355 (complex) or backend::mi355_cos / mi355_sin / mi355_cos1 (real-to-real).
 */
template <typename backend_tag, typename index = int> class fft3d {
    static_assert(backend::is_enabled<backend_tag>::value, "The requested backend is invalid or has not been enabled.");
    static constexpr bool r2r = detail::plan_kind<backend_tag>::value != HMI_KIND_C2C;

  public:
    template <typename T> using buffer_container = gpu::vector<T>;

    fft3d(box3d<index> const inbox, box3d<index> const outbox, plan_options const = default_options<backend_tag>())
        : fft3d(nullptr, inbox, outbox) {}
    fft3d(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox,
          plan_options const = default_options<backend_tag>())
        : in_(inbox), out_(outbox), stream_(stream) {
        long long il[3], ih[3], ol[3], oh[3];
        int io[3], oo[3];
        detail::unpack(inbox, il, ih, io);
        detail::unpack(outbox, ol, oh, oo);
        heffte_mi355_plan p = nullptr;
        if (r2r)
            detail::check(heffte_mi355_plan_create_r2r(detail::plan_kind<backend_tag>::value, il, ih, io, ol, oh, oo,
                                                       nullptr, &p));
        else
            detail::check(heffte_mi355_plan_create(il, ih, io, ol, oh, oo, nullptr, &p));
        plan_.reset(p);
        detail::check(heffte_mi355_set_stream(p, stream));
        const double n = (double)inbox.count();
        scale_factor_ = r2r ? 1.0 / (64.0 * n) : 1.0 / n;
        if (detail::plan_kind<backend_tag>::value == HMI_KIND_COS1)  // include/heffte_fft3d.h:643-646
            scale_factor_ = 1.0 / (64.0 * (double)(inbox.size[0] - 1) * (double)(inbox.size[1] - 1) *
                                   (double)(inbox.size[2] - 1));
    }
#ifdef HEFFTE_MI355_HAVE_MPI
    //! The reference constructor; the communicator must have exactly one rank.
    fft3d(box3d<index> const inbox, box3d<index> const outbox, MPI_Comm const comm,
          plan_options const options = default_options<backend_tag>())
        : fft3d(nullptr, inbox, outbox, comm, options) {}
    fft3d(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, MPI_Comm const comm,
          plan_options const options = default_options<backend_tag>())
        : fft3d(stream, inbox, outbox, options) {
        int size = 0;
        MPI_Comm_size(comm, &size);
        if (size != 1) throw std::invalid_argument("backend::mi355 runs a single rank (communicator of size 1)");
    }
#endif

    long long size_inbox() const { return in_.count(); }
    long long size_outbox() const { return out_.count(); }
    box3d<index> inbox() const { return in_; }
    box3d<index> outbox() const { return out_; }
    //! Workspace size in complex elements (real elements for real-to-real tags).
    size_t size_workspace() const { return (size_t)heffte_mi355_size_workspace(plan_.get()); }
    double get_scale_factor(scale scaling) const {
        return scaling == scale::symmetric ? std::sqrt(scale_factor_) : scale_factor_;
    }
    hipStream_t stream() const { return stream_; }

    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), 1, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), 1, input, output, nullptr, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), batch_size, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[],
                 scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), batch_size, input, output, nullptr, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), 1, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), 1, input, output, nullptr, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), batch_size, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[],
                  scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), batch_size, input, output, nullptr, detail::to_c(scaling)));
    }

    //! Vector API (include/heffte_fft3d.h:430-446): complex or real input, complex output
    //! (real output for the real-to-real tags).
    template <typename input_type>
    gpu::vector<typename std::conditional<r2r, input_type, typename detail::complex_of<input_type>::type>::type>
    forward(gpu::vector<input_type> const &input, scale scaling = scale::none) {
        using out_t = typename std::conditional<r2r, input_type, typename detail::complex_of<input_type>::type>::type;
        if (input.size() < (size_t)size_inbox())
            throw std::invalid_argument("The input vector is smaller than size_inbox(), i.e., not enough entries provided to fill the inbox.");
        gpu::vector<out_t> out((size_t)size_outbox());
        forward(input.data(), out.data(), scaling);
        return out;
    }
    template <typename scalar_type>
    gpu::vector<scalar_type> backward(gpu::vector<scalar_type> const &input, scale scaling = scale::none) {
        if (input.size() < (size_t)size_outbox())
            throw std::invalid_argument("The input vector is smaller than size_outbox(), i.e., not enough entries provided to fill the outbox.");
        gpu::vector<scalar_type> out((size_t)size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }
    //! Complex-to-real backward (the real part of the result), heffte's backward_real.
    template <typename scalar_type>
    gpu::vector<typename detail::real_of<scalar_type>::type> backward_real(gpu::vector<scalar_type> const &input,
                                                                          scale scaling = scale::none) {
        gpu::vector<typename detail::real_of<scalar_type>::type> out((size_t)size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }

  private:
    box3d<index> in_, out_;
    hipStream_t stream_;
    detail::plan_ptr plan_;
    double scale_factor_ = 1.0;
};

template <typename backend_tag, typename index = int> using fft2d = fft3d<backend_tag, index>;
template <typename backend_tag, typename index = int> using rtransform = fft3d<backend_tag, index>;

/*!
 * \brief Real-to-complex single-rank 3D FFT (heffte::fft3d_r2c, include/heffte_fft3d_r2c.h:45-380).
 */
template <typename backend_tag, typename index = int> class fft3d_r2c {
    static_assert(std::is_same<backend_tag, backend::mi355>::value, "fft3d_r2c needs backend::mi355");

  public:
    fft3d_r2c(box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              plan_options const = default_options<backend_tag>())
        : fft3d_r2c(nullptr, inbox, outbox, r2c_direction) {}
    fft3d_r2c(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              plan_options const = default_options<backend_tag>())
        : in_(inbox), out_(outbox), stream_(stream) {
        long long il[3], ih[3], ol[3], oh[3];
        int io[3], oo[3];
        detail::unpack(inbox, il, ih, io);
        detail::unpack(outbox, ol, oh, oo);
        heffte_mi355_plan p = nullptr;
        detail::check(heffte_mi355_plan_create_r2c(il, ih, io, ol, oh, oo, r2c_direction, nullptr, &p));
        plan_.reset(p);
        detail::check(heffte_mi355_set_stream(p, stream));
        scale_factor_ = 1.0 / (double)inbox.count();
    }
#ifdef HEFFTE_MI355_HAVE_MPI
    fft3d_r2c(box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction, MPI_Comm const comm,
              plan_options const options = default_options<backend_tag>())
        : fft3d_r2c(nullptr, inbox, outbox, r2c_direction, comm, options) {}
    fft3d_r2c(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              MPI_Comm const comm, plan_options const options = default_options<backend_tag>())
        : fft3d_r2c(stream, inbox, outbox, r2c_direction, options) {
        int size = 0;
        MPI_Comm_size(comm, &size);
        if (size != 1) throw std::invalid_argument("backend::mi355 runs a single rank (communicator of size 1)");
    }
#endif
    long long size_inbox() const { return in_.count(); }
    long long size_outbox() const { return out_.count(); }
    box3d<index> inbox() const { return in_; }
    box3d<index> outbox() const { return out_; }
    size_t size_workspace() const { return (size_t)heffte_mi355_size_workspace(plan_.get()); }
    double get_scale_factor(scale scaling) const {
        return scaling == scale::symmetric ? std::sqrt(scale_factor_) : scale_factor_;
    }

    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        static_assert(!std::is_same<input_type, output_type>::value, "fft3d_r2c::forward takes real input");
        detail::check(detail::fwd(plan_.get(), 1, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), 1, input, output, nullptr, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        detail::check(detail::fwd(plan_.get(), batch_size, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), 1, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), 1, input, output, nullptr, detail::to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        detail::check(detail::bwd(plan_.get(), batch_size, input, output, workspace, detail::to_c(scaling)));
    }
    template <typename input_type>
    gpu::vector<std::complex<input_type>> forward(gpu::vector<input_type> const &input, scale scaling = scale::none) {
        if (input.size() < (size_t)size_inbox())
            throw std::invalid_argument("The input vector is smaller than size_inbox(), i.e., not enough entries provided to fill the inbox.");
        gpu::vector<std::complex<input_type>> out((size_t)size_outbox());
        forward(input.data(), out.data(), scaling);
        return out;
    }
    template <typename scalar_type>
    gpu::vector<typename detail::real_of<scalar_type>::type> backward(gpu::vector<scalar_type> const &input,
                                                                     scale scaling = scale::none) {
        if (input.size() < (size_t)size_outbox())
            throw std::invalid_argument("The input vector is smaller than size_outbox(), i.e., not enough entries provided to fill the outbox.");
        gpu::vector<typename detail::real_of<scalar_type>::type> out((size_t)size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }

  private:
    box3d<index> in_, out_;
    hipStream_t stream_;
    detail::plan_ptr plan_;
    double scale_factor_ = 1.0;
};

}  // namespace heffte

#endif
