/*
 * heffte_backend_mi355.h -- the MI355X single-rank 3D FFT library as a heFFTe backend tag.
 *
 * Include it after heffte.h:
 *
 *     #include "heffte.h"
 *     #include "heffte_backend_mi355.h"
 *     heffte::fft3d<heffte::backend::mi355> fft(inbox, outbox, MPI_COMM_SELF);   // was backend::rocfft
 *     fft.forward(input, output, workspace, heffte::scale::symmetric);
 *
 * It adds, inside heFFTe's own namespace and without redefining anything of it:
 *   - the tags backend::mi355 (fft3d, fft3d_r2c) and backend::mi355_cos / mi355_sin / mi355_cos1
 *     (rtransform: DCT-II/III, DST-II/III, DCT-I with the conventions of backend::rocfft_cos /
 *     rocfft_sin / rocfft_cos1), and their traits as heFFTe declares them for every tag
 *     (include/heffte_common.h:225-258,264-396,439-543, include/heffte_plan_logic.h:131-176):
 *     is_enabled, buffer_traits (location tag::gpu, a device-vector container), name,
 *     uses_fft_types (false for the real-to-real tags), default_plan_options (use_reorder =
 *     false: the fused passes absorb the layout permutation);
 *   - partial specializations fft3d<backend::mi355, index> (and of the three real-to-real tags,
 *     reached through heffte::rtransform, include/heffte_fft3d.h:748-749) and
 *     fft3d_r2c<backend::mi355, index> (include/heffte_fft3d_r2c.h:45), so that the tag alone
 *     selects the library's fused passes through the C ABI of heffte_mi355.h -- heFFTe's
 *     reshapes and 1D executors are not involved.
 *
 * The specializations keep the reference's public interface for one rank: the MPI_Comm and
 * stream constructors (the communicator must hold one rank), size_inbox / size_outbox /
 * size_workspace (complex elements; real for the real-to-real tags) / size_comm_buffers (0 on
 * one rank), inbox / outbox, get_scale_factor, stream / synchronize_device, the raw-array
 * forward/backward overloads with and without workspace and batch size, and the container API
 * (forward / backward / backward_real on buffer_container).  Errors are thrown as
 * std::invalid_argument (bad boxes, sizes, types) and std::runtime_error (device failures).
 *
 * Without heffte.h, include heffte_mi355.hpp instead: it declares the parts of heFFTe's
 * interface this header builds on and then includes it (the standalone form).
 */
#ifndef HEFFTE_BACKEND_MI355_H
#define HEFFTE_BACKEND_MI355_H

#if !defined(HEFFTE_FFT3D_R2C_H) && !defined(HEFFTE_MI355_STANDALONE)
#error "include heffte.h (or, without heFFTe, heffte_mi355.hpp) before heffte_backend_mi355.h"
#endif

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "heffte_mi355.h"

#if defined(HEFFTE_FFT3D_R2C_H) && !defined(HEFFTE_MI355_HAVE_MPI)
#define HEFFTE_MI355_HAVE_MPI 1  // heffte.h includes mpi.h
#endif

namespace heffte {

namespace backend {
//! Complex transforms (fft3d, fft3d_r2c) on an MI355X.
struct mi355 {};
//! rtransform: DCT-II forward / DCT-III backward (the convention of backend::rocfft_cos).
struct mi355_cos {};
//! rtransform: DST-II forward / DST-III backward (the convention of backend::rocfft_sin).
struct mi355_sin {};
//! rtransform: DCT-I forward and backward (the convention of backend::rocfft_cos1).
struct mi355_cos1 {};

template <> struct is_enabled<mi355> : std::true_type {};
template <> struct is_enabled<mi355_cos> : std::true_type {};
template <> struct is_enabled<mi355_sin> : std::true_type {};
template <> struct is_enabled<mi355_cos1> : std::true_type {};
}  // namespace backend

namespace mi355_detail {

inline void check(int rc) {
    if (rc == HMI_SUCCESS) return;
    const std::string msg = heffte_mi355_last_error();
    if (rc == HMI_ERROR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error("heffte mi355 error " + std::to_string(rc) + ": " + msg);
}
inline void hip_check(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

/*!
 * \brief RAII device array: the buffer container of the mi355 tags (the role of
 *        heffte::gpu::device_vector, include/heffte_backend_vector.h:52-157).
 */
template <typename T> class device_vector {
  public:
    using value_type = T;
    device_vector() = default;
    explicit device_vector(size_t n) : n_(n) {
        if (n) hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
    }
    device_vector(T const *host_begin, T const *host_end) : device_vector((size_t)(host_end - host_begin)) {
        if (n_) hip_check(hipMemcpy(p_, host_begin, n_ * sizeof(T), hipMemcpyHostToDevice));
    }
    device_vector(device_vector const &o) : device_vector(o.n_) {
        if (n_) hip_check(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice));
    }
    device_vector(device_vector &&o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    device_vector &operator=(device_vector o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~device_vector() {
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

}  // namespace mi355_detail

namespace backend {
template <> struct buffer_traits<mi355> {
    using location = tag::gpu;
    template <typename T> using container = mi355_detail::device_vector<T>;
};
template <> struct buffer_traits<mi355_cos> {
    using location = tag::gpu;
    template <typename T> using container = mi355_detail::device_vector<T>;
};
template <> struct buffer_traits<mi355_sin> {
    using location = tag::gpu;
    template <typename T> using container = mi355_detail::device_vector<T>;
};
template <> struct buffer_traits<mi355_cos1> {
    using location = tag::gpu;
    template <typename T> using container = mi355_detail::device_vector<T>;
};

template <> inline std::string name<mi355>() { return "mi355"; }
template <> inline std::string name<mi355_cos>() { return "mi355-cos"; }
template <> inline std::string name<mi355_sin>() { return "mi355-sin"; }
template <> inline std::string name<mi355_cos1>() { return "mi355-cos1"; }

template <> struct uses_fft_types<mi355_cos> : std::false_type {};
template <> struct uses_fft_types<mi355_sin> : std::false_type {};
template <> struct uses_fft_types<mi355_cos1> : std::false_type {};
}  // namespace backend

//! The fused executor absorbs the layout permutation: no reorder pass is requested.
template <> struct default_plan_options<backend::mi355> { static const bool use_reorder = false; };
template <> struct default_plan_options<backend::mi355_cos> { static const bool use_reorder = false; };
template <> struct default_plan_options<backend::mi355_sin> { static const bool use_reorder = false; };
template <> struct default_plan_options<backend::mi355_cos1> { static const bool use_reorder = false; };

namespace mi355_detail {

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

inline int to_c(scale s) {
    return s == scale::none ? HMI_SCALE_NONE : (s == scale::full ? HMI_SCALE_FULL : HMI_SCALE_SYMMETRIC);
}

#ifdef HEFFTE_MI355_HAVE_MPI
inline void require_one_rank(MPI_Comm comm) {
    int size = 0;
    MPI_Comm_size(comm, &size);
    if (size != 1) throw std::invalid_argument("backend::mi355 runs a single rank (communicator of size 1)");
}
#endif

//! Common part of the specializations: the plan, the boxes, the stream and the scale factor.
template <typename index> class plan_holder {
  public:
    long long size_inbox() const { return in_.count(); }
    long long size_outbox() const { return out_.count(); }
    box3d<index> inbox() const { return in_; }
    box3d<index> outbox() const { return out_; }
    //! Workspace size in complex elements (real elements for the real-to-real tags).
    size_t size_workspace() const { return (size_t)heffte_mi355_size_workspace(plan_.get()); }
    //! One rank exchanges nothing: no communication buffers.
    size_t size_comm_buffers() const { return 0; }
    double get_scale_factor(scale scaling) const {
        return scaling == scale::symmetric ? std::sqrt(scale_factor_) : scale_factor_;
    }
    hipStream_t stream() const { return stream_; }
    void synchronize_device() const { hip_check(hipStreamSynchronize(stream_)); }

  protected:
    plan_holder(hipStream_t stream, box3d<index> const &inbox, box3d<index> const &outbox)
        : in_(inbox), out_(outbox), stream_(stream) {}
    void adopt(heffte_mi355_plan p, double scale_factor) {
        plan_.reset(p);
        check(heffte_mi355_set_stream(p, stream_));
        scale_factor_ = scale_factor;
    }
    heffte_mi355_plan plan() const { return plan_.get(); }

  private:
    box3d<index> in_, out_;
    hipStream_t stream_;
    plan_ptr plan_;
    double scale_factor_ = 1.0;
};

/*!
 * \brief fft3d for the mi355 tags (include/heffte_fft3d.h:239-692): backend::mi355 (complex
 *        transforms) or backend::mi355_cos / mi355_sin / mi355_cos1 (rtransform).
 */
template <typename backend_tag, typename index> class fft3d_impl : public plan_holder<index> {
    static constexpr bool r2r = plan_kind<backend_tag>::value != HMI_KIND_C2C;
    using base = plan_holder<index>;

  public:
    template <typename T> using buffer_container = device_vector<T>;
    using location_tag = tag::gpu;

    fft3d_impl(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox) : base(stream, inbox, outbox) {
        long long il[3], ih[3], ol[3], oh[3];
        int io[3], oo[3];
        unpack(inbox, il, ih, io);
        unpack(outbox, ol, oh, oo);
        heffte_mi355_plan p = nullptr;
        if (r2r)
            check(heffte_mi355_plan_create_r2r(plan_kind<backend_tag>::value, il, ih, io, ol, oh, oo, nullptr, &p));
        else
            check(heffte_mi355_plan_create(il, ih, io, ol, oh, oo, nullptr, &p));
        const double n = (double)inbox.count();
        double s = r2r ? 1.0 / (64.0 * n) : 1.0 / n;
        if (plan_kind<backend_tag>::value == HMI_KIND_COS1)  // include/heffte_fft3d.h:643-646
            s = 1.0 / (64.0 * (double)(inbox.size[0] - 1) * (double)(inbox.size[1] - 1) * (double)(inbox.size[2] - 1));
        this->adopt(p, s);
    }

    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        check(fwd(this->plan(), 1, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        check(fwd(this->plan(), 1, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        check(fwd(this->plan(), batch_size, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[],
                 scale scaling = scale::none) const {
        check(fwd(this->plan(), batch_size, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), 1, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        check(bwd(this->plan(), 1, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), batch_size, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), batch_size, input, output, nullptr, to_c(scaling)));
    }

    //! Container API (include/heffte_fft3d.h:430-446): complex or real input, complex output
    //! (real output for the real-to-real tags).
    template <typename input_type>
    device_vector<typename std::conditional<r2r, input_type, typename complex_of<input_type>::type>::type>
    forward(device_vector<input_type> const &input, scale scaling = scale::none) {
        using out_t = typename std::conditional<r2r, input_type, typename complex_of<input_type>::type>::type;
        if (input.size() < (size_t)this->size_inbox())
            throw std::invalid_argument("The input vector is smaller than size_inbox(), i.e., not enough entries provided to fill the inbox.");
        device_vector<out_t> out((size_t)this->size_outbox());
        forward(input.data(), out.data(), scaling);
        return out;
    }
    template <typename scalar_type>
    device_vector<scalar_type> backward(device_vector<scalar_type> const &input, scale scaling = scale::none) {
        if (input.size() < (size_t)this->size_outbox())
            throw std::invalid_argument("The input vector is smaller than size_outbox(), i.e., not enough entries provided to fill the outbox.");
        device_vector<scalar_type> out((size_t)this->size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }
    //! Complex-to-real backward (the real part of the result), heffte's backward_real.
    template <typename scalar_type>
    device_vector<typename real_of<scalar_type>::type> backward_real(device_vector<scalar_type> const &input,
                                                                    scale scaling = scale::none) {
        device_vector<typename real_of<scalar_type>::type> out((size_t)this->size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }
};

/*!
 * \brief fft3d_r2c for backend::mi355 (include/heffte_fft3d_r2c.h:45-380).
 */
template <typename index> class fft3d_r2c_impl : public plan_holder<index> {
    using base = plan_holder<index>;

  public:
    template <typename T> using buffer_container = device_vector<T>;
    using location_tag = tag::gpu;

    fft3d_r2c_impl(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction)
        : base(stream, inbox, outbox) {
        long long il[3], ih[3], ol[3], oh[3];
        int io[3], oo[3];
        unpack(inbox, il, ih, io);
        unpack(outbox, ol, oh, oo);
        heffte_mi355_plan p = nullptr;
        check(heffte_mi355_plan_create_r2c(il, ih, io, ol, oh, oo, r2c_direction, nullptr, &p));
        this->adopt(p, 1.0 / (double)inbox.count());
    }

    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        static_assert(!std::is_same<input_type, output_type>::value, "fft3d_r2c::forward takes real input");
        check(fwd(this->plan(), 1, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        check(fwd(this->plan(), 1, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[], output_type workspace[],
                 scale scaling = scale::none) const {
        check(fwd(this->plan(), batch_size, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void forward(int const batch_size, input_type const input[], output_type output[],
                 scale scaling = scale::none) const {
        check(fwd(this->plan(), batch_size, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), 1, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(input_type const input[], output_type output[], scale scaling = scale::none) const {
        check(bwd(this->plan(), 1, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[], input_type workspace[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), batch_size, input, output, workspace, to_c(scaling)));
    }
    template <typename input_type, typename output_type>
    void backward(int const batch_size, input_type const input[], output_type output[],
                  scale scaling = scale::none) const {
        check(bwd(this->plan(), batch_size, input, output, nullptr, to_c(scaling)));
    }
    template <typename input_type>
    device_vector<std::complex<input_type>> forward(device_vector<input_type> const &input,
                                                    scale scaling = scale::none) {
        if (input.size() < (size_t)this->size_inbox())
            throw std::invalid_argument("The input vector is smaller than size_inbox(), i.e., not enough entries provided to fill the inbox.");
        device_vector<std::complex<input_type>> out((size_t)this->size_outbox());
        forward(input.data(), out.data(), scaling);
        return out;
    }
    template <typename scalar_type>
    device_vector<typename real_of<scalar_type>::type> backward(device_vector<scalar_type> const &input,
                                                               scale scaling = scale::none) {
        if (input.size() < (size_t)this->size_outbox())
            throw std::invalid_argument("The input vector is smaller than size_outbox(), i.e., not enough entries provided to fill the outbox.");
        device_vector<typename real_of<scalar_type>::type> out((size_t)this->size_inbox());
        backward(input.data(), out.data(), scaling);
        return out;
    }
};

}  // namespace mi355_detail

// The tag selects the library: one specialization per tag, with the reference's constructors.
#ifdef HEFFTE_MI355_HAVE_MPI
#define HEFFTE_MI355_FFT3D_SPECIALIZATION(TAG)                                                                     \
    template <typename index> class fft3d<backend::TAG, index> : public mi355_detail::fft3d_impl<backend::TAG, index> { \
        using impl = mi355_detail::fft3d_impl<backend::TAG, index>;                                                \
                                                                                                                   \
      public:                                                                                                      \
        fft3d(box3d<index> const inbox, box3d<index> const outbox,                                                 \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(nullptr, inbox, outbox) {}                                                                      \
        fft3d(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox,                             \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(stream, inbox, outbox) {}                                                                       \
        fft3d(box3d<index> const inbox, box3d<index> const outbox, MPI_Comm const comm,                            \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(nullptr, inbox, outbox) {                                                                       \
            mi355_detail::require_one_rank(comm);                                                                  \
        }                                                                                                          \
        fft3d(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, MPI_Comm const comm,        \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(stream, inbox, outbox) {                                                                        \
            mi355_detail::require_one_rank(comm);                                                                  \
        }                                                                                                          \
    };
#else
#define HEFFTE_MI355_FFT3D_SPECIALIZATION(TAG)                                                                     \
    template <typename index> class fft3d<backend::TAG, index> : public mi355_detail::fft3d_impl<backend::TAG, index> { \
        using impl = mi355_detail::fft3d_impl<backend::TAG, index>;                                                \
                                                                                                                   \
      public:                                                                                                      \
        fft3d(box3d<index> const inbox, box3d<index> const outbox,                                                 \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(nullptr, inbox, outbox) {}                                                                      \
        fft3d(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox,                             \
              plan_options const = default_options<backend::TAG>())                                                \
            : impl(stream, inbox, outbox) {}                                                                       \
    };
#endif

HEFFTE_MI355_FFT3D_SPECIALIZATION(mi355)
HEFFTE_MI355_FFT3D_SPECIALIZATION(mi355_cos)
HEFFTE_MI355_FFT3D_SPECIALIZATION(mi355_sin)
HEFFTE_MI355_FFT3D_SPECIALIZATION(mi355_cos1)
#undef HEFFTE_MI355_FFT3D_SPECIALIZATION

template <typename index> class fft3d_r2c<backend::mi355, index> : public mi355_detail::fft3d_r2c_impl<index> {
    using impl = mi355_detail::fft3d_r2c_impl<index>;

  public:
    fft3d_r2c(box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              plan_options const = default_options<backend::mi355>())
        : impl(nullptr, inbox, outbox, r2c_direction) {}
    fft3d_r2c(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              plan_options const = default_options<backend::mi355>())
        : impl(stream, inbox, outbox, r2c_direction) {}
#ifdef HEFFTE_MI355_HAVE_MPI
    fft3d_r2c(box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction, MPI_Comm const comm,
              plan_options const = default_options<backend::mi355>())
        : impl(nullptr, inbox, outbox, r2c_direction) {
        mi355_detail::require_one_rank(comm);
    }
    fft3d_r2c(hipStream_t stream, box3d<index> const inbox, box3d<index> const outbox, int const r2c_direction,
              MPI_Comm const comm, plan_options const = default_options<backend::mi355>())
        : impl(stream, inbox, outbox, r2c_direction) {
        mi355_detail::require_one_rank(comm);
    }
#endif
};

}  // namespace heffte

#endif
