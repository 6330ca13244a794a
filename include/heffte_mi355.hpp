/*
 * heffte_mi355.hpp -- C++ front end of the MI355X single-rank 3D FFT backend.
 *
 * It gives callers of heFFTe the interface they use (include/heffte_fft3d.h,
 * include/heffte_fft3d_r2c.h, include/heffte_geometry.h) for one rank:
 *
 *     heffte::box3d<> inbox({0,0,0}, {511,511,511}, {0,1,2});
 *     heffte::box3d<> outbox({0,0,0}, {511,511,511}, {1,2,0});
 *     heffte::fft3d<heffte::backend::mi355> fft(inbox, outbox, MPI_COMM_SELF);
 *     fft.forward(input, output, workspace, heffte::scale::symmetric);
 *
 * Two ways in:
 *   - inside a heFFTe build, include heffte.h first: this header then adds only the backend
 *     (heffte_backend_mi355.h: the mi355 tags, their traits and the fft3d / fft3d_r2c
 *     specializations) to heFFTe's own declarations, so existing code switches by changing the
 *     backend tag;
 *   - without heFFTe (standalone), it declares the parts of heFFTe's interface the backend
 *     builds on -- scale, plan_options, default_options, box3d, the tag / trait templates and
 *     the fft3d / fft3d_r2c / rtransform / fft2d templates, with the reference's names,
 *     signatures and semantics -- and then includes heffte_backend_mi355.h; heffte::gpu::vector
 *     and heffte::gpu::transfer are the device container and host transfers.
 * Every transform runs in the library's HIP kernels through the C ABI of heffte_mi355.h; no
 * header holds arithmetic.
 */
#ifndef HEFFTE_MI355_HPP
#define HEFFTE_MI355_HPP

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <type_traits>
#include <vector>

#if defined(HEFFTE_FFT3D_R2C_H)

// heffte.h is already included: its declarations are used as they are
#include "heffte_backend_mi355.h"

#else

#define HEFFTE_MI355_STANDALONE 1

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

//! Location tags (include/heffte_common.h:38-45).
namespace tag {
struct cpu {};
struct gpu {};
}  // namespace tag

//! Backend traits (include/heffte_common.h:225-258,264-265,439): primary templates; the mi355
//! tags specialize them in heffte_backend_mi355.h.
namespace backend {
template <typename tag> struct is_enabled : std::false_type {};
template <typename backend_tag, typename std::enable_if<is_enabled<backend_tag>::value, void *>::type = nullptr>
struct buffer_traits {
    using location = tag::cpu;
    template <typename T> using container = std::vector<T>;
};
template <typename backend_tag> inline std::string name() { return "unknown"; }
template <typename backend_tag> struct uses_fft_types : std::true_type {};
}  // namespace backend

//! Default options of a tag (include/heffte_common.h:761).
template <typename> struct default_plan_options {};

//! heffte::plan_options (include/heffte_plan_logic.h:131-176); accepted, no effect on one rank.
struct plan_options {
    template <typename backend_tag>
    plan_options(backend_tag const)
        : use_reorder(default_plan_options<backend_tag>::use_reorder), algorithm(reshape_algorithm::alltoallv),
          use_pencils(true), use_gpu_aware(true) {}
    plan_options(bool reorder, reshape_algorithm alg, bool pencils)
        : use_reorder(reorder), algorithm(alg), use_pencils(pencils), use_gpu_aware(true) {}
    bool use_reorder;
    reshape_algorithm algorithm;
    bool use_pencils;
    bool use_gpu_aware;
};

//! default_options (include/heffte_plan_logic.h:248-251).
template <typename backend_tag> plan_options default_options() { return plan_options(backend_tag()); }

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

//! The transform templates (include/heffte_fft3d.h:239, :703, :749; include/heffte_fft3d_r2c.h:45);
//! only the specializations of heffte_backend_mi355.h are defined.
template <typename backend_tag, typename index = int> class fft3d;
template <typename backend_tag, typename index = int> class fft3d_r2c;
template <typename backend_tag, typename index = int> using fft2d = fft3d<backend_tag, index>;
template <typename backend_tag, typename index = int> using rtransform = fft3d<backend_tag, index>;

}  // namespace heffte

#include "heffte_backend_mi355.h"

namespace heffte {
namespace gpu {
//! Device container of the mi355 tags (the role of heffte::gpu::vector).
template <typename T> using vector = mi355_detail::device_vector<T>;

//! gpu::transfer::load / unload (include/heffte_backend_data_transfer.h:29-177).
struct transfer {
    template <typename T> static vector<T> load(std::vector<T> const &h) { return vector<T>(h.data(), h.data() + h.size()); }
    template <typename T> static std::vector<T> unload(vector<T> const &d) {
        std::vector<T> h(d.size());
        if (!h.empty())
            mi355_detail::hip_check(hipMemcpy(h.data(), d.data(), d.size() * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
};
}  // namespace gpu
}  // namespace heffte

#endif  // heffte.h included
#endif
