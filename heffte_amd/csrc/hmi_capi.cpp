// extern "C" boundary of libheffte_mi355.so (declared in include/heffte_mi355.h).
// Every entry point catches C++ exceptions and turns them into status codes,
// as heFFTe's C binding does (src/heffte_c.cpp:248-273).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/heffte_mi355.h"
#include "hmi_batch.h"
#include "hmi_plan.h"

struct heffte_mi355_plan_s {
    uint32_t magic;
    hmi::fft_plan *impl;
};

struct heffte_mi355_batch_s {
    uint32_t magic;
    hmi::batch_plan *impl;
};

namespace {

constexpr uint32_t kMagic = 0x4d493335u;  // "MI35"
thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

template <typename F> int guarded(F &&f) {
    try {
        g_last_error.clear();
        f();
        return HMI_SUCCESS;
    } catch (const std::invalid_argument &e) {
        return fail(HMI_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc &) {
        return fail(HMI_ERROR_DEVICE, "host allocation failed");
    } catch (const std::runtime_error &e) {
        return fail(HMI_ERROR_DEVICE, e.what());
    } catch (...) {
        return fail(HMI_ERROR_DEVICE, "unknown error");
    }
}

hmi::box3 make_box(const long long low[3], const long long high[3], const int *order) {
    if (!low || !high) throw std::invalid_argument("null box bounds");
    hmi::box3 b;
    for (int d = 0; d < 3; d++) {
        b.low[d] = low[d];
        b.high[d] = high[d];
        b.order[d] = order ? order[d] : d;
    }
    return b;
}

bool valid(heffte_mi355_plan p) { return p && p->magic == kMagic && p->impl; }

int create(hmi::kind k, const long long *il, const long long *ih, const int *io, const long long *ol,
           const long long *oh, const int *oo, int r2c_dir, heffte_mi355_plan *plan) {
    if (!plan) return fail(HMI_ERROR_INVALID_ARGUMENT, "null plan pointer");
    *plan = nullptr;
    return guarded([&] {
        hmi::box3 in = make_box(il, ih, io), out = make_box(ol, oh, oo);
        auto *impl = new hmi::fft_plan(k, in, out, r2c_dir);
        *plan = new heffte_mi355_plan_s{kMagic, impl};
    });
}

int run(heffte_mi355_plan plan, int dir, bool dbl, bool in_real, bool out_real, int batch, const void *in, void *out,
        void *ws, int scale) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (scale < 0 || scale > 2) return fail(HMI_ERROR_INVALID_ARGUMENT, "scale must be 0, 1 or 2");
    return guarded([&] { plan->impl->execute(dir, dbl, in_real, out_real, batch, in, out, ws, scale); });
}

int run_checked(heffte_mi355_plan plan, bool r2r, int dir, bool dbl, bool in_real, bool out_real, int batch,
                const void *in, void *out, void *ws, int scale) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    const hmi::kind k = plan->impl->get_kind();
    const bool is_r2r = (k == hmi::kind::cos || k == hmi::kind::sin || k == hmi::kind::cos1);
    if (r2r != is_r2r)
        return fail(HMI_ERROR_INVALID_ARGUMENT, r2r ? "real-to-real call on a complex plan"
                                                    : "complex call on a real-to-real plan");
    if (k == hmi::kind::r2c && ((dir < 0 && !in_real) || (dir > 0 && !out_real)))
        return fail(HMI_ERROR_INVALID_ARGUMENT, "an r2c plan takes real input forward and gives real output backward");
    return run(plan, dir, dbl, in_real, out_real, batch, in, out, ws, scale);
}

int convolve_spectrum(heffte_mi355_plan plan, bool dbl, int batch, const void *x, const void *yhat, void *out,
                      void *ws, int scale) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (plan->impl->get_kind() != hmi::kind::c2c) return fail(HMI_ERROR_INVALID_ARGUMENT, "convolution needs a c2c plan");
    if (scale < 0 || scale > 2) return fail(HMI_ERROR_INVALID_ARGUMENT, "scale must be 0, 1 or 2");
    return guarded([&] { plan->impl->convolve_spectrum(dbl, batch, x, yhat, out, ws, scale); });
}

int convolve(heffte_mi355_plan plan, bool dbl, int batch, const void *x, const void *y, void *out, void *ws,
             int scale) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (plan->impl->get_kind() != hmi::kind::c2c) return fail(HMI_ERROR_INVALID_ARGUMENT, "convolution needs a c2c plan");
    if (scale < 0 || scale > 2) return fail(HMI_ERROR_INVALID_ARGUMENT, "scale must be 0, 1 or 2");
    return guarded([&] { plan->impl->convolve(dbl, batch, x, y, out, ws, scale); });
}

}  // namespace

extern "C" {

int heffte_mi355_set_default_options(heffte_mi355_plan_options *options) {
    if (!options) return fail(HMI_ERROR_INVALID_ARGUMENT, "null options");
    // default_plan_options of the new tag: the fused executor absorbs the
    // permutation, so no reorder pass is requested.
    options->use_reorder = 0;
    options->algorithm = 0;
    options->use_pencils = 1;
    options->use_gpu_aware = 1;
    return HMI_SUCCESS;
}

int heffte_mi355_plan_create(long long const inbox_low[3], long long const inbox_high[3], int const *inbox_order,
                             long long const outbox_low[3], long long const outbox_high[3], int const *outbox_order,
                             heffte_mi355_plan_options const *, heffte_mi355_plan *plan) {
    return create(hmi::kind::c2c, inbox_low, inbox_high, inbox_order, outbox_low, outbox_high, outbox_order, -1,
                  plan);
}

int heffte_mi355_plan_create_r2c(long long const inbox_low[3], long long const inbox_high[3], int const *inbox_order,
                                 long long const outbox_low[3], long long const outbox_high[3],
                                 int const *outbox_order, int r2c_direction, heffte_mi355_plan_options const *,
                                 heffte_mi355_plan *plan) {
    return create(hmi::kind::r2c, inbox_low, inbox_high, inbox_order, outbox_low, outbox_high, outbox_order,
                  r2c_direction, plan);
}

int heffte_mi355_plan_create_r2r(int kind, long long const inbox_low[3], long long const inbox_high[3],
                                 int const *inbox_order, long long const outbox_low[3],
                                 long long const outbox_high[3], int const *outbox_order,
                                 heffte_mi355_plan_options const *, heffte_mi355_plan *plan) {
    if (kind != HMI_KIND_COS && kind != HMI_KIND_SIN && kind != HMI_KIND_COS1) {
        if (plan) *plan = nullptr;
        return fail(HMI_ERROR_INVALID_ARGUMENT, "r2r kind must be HMI_KIND_COS, HMI_KIND_SIN or HMI_KIND_COS1");
    }
    const hmi::kind k = kind == HMI_KIND_COS ? hmi::kind::cos : (kind == HMI_KIND_SIN ? hmi::kind::sin : hmi::kind::cos1);
    return create(k, inbox_low, inbox_high, inbox_order, outbox_low, outbox_high, outbox_order, -1, plan);
}

int heffte_mi355_plan_destroy(heffte_mi355_plan plan) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    delete plan->impl;
    plan->impl = nullptr;
    plan->magic = 0;
    delete plan;
    return HMI_SUCCESS;
}

long long heffte_mi355_size_inbox(heffte_mi355_plan plan) { return valid(plan) ? plan->impl->size_inbox() : -1; }
long long heffte_mi355_size_outbox(heffte_mi355_plan plan) { return valid(plan) ? plan->impl->size_outbox() : -1; }
long long heffte_mi355_size_workspace(heffte_mi355_plan plan) {
    return valid(plan) ? plan->impl->size_workspace() : -1;
}

int heffte_mi355_get_kind(heffte_mi355_plan plan) { return valid(plan) ? (int)plan->impl->get_kind() : -1; }

int heffte_mi355_set_stream(heffte_mi355_plan plan, void *hip_stream) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    plan->impl->set_stream(static_cast<hipStream_t>(hip_stream));
    return HMI_SUCCESS;
}

int heffte_mi355_set_precision(heffte_mi355_plan plan, int precision) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (precision < HMI_PRECISION_FP32 || precision > HMI_PRECISION_SPLIT_BF16)
        return fail(HMI_ERROR_INVALID_ARGUMENT, "unknown precision form");
    plan->impl->set_precision(precision);
    return HMI_SUCCESS;
}

int heffte_mi355_plan_matrix_core_passes(heffte_mi355_plan plan, int direction, int in_place, int *mfma_passes,
                                         int *total_passes) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (!mfma_passes || !total_passes) return fail(HMI_ERROR_INVALID_ARGUMENT, "null output pointer");
    if (direction != -1 && direction != 1) return fail(HMI_ERROR_INVALID_ARGUMENT, "direction must be -1 or +1");
    return guarded([&] { *mfma_passes = plan->impl->matrix_core_passes(direction, in_place != 0, total_passes); });
}

int heffte_mi355_plan_prepare(heffte_mi355_plan plan, int double_precision) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    return guarded([&] { plan->impl->prepare(double_precision != 0); });
}

int heffte_mi355_plan_set_timing(heffte_mi355_plan plan, int enable) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    return guarded([&] { plan->impl->set_timing(enable != 0); });
}

char const *heffte_mi355_plan_timing_report(heffte_mi355_plan plan) {
    static thread_local std::string report;
    report.clear();
    if (!valid(plan)) {
        fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
        return nullptr;
    }
    const int rc = guarded([&] { report = plan->impl->timing_report(); });
    return rc == HMI_SUCCESS ? report.c_str() : nullptr;
}

long long heffte_mi355_size_convolve_workspace(heffte_mi355_plan plan) {
    return valid(plan) && plan->impl->get_kind() == hmi::kind::c2c ? plan->impl->convolve_workspace() : -1;
}

int heffte_mi355_forward_c2c(heffte_mi355_plan p, int b, void const *in, void *out, void *ws, int s) {
    return run_checked(p, false, -1, false, false, false, b, in, out, ws, s);
}
int heffte_mi355_forward_z2z(heffte_mi355_plan p, int b, void const *in, void *out, void *ws, int s) {
    return run_checked(p, false, -1, true, false, false, b, in, out, ws, s);
}
int heffte_mi355_forward_s2c(heffte_mi355_plan p, int b, float const *in, void *out, void *ws, int s) {
    return run_checked(p, false, -1, false, true, false, b, in, out, ws, s);
}
int heffte_mi355_forward_d2z(heffte_mi355_plan p, int b, double const *in, void *out, void *ws, int s) {
    return run_checked(p, false, -1, true, true, false, b, in, out, ws, s);
}
int heffte_mi355_backward_c2c(heffte_mi355_plan p, int b, void const *in, void *out, void *ws, int s) {
    return run_checked(p, false, 1, false, false, false, b, in, out, ws, s);
}
int heffte_mi355_backward_z2z(heffte_mi355_plan p, int b, void const *in, void *out, void *ws, int s) {
    return run_checked(p, false, 1, true, false, false, b, in, out, ws, s);
}
int heffte_mi355_backward_c2s(heffte_mi355_plan p, int b, void const *in, float *out, void *ws, int s) {
    return run_checked(p, false, 1, false, false, true, b, in, out, ws, s);
}
int heffte_mi355_backward_z2d(heffte_mi355_plan p, int b, void const *in, double *out, void *ws, int s) {
    return run_checked(p, false, 1, true, false, true, b, in, out, ws, s);
}
int heffte_mi355_convolve_c2c(heffte_mi355_plan p, int b, void const *x, void const *y, void *out, void *ws, int s) {
    return convolve(p, false, b, x, y, out, ws, s);
}
int heffte_mi355_convolve_z2z(heffte_mi355_plan p, int b, void const *x, void const *y, void *out, void *ws, int s) {
    return convolve(p, true, b, x, y, out, ws, s);
}
int heffte_mi355_convolve_spectrum_c2c(heffte_mi355_plan p, int b, void const *x, void const *yhat, void *out,
                                       void *ws, int s) {
    return convolve_spectrum(p, false, b, x, yhat, out, ws, s);
}
int heffte_mi355_convolve_spectrum_z2z(heffte_mi355_plan p, int b, void const *x, void const *yhat, void *out,
                                       void *ws, int s) {
    return convolve_spectrum(p, true, b, x, yhat, out, ws, s);
}
int heffte_mi355_forward_s2s(heffte_mi355_plan p, int b, float const *in, float *out, void *ws, int s) {
    return run_checked(p, true, -1, false, true, true, b, in, out, ws, s);
}
int heffte_mi355_forward_d2d(heffte_mi355_plan p, int b, double const *in, double *out, void *ws, int s) {
    return run_checked(p, true, -1, true, true, true, b, in, out, ws, s);
}
int heffte_mi355_backward_s2s(heffte_mi355_plan p, int b, float const *in, float *out, void *ws, int s) {
    return run_checked(p, true, 1, false, true, true, b, in, out, ws, s);
}
int heffte_mi355_backward_d2d(heffte_mi355_plan p, int b, double const *in, double *out, void *ws, int s) {
    return run_checked(p, true, 1, true, true, true, b, in, out, ws, s);
}

int heffte_mi355_execute_c2c_noalloc(heffte_mi355_plan plan, int direction, int double_precision, int batch,
                                     void const *input, void *output, int scale) {
    if (!valid(plan)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed plan");
    if (plan->impl->get_kind() != hmi::kind::c2c)
        return fail(HMI_ERROR_INVALID_ARGUMENT, "the no-allocation call takes a c2c plan");
    if (direction != -1 && direction != 1) return fail(HMI_ERROR_INVALID_ARGUMENT, "direction must be -1 or +1");
    if (scale < 0 || scale > 2) return fail(HMI_ERROR_INVALID_ARGUMENT, "scale must be 0, 1 or 2");
    return guarded([&] {
        plan->impl->execute(direction, double_precision != 0, false, false, batch, input, output, nullptr, scale,
                            nullptr, true);
    });
}

int heffte_mi355_reshape(int elem_bytes, long long const low[3], long long const high[3], int const in_order[3],
                         int const out_order[3], int batch, void const *input, void *output, void *hip_stream) {
    return guarded([&] {
        if (!in_order || !out_order) throw std::invalid_argument("null order");
        hmi::box3 b = make_box(low, high, in_order);
        hmi::permute(elem_bytes, b, out_order, batch, input, output, static_cast<hipStream_t>(hip_stream));
    });
}

static bool valid_batch(heffte_mi355_batch b) { return b && b->magic == kMagic && b->impl; }

int heffte_mi355_batch_create(int count, int const *sizes, long long const *in_offsets, long long const *out_offsets,
                              heffte_mi355_batch *batch) {
    if (!batch) return fail(HMI_ERROR_INVALID_ARGUMENT, "null batch pointer");
    *batch = nullptr;
    return guarded([&] {
        auto *impl = new hmi::batch_plan(count, sizes, in_offsets, out_offsets);
        *batch = new heffte_mi355_batch_s{kMagic, impl};
    });
}

int heffte_mi355_batch_destroy(heffte_mi355_batch batch) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    delete batch->impl;
    batch->impl = nullptr;
    batch->magic = 0;
    delete batch;
    return HMI_SUCCESS;
}

int heffte_mi355_batch_set_stream(heffte_mi355_batch batch, void *hip_stream) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    batch->impl->set_stream(static_cast<hipStream_t>(hip_stream));
    return HMI_SUCCESS;
}

int heffte_mi355_batch_set_graph(heffte_mi355_batch batch, int enable) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    batch->impl->set_graph(enable != 0);
    return HMI_SUCCESS;
}

int heffte_mi355_batch_set_autotune(heffte_mi355_batch batch, int enable) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    batch->impl->set_autotune(enable != 0);
    return HMI_SUCCESS;
}

int heffte_mi355_batch_set_fused(heffte_mi355_batch batch, int enable) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    batch->impl->set_fused(enable != 0);
    return HMI_SUCCESS;
}

int heffte_mi355_batch_fused_count(heffte_mi355_batch batch, int double_precision, int *count) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    if (!count) return fail(HMI_ERROR_INVALID_ARGUMENT, "null count pointer");
    *count = batch->impl->fused_count(double_precision != 0);
    return HMI_SUCCESS;
}

long long heffte_mi355_batch_size_input(heffte_mi355_batch batch) {
    return valid_batch(batch) ? batch->impl->input_elems() : -1;
}
long long heffte_mi355_batch_size_output(heffte_mi355_batch batch) {
    return valid_batch(batch) ? batch->impl->output_elems() : -1;
}

long long heffte_mi355_batch_graph_captures(heffte_mi355_batch batch) {
    return valid_batch(batch) ? batch->impl->graph_captures() : -1;
}

static int batch_run(heffte_mi355_batch batch, int dir, bool dbl, void const *in, void *out, int scale) {
    if (!valid_batch(batch)) return fail(HMI_ERROR_CORRUPTED_PLAN, "invalid or destroyed batch");
    if (scale < 0 || scale > 2) return fail(HMI_ERROR_INVALID_ARGUMENT, "scale must be 0, 1 or 2");
    return guarded([&] { batch->impl->execute(dir, dbl, in, out, scale); });
}

int heffte_mi355_batch_forward_c2c(heffte_mi355_batch b, void const *in, void *out, int s) {
    return batch_run(b, -1, false, in, out, s);
}
int heffte_mi355_batch_forward_z2z(heffte_mi355_batch b, void const *in, void *out, int s) {
    return batch_run(b, -1, true, in, out, s);
}
int heffte_mi355_batch_backward_c2c(heffte_mi355_batch b, void const *in, void *out, int s) {
    return batch_run(b, 1, false, in, out, s);
}
int heffte_mi355_batch_backward_z2z(heffte_mi355_batch b, void const *in, void *out, int s) {
    return batch_run(b, 1, true, in, out, s);
}

char const *heffte_mi355_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
