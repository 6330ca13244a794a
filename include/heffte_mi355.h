/*
 * heffte_mi355.h -- C ABI of the MI355X (gfx950) single-rank 3D FFT backend.
 *
 * The entry points mirror heFFTe's C binding (include/heffte_c.h) for a
 * communicator of size one: plan creation from boxes with an index order,
 * size queries, and forward/backward calls with a caller-provided workspace and
 * a scale mode.  Differences, all forced by the boundary being a plain C ABI:
 *   - every call returns a status code (no C++ exception crosses the ABI);
 *   - there is no MPI_Comm argument: the plan covers one rank that owns the
 *     whole world, so inbox and outbox hold the same indexes;
 *   - box indexes are long long (heFFTe's box3d<long long> covers box3d<int>);
 *   - a batch size argument carries fft3d::forward(batch_size, ...)
 *     (include/heffte_fft3d.h:392-402) through the ABI;
 *   - the HIP stream is bound to the plan (fft3d(gpu_stream, ...),
 *     include/heffte_fft3d.h:287-290) by heffte_mi355_set_stream.
 *
 * All pointers passed to the forward/backward calls are device pointers.
 * Calls are asynchronous on the plan's stream, as heFFTe's GPU backends are.
 */
#ifndef HEFFTE_MI355_H
#define HEFFTE_MI355_H

#ifdef __cplusplus
extern "C" {
#endif

/* Backend id a maintainer would add next to Heffte_BACKEND_ROCFFT
 * (include/heffte_c_defines.h:82). */
#define Heffte_BACKEND_MI355 12

/* Status codes; 0-3 keep the meaning of include/heffte_c.h:49-50,74. */
#define HMI_SUCCESS 0
#define HMI_ERROR_INVALID_BACKEND 1
#define HMI_ERROR_INVALID_ARGUMENT 2
#define HMI_ERROR_CORRUPTED_PLAN 3
#define HMI_ERROR_DEVICE 4

/* Scale modes, same values as Heffte_SCALE_* (include/heffte_c_defines.h:152-162). */
#define HMI_SCALE_NONE 0
#define HMI_SCALE_FULL 1
#define HMI_SCALE_SYMMETRIC 2

/* Transform kinds. */
#define HMI_KIND_C2C 0 /* heffte::fft3d                                   */
#define HMI_KIND_R2C 1 /* heffte::fft3d_r2c                               */
#define HMI_KIND_COS 2 /* heffte::rtransform<backend::*_cos> (DCT-II/III) */
#define HMI_KIND_SIN 3 /* heffte::rtransform<backend::*_sin> (DST-II/III) */
#define HMI_KIND_COS1 4 /* heffte::rtransform<backend::rocfft_cos1> (DCT-I) */

/* Same layout as heffte_plan_options (include/heffte_c_defines.h:113-122).
 * On a single rank the options do not change any result; they are accepted
 * so that callers can pass their existing option structs. */
typedef struct {
    int use_reorder;
    int algorithm;
    int use_pencils;
    int use_gpu_aware;
} heffte_mi355_plan_options;

typedef struct heffte_mi355_plan_s *heffte_mi355_plan;

/* Replaces heffte_set_default_options (include/heffte_c.h:25). */
int heffte_mi355_set_default_options(heffte_mi355_plan_options *options);

/* Replaces heffte_plan_create (include/heffte_c.h:52-54).  inbox/outbox
 * orders may be NULL for (0, 1, 2).  options may be NULL. */
int heffte_mi355_plan_create(long long const inbox_low[3], long long const inbox_high[3], int const *inbox_order,
                             long long const outbox_low[3], long long const outbox_high[3], int const *outbox_order,
                             heffte_mi355_plan_options const *options, heffte_mi355_plan *plan);

/* Replaces heffte_plan_create_r2c (include/heffte_c.h:63-65). */
int heffte_mi355_plan_create_r2c(long long const inbox_low[3], long long const inbox_high[3], int const *inbox_order,
                                 long long const outbox_low[3], long long const outbox_high[3],
                                 int const *outbox_order, int r2c_direction,
                                 heffte_mi355_plan_options const *options, heffte_mi355_plan *plan);

/* Replaces the construction of heffte::rtransform<tag> (include/heffte_fft3d.h:748-749;
 * the reference C binding has no r2r entry).  kind is HMI_KIND_COS, HMI_KIND_SIN or
 * HMI_KIND_COS1; DCT-I needs at least 2 points along every axis and scales by
 * 1 / (64 prod(n_d - 1)) under HMI_SCALE_FULL (include/heffte_fft3d.h:643-646). */
int heffte_mi355_plan_create_r2r(int kind, long long const inbox_low[3], long long const inbox_high[3],
                                 int const *inbox_order, long long const outbox_low[3],
                                 long long const outbox_high[3], int const *outbox_order,
                                 heffte_mi355_plan_options const *options, heffte_mi355_plan *plan);

/* Replaces heffte_plan_destroy (include/heffte_c.h:75). */
int heffte_mi355_plan_destroy(heffte_mi355_plan plan);

/* Replace heffte_size_inbox / _outbox / _workspace (include/heffte_c.h:81-91).
 * The workspace is counted in elements of the transform's complex type (real
 * type for r2r), per transform: a batch needs batch * size_workspace. */
long long heffte_mi355_size_inbox(heffte_mi355_plan plan);
long long heffte_mi355_size_outbox(heffte_mi355_plan plan);
long long heffte_mi355_size_workspace(heffte_mi355_plan plan);

/* Replaces heffte_is_r2c (include/heffte_c.h:102); returns the HMI_KIND_* of the plan, -1 on error. */
int heffte_mi355_get_kind(heffte_mi355_plan plan);

/* Binds a hipStream_t (NULL = default stream); replaces the stream argument of
 * the fft3d constructor (include/heffte_fft3d.h:287-290). */
int heffte_mi355_set_stream(heffte_mi355_plan plan, void *hip_stream);

/* Precision form of the complex<float> c2c passes (no heFFTe counterpart):
 *   HMI_PRECISION_FP32       vector-ALU FP32 (default)
 *   HMI_PRECISION_SPLIT_FP16 512-point passes as 32x16 DFT GEMMs on MFMA, operands split
 *                            into two FP16 terms (3 products), FP32 accumulation,
 *                            a power-of-two scale per 32- or 16-point DFT
 *   HMI_PRECISION_SPLIT_BF16 the same with three BF16 terms (6 products)
 * Other lengths and complex<double> always use the vector-ALU passes. */
#define HMI_PRECISION_FP32 0
#define HMI_PRECISION_SPLIT_FP16 1
#define HMI_PRECISION_SPLIT_BF16 2
int heffte_mi355_set_precision(heffte_mi355_plan plan, int precision);
/* How many passes of the plan's complex<float> transform in the given direction (-1 forward,
 * +1 backward; in_place = input and output the same array) run on the matrix cores under the
 * current precision form (*mfma_passes) out of all its passes (*total_passes).  The passes of
 * other lengths keep the vector-ALU form whatever the precision form; this reports which.  No
 * heFFTe counterpart. */
int heffte_mi355_plan_matrix_core_passes(heffte_mi355_plan plan, int direction, int in_place, int *mfma_passes,
                                         int *total_passes);

/* Uploads the plan's device tables (twiddles, and the split-precision operands of the current
 * precision form) for every transform direction and overload, in single (double_precision = 0)
 * or double precision.  No heFFTe counterpart: heFFTe's rocFFT executors create their plans
 * lazily on the first call (include/heffte_backend_rocm.h:705-715).  Without it the tables are
 * uploaded on a plan's first call; a first call on a stream under capture, which cannot
 * record that synchronous upload, fails with HMI_ERROR_DEVICE and enqueues nothing.  Call it
 * again after heffte_mi355_set_precision selects a split-precision form. */
int heffte_mi355_plan_prepare(heffte_mi355_plan plan, int double_precision);

/* Per-pass timing (diagnostic, no heFFTe counterpart; heFFTe's add_trace/finalize_tracing,
 * include/heffte_trace.h, record host-side events per stage instead).  enable = 1 makes every
 * later call on the plan record a hipEvent on the plan's stream before its first pass and after
 * each pass; 0 stops; either drops the records so far.  The report waits for the recorded events
 * and returns one line per pass kernel, "kernel name<TAB>launches<TAB>total milliseconds", over
 * every call since timing was enabled (NULL on error; the text stays valid until the next report
 * on the calling thread). */
int heffte_mi355_plan_set_timing(heffte_mi355_plan plan, int enable);
char const *heffte_mi355_plan_timing_report(heffte_mi355_plan plan);

/* Forward transforms; replace heffte_forward_{s2c,c2c,d2z,z2z}_buffered
 * (include/heffte_c.h:164-179) with a batch size.  workspace may be NULL, in
 * which case it is allocated for the call (fft3d::forward(in, out),
 * include/heffte_fft3d.h:353-360).  in == out is allowed for c2c/z2z. */
int heffte_mi355_forward_c2c(heffte_mi355_plan plan, int batch, void const *input, void *output, void *workspace,
                             int scale);
int heffte_mi355_forward_z2z(heffte_mi355_plan plan, int batch, void const *input, void *output, void *workspace,
                             int scale);
int heffte_mi355_forward_s2c(heffte_mi355_plan plan, int batch, float const *input, void *output, void *workspace,
                             int scale);
int heffte_mi355_forward_d2z(heffte_mi355_plan plan, int batch, double const *input, void *output, void *workspace,
                             int scale);

/* Backward transforms; replace heffte_backward_{c2s,c2c,z2d,z2z}_buffered
 * (include/heffte_c.h:241-256). */
int heffte_mi355_backward_c2c(heffte_mi355_plan plan, int batch, void const *input, void *output, void *workspace,
                              int scale);
int heffte_mi355_backward_z2z(heffte_mi355_plan plan, int batch, void const *input, void *output, void *workspace,
                              int scale);
int heffte_mi355_backward_c2s(heffte_mi355_plan plan, int batch, void const *input, float *output, void *workspace,
                              int scale);
int heffte_mi355_backward_z2d(heffte_mi355_plan plan, int batch, void const *input, double *output, void *workspace,
                              int scale);

/* Circular convolution through the transform (no heFFTe entry point: the sequence of
 * benchmarks/convolution.cpp:86-95 -- forward, pointwise product, backward -- in one call):
 *   output = backward(forward(x) * forward(y))
 * on a c2c plan, with the backward scale applied (HMI_SCALE_FULL gives the circular
 * convolution of x and y; HMI_SCALE_NONE gives N times it).  x, y and output are in the
 * inbox layout; output may alias x or y.  The product is folded into the first backward
 * pass, so the product is never written back; x == y (one array) runs one forward
 * transform and squares its spectrum on load.  workspace holds
 * batch * heffte_mi355_size_convolve_workspace(plan) complex elements or is NULL (allocated
 * for the call). */
long long heffte_mi355_size_convolve_workspace(heffte_mi355_plan plan);
int heffte_mi355_convolve_c2c(heffte_mi355_plan plan, int batch, void const *x, void const *y, void *output,
                              void *workspace, int scale);
int heffte_mi355_convolve_z2z(heffte_mi355_plan plan, int batch, void const *x, void const *y, void *output,
                              void *workspace, int scale);
/* The same with the second operand's spectrum computed once by the caller (a filter applied to
 * many inputs): spectrum = forward(y) with HMI_SCALE_NONE, in the outbox layout, and
 *   output = backward(forward(x) * spectrum)
 * runs one forward and one backward transform.  workspace holds batch *
 * heffte_mi355_size_workspace(plan) elements or is NULL; output may alias x, not spectrum. */
int heffte_mi355_convolve_spectrum_c2c(heffte_mi355_plan plan, int batch, void const *x, void const *spectrum,
                                       void *output, void *workspace, int scale);
int heffte_mi355_convolve_spectrum_z2z(heffte_mi355_plan plan, int batch, void const *x, void const *spectrum,
                                       void *output, void *workspace, int scale);

/* No-allocation variant of heffte_mi355_forward_c2c / _z2z and heffte_mi355_backward_c2c / _z2z
 * (direction -1 forward, +1 backward; double_precision selects complex<double>).  It has no
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The call allocates
 * nothing: the workspace is a static device array of HMI_NOALLOC_SCRATCH_BYTES per GPU, held
 * from the first such call on that GPU until the process ends, and the plan's device tables
 * must have been uploaded by heffte_mi355_plan_prepare (same precision) beforehand.  A call
 * whose workspace (batch * heffte_mi355_size_workspace(plan) complex elements) exceeds the
 * scratch, or whose tables are not resident, fails with HMI_ERROR_INVALID_ARGUMENT /
 * HMI_ERROR_DEVICE and enqueues nothing.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams, other threads) race, so serialize them. */
#define HMI_NOALLOC_SCRATCH_BYTES (64ull << 20)
int heffte_mi355_execute_c2c_noalloc(heffte_mi355_plan plan, int direction, int double_precision, int batch,
                                     void const *input, void *output, int scale);

/* Real-to-real transforms of an HMI_KIND_COS/SIN/COS1 plan; replace
 * rtransform::forward/backward(float/double data, workspace)
 * (include/heffte_fft3d.h:374-384,480-490 with a real scalar type). */
int heffte_mi355_forward_s2s(heffte_mi355_plan plan, int batch, float const *input, float *output, void *workspace,
                             int scale);
int heffte_mi355_forward_d2d(heffte_mi355_plan plan, int batch, double const *input, double *output,
                             void *workspace, int scale);
int heffte_mi355_backward_s2s(heffte_mi355_plan plan, int batch, float const *input, float *output,
                              void *workspace, int scale);
int heffte_mi355_backward_d2d(heffte_mi355_plan plan, int batch, double const *input, double *output,
                              void *workspace, int scale);

/* In-node reshape: moves a box stored in in_order into out_order (same
 * indexes), element by element and bit-exact.  Replaces
 * reshape3d_transpose::apply (include/heffte_reshape3d.h:465-484) and the
 * transpose_unpacker kernel (src/heffte_backend_rocm.hip.cpp:91-134,407-444).
 * elem_bytes is 4, 8 or 16; input and output must not overlap. */
int heffte_mi355_reshape(int elem_bytes, long long const low[3], long long const high[3], int const in_order[3],
                         int const out_order[3], int batch, void const *input, void *output, void *hip_stream);

/* ---- Batches of independent transforms of different sizes (no heFFTe counterpart:
 * heFFTe batches share one geometry, include/heffte_fft3d.h:392-402).  Each transform is a
 * c2c box of sizes[3*i .. 3*i+2] in order (0, 1, 2) stored at element offset in_offsets[i]
 * of the input array and out_offsets[i] of the output array (NULL = packed back to back;
 * out_offsets NULL = same as in_offsets).  The forward/backward calls run phases over device
 * work lists (the fused x-y phase, the separate x and y phases for the transforms it does not
 * take, then the z phase; each one launch per register class of tile), replayed from a captured
 * hipGraph unless heffte_mi355_batch_set_graph(batch, 0) was called; one instantiated graph is
 * kept per (direction, scale, input, output, stream, fused setting), so calls that alternate
 * between a few argument sets replay without capturing again.  scale applies per transform
 * (1/N_i).
 * Aliasing: the output ranges of the transforms must be disjoint (checked at creation);
 * input == output only when in_offsets[i] == out_offsets[i] for every i, otherwise the input and
 * output ranges must not overlap (checked per call, HMI_ERROR_INVALID_ARGUMENT). */
typedef struct heffte_mi355_batch_s *heffte_mi355_batch;

int heffte_mi355_batch_create(int count, int const *sizes, long long const *in_offsets,
                              long long const *out_offsets, heffte_mi355_batch *batch);
int heffte_mi355_batch_destroy(heffte_mi355_batch batch);
int heffte_mi355_batch_set_stream(heffte_mi355_batch batch, void *hip_stream);
int heffte_mi355_batch_set_graph(heffte_mi355_batch batch, int enable);
/* Autotuning (on by default): on the first forward/backward call of each precision, every
 * distinct side length of the batch gets its radix plan, lines per tile and LDS staging (the
 * compile-time tile or a runtime plan) by timing the candidates on the GPU; choices are cached
 * per (length, precision) in memory and in $HMI_TUNE_CACHE (default
 * $HOME/.cache/heffte_mi355/batch_tuning.txt; a file written by another build of the library is
 * ignored and rewritten).  enable = 0 takes the untimed defaults.  The
 * first call then runs timing launches on a stream of its own: do it outside stream capture. */
int heffte_mi355_batch_set_autotune(heffte_mi355_batch batch, int enable);
/* Fused x-y phase (on by default): a transform whose x and y lengths both have compile-time
 * tiles and whose x-y plane fits the LDS of one block (10240 complex<float> / 5120
 * complex<double> elements, or 20480 complex<float> elements in the 512-thread build) runs its
 * x and y axes in one pass over its planes (one read and one write of the array instead of two);
 * the others take the separate x and y phases.  With it, a transform whose z length has a
 * compile-time tile takes the column-block z phase (the z lines of up to 10240 / n2 consecutive
 * (x, y) positions per block of LDS; 5120 for complex<double>).  enable = 0 sends every
 * transform through the three separate phases.  heffte_mi355_batch_fused_count stores in *count
 * how many transforms the fused x-y phase takes in the given precision (-1 before the first
 * forward/backward call of that precision). */
int heffte_mi355_batch_set_fused(heffte_mi355_batch batch, int enable);
int heffte_mi355_batch_fused_count(heffte_mi355_batch batch, int double_precision, int *count);
/* Elements (complex) the input / output arrays must hold. */
long long heffte_mi355_batch_size_input(heffte_mi355_batch batch);
long long heffte_mi355_batch_size_output(heffte_mi355_batch batch);
int heffte_mi355_batch_forward_c2c(heffte_mi355_batch batch, void const *input, void *output, int scale);
int heffte_mi355_batch_forward_z2z(heffte_mi355_batch batch, void const *input, void *output, int scale);
int heffte_mi355_batch_backward_c2c(heffte_mi355_batch batch, void const *input, void *output, int scale);
int heffte_mi355_batch_backward_z2z(heffte_mi355_batch batch, void const *input, void *output, int scale);
/* Number of hipGraph captures the batch has made (diagnostic: a replayed call adds none). */
long long heffte_mi355_batch_graph_captures(heffte_mi355_batch batch);

/* Text of the last error raised on the calling thread ("" if none). */
char const *heffte_mi355_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HEFFTE_MI355_H */
