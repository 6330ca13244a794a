/*
 * heffte_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in FP64 (with long-double twiddles), of what heFFTe 2.4.1
 * computes on its single-rank 3D FFT path.  It is the checker for the GPU
 * library in heffte_amd/: only the tests, the smoke check and the CPU-baseline
 * leg of bench.py load it.  Nothing in the product links or calls it.
 *
 * Reference semantics restated here (paths relative to the heFFTe tree):
 *   box3d layout / order ......... include/heffte_geometry.h:67-133
 *   in-node transpose map ........ src/heffte_reshape3d.cpp:179-206,
 *                                  include/heffte_pack3d.h:134-196
 *   c2c / r2c / c2r pipeline ..... src/heffte_compute_transform.cpp:15-258
 *   scale factors ................ include/heffte_fft3d.h:545-547,577,635-650
 *   DCT (cos) conventions ........ include/heffte_r2r_executor.h:43-87,221-245
 *   sign convention .............. include/heffte_backend_fftw.h:143-147
 *                                  (FFTW_FORWARD = exp(-2 pi i jk/n))
 *
 * Every function is pinned by tests/test_oracle_*.py against the reference's
 * own known-answer vectors (test/test_units_nompi.cpp, test/test_cos.cpp),
 * brute-force long-double DFTs and numpy.fft; see the per-function notes.
 *
 * Complex data is interleaved (re, im) doubles.
 */
#ifndef HEFFTE_ORACLE_H
#define HEFFTE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A box of global indexes, inclusive low/high, order = (fast, mid, slow). */
typedef struct {
    long long low[3];
    long long high[3];
    int order[3];
} orc_box;

/* Number of indexes in the box (0 if empty). */
long long orc_box_count(const orc_box *b);

/* Memory stride of global axis d inside a box stored in its order. */
void orc_box_strides(const orc_box *b, long long stride[3]);

/* Threads used by the 3D routines (OpenMP); returns the value in effect. */
int orc_set_threads(int nthreads);

/* In-node transpose (reshape3d_transpose): dst (box `out`) receives the
 * elements of src (box `in`, same index set, possibly other order).
 * elem_bytes is the size of one element.  Returns 0 or -1 on box mismatch. */
int orc_transpose(const orc_box *in, const orc_box *out, int elem_bytes,
                  const void *src, void *dst);

/* Brute-force DFT in long double, O(n^2), exact index reduction (jk mod n).
 * sign = -1 forward, +1 backward (unnormalised). */
void orc_dft1d_ld(int n, int sign, const double *in, double *out);

/* FP64 mixed-radix FFT of one contiguous complex line (out-of-place). */
void orc_fft1d(int n, int sign, const double *in, double *out);

/* 3D c2c (fft3d::forward/backward with complex in/out).
 * in is stored in inbox order, out in outbox order; the two boxes must hold the
 * same indexes.  scale multiplies every output (1.0 = scale::none).
 * batch transforms are contiguous with strides count(inbox)/count(outbox). */
int orc_fft3d_c2c(const orc_box *inbox, const orc_box *outbox, int sign,
                  double scale, int batch, const double *in, double *out);

/* fft3d forward with real input (complex output, full spectrum). */
int orc_fft3d_real_c2c(const orc_box *inbox, const orc_box *outbox,
                       double scale, int batch, const double *in_real, double *out);

/* fft3d_r2c forward: real inbox -> half-spectrum outbox (= inbox.r2c(dir)). */
int orc_fft3d_r2c(const orc_box *inbox, const orc_box *outbox, int r2c_dir,
                  double scale, int batch, const double *in_real, double *out);

/* fft3d_r2c backward: half spectrum (outbox) -> real (inbox).  The missing
 * half is filled by Hermitian symmetry X[-k] = conj(X[k]); for the planes
 * k_dir = 0 and k_dir = n/2 (even n) only entries with k_dir <= n/2 are read,
 * as a c2r transform does. */
int orc_fft3d_c2r(const orc_box *inbox, const orc_box *outbox, int r2c_dir,
                  double scale, int batch, const double *in, double *out_real);

/* rtransform<..._cos>: DCT-II along every axis on forward (sign=-1),
 * 2*DCT-III on backward (sign=+1), the rocFFT/stock convention of
 * include/heffte_r2r_executor.h.  Real in/out, boxes as for c2c. */
int orc_r2r_cos(const orc_box *inbox, const orc_box *outbox, int sign,
                double scale, int batch, const double *in, double *out);

/* One line of the r2r transform (kind 0 = cos, 1 = sin, 2 = cos1) of length n, as the
 * 1D executor real2real_executor applies it (include/heffte_r2r_executor.h:221-245). */
void orc_r2r_1d(int kind, int sign, int n, const double *x, double *y);

/* Same for the sine transform (DST-II forward, 2*DST-III backward). */
int orc_r2r_sin(const orc_box *inbox, const orc_box *outbox, int sign,
                double scale, int batch, const double *in, double *out);

/* rtransform<rocfft_cos1>: DCT-I (REDFT00) along every axis on forward, 2 x
 * REDFT00 on backward (src/heffte_backend_rocm.hip.cpp:269-328); every extent
 * must be at least 2 (returns -1 otherwise). */
int orc_r2r_cos1(const orc_box *inbox, const orc_box *outbox, int sign,
                 double scale, int batch, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
