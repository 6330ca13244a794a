/*
 * heffte_oracle.c -- TEST INFRASTRUCTURE ONLY (see heffte_oracle.h).
 *
 * FP64 CPU restatement of heFFTe's single-rank 3D transform semantics.  The 1D
 * FFT is a recursive mixed-radix decimation-in-time Cooley-Tukey whose roots of
 * unity are evaluated once in long double and rounded to double; every twiddle
 * is read from that table (no recurrences), so the error stays at a few ulp
 * times log(n) -- the property the heFFTe stock backend lacks in float
 * (include/stock_fft/heffte_stock_algos.h:34-40 builds twiddles by recurrence).
 */
#include "heffte_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static const long double ORC_PI = 3.141592653589793238462643383279502884L;

/* ------------------------------------------------------------------ boxes */

long long orc_box_count(const orc_box *b) {
    long long c = 1;
    for (int d = 0; d < 3; d++) {
        long long s = b->high[d] - b->low[d] + 1;
        if (s <= 0) return 0;
        c *= s;
    }
    return c;
}

/* heffte_geometry.h:124 osize(i) = size[order[i]]; linear index
 * k*osize0*osize1 + j*osize0 + i over (order[0], order[1], order[2]). */
void orc_box_strides(const orc_box *b, long long stride[3]) {
    long long size[3];
    for (int d = 0; d < 3; d++) size[d] = b->high[d] - b->low[d] + 1;
    stride[b->order[0]] = 1;
    stride[b->order[1]] = size[b->order[0]];
    stride[b->order[2]] = size[b->order[0]] * size[b->order[1]];
}

static int same_indexes(const orc_box *a, const orc_box *b) {
    for (int d = 0; d < 3; d++)
        if (a->low[d] != b->low[d] || a->high[d] != b->high[d]) return 0;
    return 1;
}

static void box_sizes(const orc_box *b, long long size[3]) {
    for (int d = 0; d < 3; d++) size[d] = b->high[d] - b->low[d] + 1;
}

int orc_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* include/heffte_reshape3d.h:443-487 with the map of
 * src/heffte_reshape3d.cpp:179-206: every global index keeps its value, only
 * its position changes from the source order to the destination order. */
int orc_transpose(const orc_box *in, const orc_box *out, int elem_bytes,
                  const void *src, void *dst) {
    if (!same_indexes(in, out)) return -1;
    long long size[3], si[3], so[3];
    box_sizes(in, size);
    orc_box_strides(in, si);
    orc_box_strides(out, so);
    const char *s = (const char *)src;
    char *o = (char *)dst;
    for (long long k = 0; k < size[2]; k++)
        for (long long j = 0; j < size[1]; j++)
            for (long long i = 0; i < size[0]; i++) {
                long long a = i * si[0] + j * si[1] + k * si[2];
                long long b = i * so[0] + j * so[1] + k * so[2];
                memcpy(o + b * elem_bytes, s + a * elem_bytes, (size_t)elem_bytes);
            }
    return 0;
}

/* ------------------------------------------------------------- 1D kernels */

void orc_dft1d_ld(int n, int sign, const double *in, double *out) {
    for (int k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (int j = 0; j < n; j++) {
            long long e = ((long long)j * (long long)k) % n;
            long double ang = 2.0L * ORC_PI * (long double)e / (long double)n;
            long double c = cosl(ang), s = (long double)sign * sinl(ang);
            long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

typedef struct {
    int n, sign;
    cplx *roots;       /* roots[j] = exp(sign * 2 pi i j / n), j < n */
    int nfac, fac[64]; /* radix sequence, outermost first */
} fft_plan;

static void plan_init(fft_plan *p, int n, int sign) {
    p->n = n;
    p->sign = sign;
    p->roots = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) {
        long double ang = 2.0L * ORC_PI * (long double)j / (long double)n;
        p->roots[j].re = (double)cosl(ang);
        p->roots[j].im = (double)((long double)sign * sinl(ang));
    }
    p->nfac = 0;
    int m = n;
    while (m > 1) {
        int r;
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
        else {
            r = 3;
            while (m % r) r += 2;
        }
        p->fac[p->nfac++] = r;
        m /= r;
    }
}

static void plan_free(fft_plan *p) { free(p->roots); p->roots = NULL; }

/* out[0..n) = DFT_n of in[0], in[is], in[2 is], ...; rs = N / n (root stride).
 * Decimation in time: n = r * m, Y_q = DFT_m(x[r j + q]), then
 * X[k + m s] = sum_q (w_n^{qk} Y_q[k]) w_r^{qs}.  The combine is in place
 * because the index sets {q m + k} and {k + m s} coincide for fixed k. */
static void fft_rec(const fft_plan *p, int depth, int n, const cplx *in, long long is,
                    cplx *out, long long rs, cplx *scratch) {
    if (n == 1) { out[0] = in[0]; return; }
    int r = p->fac[depth];
    int m = n / r;
    for (int q = 0; q < r; q++)
        fft_rec(p, depth + 1, m, in + q * is, is * r, out + (long long)q * m, rs * r, scratch);
    const int N = p->n;
    const cplx *w = p->roots;
    cplx *y = scratch; /* r values */
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < r; q++) {
            cplx v = out[(long long)q * m + k];
            long long e = ((long long)q * k * rs) % N;
            cplx t = w[e];
            y[q].re = v.re * t.re - v.im * t.im;
            y[q].im = v.re * t.im + v.im * t.re;
        }
        if (r == 2) {
            out[k].re = y[0].re + y[1].re;  out[k].im = y[0].im + y[1].im;
            out[k + m].re = y[0].re - y[1].re;  out[k + m].im = y[0].im - y[1].im;
        } else if (r == 4) {
            /* w_4 = exp(sign i pi/2) = sign * i */
            double sg = (double)p->sign;
            cplx a = {y[0].re + y[2].re, y[0].im + y[2].im};
            cplx b = {y[0].re - y[2].re, y[0].im - y[2].im};
            cplx c = {y[1].re + y[3].re, y[1].im + y[3].im};
            cplx d = {y[1].re - y[3].re, y[1].im - y[3].im};
            cplx di = {-sg * d.im, sg * d.re}; /* (sign i) * d */
            out[k].re = a.re + c.re;          out[k].im = a.im + c.im;
            out[k + m].re = b.re + di.re;     out[k + m].im = b.im + di.im;
            out[k + 2 * m].re = a.re - c.re;  out[k + 2 * m].im = a.im - c.im;
            out[k + 3 * m].re = b.re - di.re; out[k + 3 * m].im = b.im - di.im;
        } else {
            long long step = (long long)N / r;
            for (int s = 0; s < r; s++) {
                double sr = 0.0, si = 0.0;
                for (int q = 0; q < r; q++) {
                    cplx t = w[(((long long)q * s) % r) * step];
                    sr += y[q].re * t.re - y[q].im * t.im;
                    si += y[q].re * t.im + y[q].im * t.re;
                }
                out[k + (long long)m * s].re = sr;
                out[k + (long long)m * s].im = si;
            }
        }
    }
}

static void plan_exec(const fft_plan *p, const cplx *in, long long is, cplx *out, cplx *scratch) {
    fft_rec(p, 0, p->n, in, is, out, 1, scratch);
}

static int max_factor(const fft_plan *p) {
    int r = 1;
    for (int i = 0; i < p->nfac; i++) if (p->fac[i] > r) r = p->fac[i];
    return r;
}

void orc_fft1d(int n, int sign, const double *in, double *out) {
    if (n <= 0) return;
    fft_plan p;
    plan_init(&p, n, sign);
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)max_factor(&p));
    plan_exec(&p, (const cplx *)in, 1, (cplx *)out, scratch);
    free(scratch);
    plan_free(&p);
}

/* In-place FFT along global axis `axis` of a canonical (x fastest) array. */
static void fft_axis(cplx *data, const long long size[3], int axis, int sign) {
    long long n = size[axis];
    if (n <= 1) return;
    long long stride[3] = {1, size[0], size[0] * size[1]};
    int o1 = (axis == 0) ? 1 : 0;
    int o2 = (axis == 2) ? 1 : 2;
    long long nlines = size[o1] * size[o2];
    fft_plan p;
    plan_init(&p, (int)n, sign);
    int mf = max_factor(&p);
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)mf);
#pragma omp for schedule(static)
        for (long long l = 0; l < nlines; l++) {
            long long a = l % size[o1], b = l / size[o1];
            cplx *base = data + a * stride[o1] + b * stride[o2];
            for (long long i = 0; i < n; i++) line[i] = base[i * stride[axis]];
            plan_exec(&p, line, 1, res, scratch);
            for (long long i = 0; i < n; i++) base[i * stride[axis]] = res[i];
        }
        free(line);
        free(res);
        free(scratch);
    }
    plan_free(&p);
}

/* Copy between a box stored in its own order and the canonical layout. */
static void gather_box(const orc_box *b, const cplx *src, cplx *canon) {
    long long size[3], s[3];
    box_sizes(b, size);
    orc_box_strides(b, s);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < size[2]; k++)
        for (long long j = 0; j < size[1]; j++)
            for (long long i = 0; i < size[0]; i++)
                canon[(k * size[1] + j) * size[0] + i] = src[i * s[0] + j * s[1] + k * s[2]];
}

static void scatter_box(const orc_box *b, const cplx *canon, cplx *dst, double scale) {
    long long size[3], s[3];
    box_sizes(b, size);
    orc_box_strides(b, s);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < size[2]; k++)
        for (long long j = 0; j < size[1]; j++)
            for (long long i = 0; i < size[0]; i++) {
                cplx v = canon[(k * size[1] + j) * size[0] + i];
                v.re *= scale;
                v.im *= scale;
                dst[i * s[0] + j * s[1] + k * s[2]] = v;
            }
}

/* src/heffte_compute_transform.cpp:15-127: reshape -> 1D FFTs along the three
 * axes -> reshape, then include/heffte_fft3d.h:662-678 multiplies by the scale.
 * The result does not depend on the intermediate layouts, so the restatement
 * transforms a canonical copy. */
int orc_fft3d_c2c(const orc_box *inbox, const orc_box *outbox, int sign,
                  double scale, int batch, const double *in, double *out) {
    if (!same_indexes(inbox, outbox)) return -1;
    long long size[3];
    box_sizes(inbox, size);
    long long cnt = orc_box_count(inbox);
    cplx *canon = (cplx *)malloc(sizeof(cplx) * (size_t)(cnt > 0 ? cnt : 1));
    for (int bt = 0; bt < batch; bt++) {
        gather_box(inbox, (const cplx *)in + bt * cnt, canon);
        for (int d = 0; d < 3; d++) fft_axis(canon, size, d, sign);
        scatter_box(outbox, canon, (cplx *)out + bt * cnt, scale);
    }
    free(canon);
    return 0;
}

int orc_fft3d_real_c2c(const orc_box *inbox, const orc_box *outbox,
                       double scale, int batch, const double *in_real, double *out) {
    long long cnt = orc_box_count(inbox);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(cnt > 0 ? cnt : 1));
    int rc = 0;
    for (int bt = 0; bt < batch && rc == 0; bt++) {
        for (long long i = 0; i < cnt; i++) { tmp[i].re = in_real[bt * cnt + i]; tmp[i].im = 0.0; }
        rc = orc_fft3d_c2c(inbox, outbox, -1, scale, 1, (const double *)tmp, out + 2 * bt * cnt);
    }
    free(tmp);
    return rc;
}

/* include/heffte_geometry.h:96-104: box3d::r2c keeps size/2 + 1 entries. */
static int check_r2c_boxes(const orc_box *rbox, const orc_box *cbox, int dir) {
    if (dir < 0 || dir > 2) return -1;
    for (int d = 0; d < 3; d++) {
        long long hi = rbox->high[d];
        if (d == dir) hi = rbox->low[d] + (rbox->high[d] - rbox->low[d] + 1) / 2;
        if (cbox->low[d] != rbox->low[d] || cbox->high[d] != hi) return -1;
    }
    return 0;
}

int orc_fft3d_r2c(const orc_box *inbox, const orc_box *outbox, int r2c_dir,
                  double scale, int batch, const double *in_real, double *out) {
    if (check_r2c_boxes(inbox, outbox, r2c_dir)) return -1;
    long long size[3], hsize[3], so[3];
    box_sizes(inbox, size);
    box_sizes(outbox, hsize);
    orc_box_strides(outbox, so);
    long long cnt = orc_box_count(inbox), hcnt = orc_box_count(outbox);
    cplx *canon = (cplx *)malloc(sizeof(cplx) * (size_t)(cnt > 0 ? cnt : 1));
    for (int bt = 0; bt < batch; bt++) {
        long long si[3];
        orc_box_strides(inbox, si);
        const double *src = in_real + bt * cnt;
        for (long long k = 0; k < size[2]; k++)
            for (long long j = 0; j < size[1]; j++)
                for (long long i = 0; i < size[0]; i++) {
                    cplx v = {src[i * si[0] + j * si[1] + k * si[2]], 0.0};
                    canon[(k * size[1] + j) * size[0] + i] = v;
                }
        for (int d = 0; d < 3; d++) fft_axis(canon, size, d, -1);
        cplx *dst = (cplx *)out + bt * hcnt;
        for (long long k = 0; k < hsize[2]; k++)
            for (long long j = 0; j < hsize[1]; j++)
                for (long long i = 0; i < hsize[0]; i++) {
                    cplx v = canon[(k * size[1] + j) * size[0] + i];
                    v.re *= scale;
                    v.im *= scale;
                    dst[i * so[0] + j * so[1] + k * so[2]] = v;
                }
    }
    free(canon);
    return 0;
}

int orc_fft3d_c2r(const orc_box *inbox, const orc_box *outbox, int r2c_dir,
                  double scale, int batch, const double *in, double *out_real) {
    /* inbox = real box, outbox = complex half box (as in fft3d_r2c). */
    if (check_r2c_boxes(inbox, outbox, r2c_dir)) return -1;
    long long size[3], hsize[3], sh[3], sr[3];
    box_sizes(inbox, size);
    box_sizes(outbox, hsize);
    orc_box_strides(outbox, sh);
    orc_box_strides(inbox, sr);
    long long cnt = orc_box_count(inbox), hcnt = orc_box_count(outbox);
    cplx *canon = (cplx *)malloc(sizeof(cplx) * (size_t)(cnt > 0 ? cnt : 1));
    for (int bt = 0; bt < batch; bt++) {
        const cplx *src = (const cplx *)in + bt * hcnt;
        for (long long k = 0; k < size[2]; k++)
            for (long long j = 0; j < size[1]; j++)
                for (long long i = 0; i < size[0]; i++) {
                    long long idx[3] = {i, j, k};
                    cplx v;
                    if (idx[r2c_dir] < hsize[r2c_dir]) {
                        v = src[i * sh[0] + j * sh[1] + k * sh[2]];
                    } else {
                        long long m[3];
                        for (int d = 0; d < 3; d++) m[d] = (size[d] - idx[d]) % size[d];
                        v = src[m[0] * sh[0] + m[1] * sh[1] + m[2] * sh[2]];
                        v.im = -v.im;
                    }
                    canon[(k * size[1] + j) * size[0] + i] = v;
                }
        for (int d = 0; d < 3; d++) fft_axis(canon, size, d, +1);
        double *dst = out_real + bt * cnt;
        for (long long k = 0; k < size[2]; k++)
            for (long long j = 0; j < size[1]; j++)
                for (long long i = 0; i < size[0]; i++)
                    dst[i * sr[0] + j * sr[1] + k * sr[2]] = scale * canon[(k * size[1] + j) * size[0] + i].re;
    }
    free(canon);
    return 0;
}

/* ------------------------------------------------------------------ r2r */

/* Length of the extended signal of one r2r line: 4N for the DCT-II/DST-II
 * pair, 4(N-1) for DCT-I (cos1_pre_pos_processor::compute_extended_length,
 * include/heffte_backend_rocm.h:165-181). */
static int r2r_ext_len(int kind, int n) { return kind == 2 ? 4 * (n - 1) : 4 * n; }

/* One DCT-I line (kind 2) as the rocFFT backend computes it, through the
 * extended signal of length L = 4(N-1):
 *   forward  src/heffte_backend_rocm.hip.cpp:269-297: real (x0 0 x1 0 .. x_{N-1} 0 .. x1 0),
 *            y[k] = Re FFT[k] (FFTW's REDFT00);
 *   backward src/heffte_backend_rocm.hip.cpp:299-328: half spectrum c[j] = x[j],
 *            c[2(N-1)-j] = x[j], real inverse FFT, y[i] = r[2i] (2 x REDFT00). */
static void cos1_line(const fft_plan *pl, int sign, int n, const double *x, double *y,
                      cplx *buf, cplx *res, cplx *scratch) {
    int L = 4 * (n - 1);
    for (int i = 0; i < L; i++) { buf[i].re = 0.0; buf[i].im = 0.0; }
    if (sign < 0) {
        for (int i = 0; i < n; i++) buf[2 * i].re = x[i];
        for (int i = 1; i < n; i++) buf[L - 2 * i].re = x[i];
        plan_exec(pl, buf, 1, res, scratch);
        for (int i = 0; i < n; i++) y[i] = res[i].re;
    } else {
        for (int j = 0; j < n; j++) { buf[j].re = x[j]; buf[2 * (n - 1) - j].re = x[j]; }
        for (int i = 2 * (n - 1) + 1; i < L; i++) { buf[i].re = buf[L - i].re; buf[i].im = -buf[L - i].im; }
        plan_exec(pl, buf, 1, res, scratch);
        for (int i = 0; i < n; i++) y[i] = res[2 * i].re;
    }
}

/* One DCT/DST line through the extended 4N signal exactly as
 * include/heffte_r2r_executor.h:43-134 defines it (pre-process, length-4N
 * real FFT, post-process).  kind 0 = cos, 1 = sin, 2 = cos1 (DCT-I). */
static void r2r_line(const fft_plan *p4, int kind, int sign, int n, const double *x, double *y,
                     cplx *buf, cplx *res, cplx *scratch) {
    int L = 4 * n;
    if (kind == 2) {
        cos1_line(p4, sign, n, x, y, buf, res, scratch);
        return;
    }
    if (sign < 0) {
        /* forward: real extended signal, FFT, post */
        for (int i = 0; i < L; i++) { buf[i].re = 0.0; buf[i].im = 0.0; }
        if (kind == 0) {
            /* cpu_cos_pre_pos_processor::pre_forward */
            for (int i = 0; i < n; i++) { buf[2 * i].re = 0.0; buf[2 * i + 1].re = x[i]; }
            buf[2 * n].re = 0.0;
            for (int i = 0; i < 2 * n; i++) buf[4 * n - i].re = buf[i].re;
        } else {
            /* cpu_sin_pre_pos_processor::pre_forward */
            for (int i = 0; i < n; i++) { buf[2 * i].re = 0.0; buf[2 * i + 1].re = x[i]; }
            buf[2 * n].re = 0.0;
            for (int i = 0; i < n; i++) { buf[4 * n - 2 * i].re = 0.0; buf[4 * n - 2 * i - 1].re = -x[i]; }
        }
        plan_exec(p4, buf, 1, res, scratch);
        if (kind == 0) for (int i = 0; i < n; i++) y[i] = res[i].re;
        else for (int i = 0; i < n; i++) y[i] = -res[i + 1].im;
    } else {
        /* backward: half spectrum of 2n+1 bins, Hermitian fill, inverse FFT */
        cplx *h = buf; /* uses first 2n+1 entries, then expanded in res */
        for (int i = 0; i <= 2 * n; i++) { h[i].re = 0.0; h[i].im = 0.0; }
        if (kind == 0) {
            /* cpu_cos_pre_pos_processor::pre_backward */
            for (int i = 0; i < n; i++) h[i].re = x[i];
            h[n].re = 0.0;
            int idx = n - 1;
            for (int i = n + 1; i < 2 * n + 1; i++) { h[i].re = -x[idx]; idx--; }
        } else {
            /* cpu_sin_pre_pos_processor::pre_backward */
            h[0].re = 0.0; h[0].im = 0.0;
            for (int i = 0; i < n; i++) { h[i + 1].re = 0.0; h[i + 1].im = -x[i]; }
            h[2 * n].re = 0.0; h[2 * n].im = 0.0;
            for (int i = 0; i < n - 1; i++) { h[n + i + 1].re = 0.0; h[n + i + 1].im = -x[n - i - 2]; }
        }
        for (int i = 2 * n + 1; i < L; i++) { h[i].re = h[L - i].re; h[i].im = -h[L - i].im; }
        plan_exec(p4, h, 1, res, scratch);
        /* post_backward: odd samples of the real result */
        for (int i = 0; i < n; i++) y[i] = res[2 * i + 1].re;
    }
}

void orc_r2r_1d(int kind, int sign, int n, const double *x, double *y) {
    if (n <= 0 || (kind == 2 && n < 2)) return;
    fft_plan p4;
    plan_init(&p4, r2r_ext_len(kind, n), sign);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * n + 1));
    cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * n));
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)max_factor(&p4));
    r2r_line(&p4, kind, sign, n, x, y, buf, res, scratch);
    free(buf); free(res); free(scratch);
    plan_free(&p4);
}

static int r2r_common(int kind, const orc_box *inbox, const orc_box *outbox, int sign,
                      double scale, int batch, const double *in, double *out) {
    if (!same_indexes(inbox, outbox)) return -1;
    long long size[3], si[3], so[3];
    box_sizes(inbox, size);
    if (kind == 2 && (size[0] < 2 || size[1] < 2 || size[2] < 2)) return -1;
    orc_box_strides(inbox, si);
    orc_box_strides(outbox, so);
    long long cnt = orc_box_count(inbox);
    double *canon = (double *)malloc(sizeof(double) * (size_t)(cnt > 0 ? cnt : 1));
    for (int bt = 0; bt < batch; bt++) {
        const double *src = in + bt * cnt;
        for (long long k = 0; k < size[2]; k++)
            for (long long j = 0; j < size[1]; j++)
                for (long long i = 0; i < size[0]; i++)
                    canon[(k * size[1] + j) * size[0] + i] = src[i * si[0] + j * si[1] + k * si[2]];
        long long cs[3] = {1, size[0], size[0] * size[1]};
        for (int axis = 0; axis < 3; axis++) {
            int n = (int)size[axis];
            int o1 = (axis == 0) ? 1 : 0;
            int o2 = (axis == 2) ? 1 : 2;
            long long nlines = size[o1] * size[o2];
            fft_plan p4;
            plan_init(&p4, r2r_ext_len(kind, n), sign);
            int mf = max_factor(&p4);
#pragma omp parallel
            {
                double *x = (double *)malloc(sizeof(double) * (size_t)n);
                double *y = (double *)malloc(sizeof(double) * (size_t)n);
                cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * n + 1));
                cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * n));
                cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)mf);
#pragma omp for schedule(static)
                for (long long l = 0; l < nlines; l++) {
                    long long a = l % size[o1], b = l / size[o1];
                    double *base = canon + a * cs[o1] + b * cs[o2];
                    for (int i = 0; i < n; i++) x[i] = base[i * cs[axis]];
                    r2r_line(&p4, kind, sign, n, x, y, buf, res, scratch);
                    for (int i = 0; i < n; i++) base[i * cs[axis]] = y[i];
                }
                free(x); free(y); free(buf); free(res); free(scratch);
            }
            plan_free(&p4);
        }
        double *dst = out + bt * cnt;
        for (long long k = 0; k < size[2]; k++)
            for (long long j = 0; j < size[1]; j++)
                for (long long i = 0; i < size[0]; i++)
                    dst[i * so[0] + j * so[1] + k * so[2]] = scale * canon[(k * size[1] + j) * size[0] + i];
    }
    free(canon);
    return 0;
}

int orc_r2r_cos(const orc_box *inbox, const orc_box *outbox, int sign,
                double scale, int batch, const double *in, double *out) {
    return r2r_common(0, inbox, outbox, sign, scale, batch, in, out);
}

int orc_r2r_sin(const orc_box *inbox, const orc_box *outbox, int sign,
                double scale, int batch, const double *in, double *out) {
    return r2r_common(1, inbox, outbox, sign, scale, batch, in, out);
}

int orc_r2r_cos1(const orc_box *inbox, const orc_box *outbox, int sign,
                 double scale, int batch, const double *in, double *out) {
    return r2r_common(2, inbox, outbox, sign, scale, batch, in, out);
}
