// Device-side building blocks for the gfx950 FFT passes: complex arithmetic on
// float2/double2 and the forward radix-R DFT butterflies.  The backward
// transform is obtained by conjugating the input on load and the output on
// store, so only the forward butterflies (w = exp(-2 pi i / R)) exist.
#pragma once
#include <hip/hip_runtime.h>

namespace hmi {

template <typename R> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };

// Streaming (non-temporal) global accesses for the pass data: every element is read once and
// written once per pass, so the data should not displace the cached twiddle tables.
typedef float hmi_f2v __attribute__((ext_vector_type(2)));
typedef double hmi_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2 *p) {
    const hmi_f2v v = __builtin_nontemporal_load(reinterpret_cast<const hmi_f2v *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_stream(const double2 *p) {
    const hmi_d2v v = __builtin_nontemporal_load(reinterpret_cast<const hmi_d2v *>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float ld_stream(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(float2 *p, float2 v) {
    const hmi_f2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<hmi_f2v *>(p));
}
__device__ __forceinline__ void st_stream(double2 *p, double2 v) {
    const hmi_d2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<hmi_d2v *>(p));
}
__device__ __forceinline__ void st_stream(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }
// Numbering of the tiles of a pass whose load or store walks the column axis in chunks of T
// lines.  Workgroups go to the 8 XCDs round robin (workgroup i to XCD i mod 8), each XCD with
// an L2 of its own; K tiles that are neighbours along the column axis go to K workgroups of
// one XCD dispatched one after another: workgroup 8 K g + x + 8 h (XCD x, h < K) takes tile
// 8 K g + K x + h, and the last, incomplete group of 8 K workgroups keeps its own numbering.
// The passes choose K so that one XCD writes (or reads) 1 KiB runs of each row: K = 8 for
// 128-byte chunks (512^3 complex<double>, dominant pass 801 -> 750 us; complex<float> 397 ->
// 390 us; r2c 1024x1024x512 backward 2.97 -> 2.79 ms), and for chunks of 32 bytes or less the
// K = 128 / chunk tiles that share a cache line (4096-point transposing pass 1590 -> 1450 us);
// 64-byte chunks group in pairs only where a chunk straddles two cache lines (a line or
// third-axis stride of the chunked side that is not a multiple of 128 bytes; plain 64-byte
// chunks in pairs: 2048 points complex<float> unchanged, complex<double> 11 % slower).
template <int K> __device__ __forceinline__ unsigned xcd_grouped_tile(unsigned b, unsigned nb) {
    static_assert(K >= 1, "a group holds at least one tile per XCD");
    constexpr unsigned G = 8 * K;
    const unsigned g = b / G;
    if (K == 1 || (g + 1) * G > nb) return b;
    const unsigned r = b % G;
    return g * G + K * (r % 8) + r / 8;
}

// NT selects the streaming form; passes whose output fits the 256 MB Infinity Cache keep the
// default policy so the next pass can read it from there.
template <bool NT, typename T> __device__ __forceinline__ T ldg(const T *p) {
    if constexpr (NT) return ld_stream(p);
    else return *p;
}
template <bool NT, typename T> __device__ __forceinline__ void stg(T *p, T v) {
    if constexpr (NT) st_stream(p, v);
    else *p = v;
}

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) re, decltype(C::x) im) {
    C c;
    c.x = re;
    c.y = im;
    return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (-i)
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return cmk<C>(a.y, -a.x); }
// a * (+i)
template <typename C> __device__ __forceinline__ C cmul_pi(C a) { return cmk<C>(-a.y, a.x); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }

// exp(-2 pi i k / 16) for the constants used by radix 8 and 16.
#define HMI_C16_1 0.92387953251128675613
#define HMI_S16_1 0.38268343236508977173
#define HMI_SQRT_HALF 0.70710678118654752440

template <int R> struct dft;

template <> struct dft<2> {
    template <typename C> __device__ __forceinline__ static void run(C *x) {
        C a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};

// y1 = (x0 - x2) - i (x1 - x3), y3 = (x0 - x2) + i (x1 - x3)
template <> struct dft<4> {
    template <typename C> __device__ __forceinline__ static void run4(C &x0, C &x1, C &x2, C &x3) {
        C t0 = cadd(x0, x2), t1 = csub(x0, x2);
        C t2 = cadd(x1, x3), t3 = cmul_mi(csub(x1, x3));
        x0 = cadd(t0, t2);
        x2 = csub(t0, t2);
        x1 = cadd(t1, t3);
        x3 = csub(t1, t3);
    }
    template <typename C> __device__ __forceinline__ static void run(C *x) { run4(x[0], x[1], x[2], x[3]); }
};

// 8 = 2 x 4: a_j = x_j +- x_{j+4}; the odd half is twiddled by w8^j.
template <> struct dft<8> {
    template <typename C> __device__ __forceinline__ static void run(C *x) {
        using T = decltype(C::x);
        const T h = (T)HMI_SQRT_HALF;
        C a0 = cadd(x[0], x[4]), b0 = csub(x[0], x[4]);
        C a1 = cadd(x[1], x[5]), b1 = csub(x[1], x[5]);
        C a2 = cadd(x[2], x[6]), b2 = csub(x[2], x[6]);
        C a3 = cadd(x[3], x[7]), b3 = csub(x[3], x[7]);
        // b1 *= w8 = (1 - i)/sqrt2 ; b2 *= -i ; b3 *= w8^3 = (-1 - i)/sqrt2
        b1 = cmk<C>((b1.x + b1.y) * h, (b1.y - b1.x) * h);
        b2 = cmul_mi(b2);
        b3 = cmk<C>((b3.y - b3.x) * h, -(b3.x + b3.y) * h);
        dft<4>::run4(a0, a1, a2, a3);
        dft<4>::run4(b0, b1, b2, b3);
        x[0] = a0; x[2] = a1; x[4] = a2; x[6] = a3;
        x[1] = b0; x[3] = b1; x[5] = b2; x[7] = b3;
    }
};

// 16 = 4 x 4: DFT4 over h of x[p + 4h], twiddle by w16^{p m}, DFT4 over p.
template <> struct dft<16> {
    template <typename C> __device__ __forceinline__ static void run(C *x) {
        using T = decltype(C::x);
        const T c1 = (T)HMI_C16_1, s1 = (T)HMI_S16_1, h = (T)HMI_SQRT_HALF;
        for (int p = 0; p < 4; p++) dft<4>::run4(x[p], x[p + 4], x[p + 8], x[p + 12]);
        // row p=1: m=1..3 -> w16^1, w16^2, w16^3
        x[5] = cmul(x[5], cmk<C>(c1, -s1));
        x[9] = cmk<C>((x[9].x + x[9].y) * h, (x[9].y - x[9].x) * h);
        x[13] = cmul(x[13], cmk<C>(s1, -c1));
        // row p=2: w16^2, w16^4, w16^6
        x[6] = cmk<C>((x[6].x + x[6].y) * h, (x[6].y - x[6].x) * h);
        x[10] = cmul_mi(x[10]);
        x[14] = cmk<C>((x[14].y - x[14].x) * h, -(x[14].x + x[14].y) * h);
        // row p=3: w16^3, w16^6, w16^9
        x[7] = cmul(x[7], cmk<C>(s1, -c1));
        x[11] = cmk<C>((x[11].y - x[11].x) * h, -(x[11].x + x[11].y) * h);
        x[15] = cmul(x[15], cmk<C>(-c1, s1));
        C y[16];
        for (int m = 0; m < 4; m++) {
            C a = x[4 * m], b = x[4 * m + 1], c = x[4 * m + 2], d = x[4 * m + 3];
            dft<4>::run4(a, b, c, d);
            y[m] = a; y[m + 4] = b; y[m + 8] = c; y[m + 12] = d;
        }
        for (int q = 0; q < 16; q++) x[q] = y[q];
    }
};

// Odd prime radices: y_q = x0 + sum_{p=1}^{(R-1)/2} (x_p + x_{R-p}) cos(2 pi p q / R)
//                               - i (x_p - x_{R-p}) sin(2 pi p q / R).
template <int R> struct odd_consts;
template <> struct odd_consts<3> {
    __device__ static constexpr double c(int k) { return k == 1 ? -0.5 : 0.0; }
    __device__ static constexpr double s(int k) { return k == 1 ? 0.8660254037844386467637 : 0.0; }
};
template <> struct odd_consts<5> {
    __device__ static constexpr double c(int k) {
        return k == 1 ? 0.3090169943749474241023 : -0.8090169943749474241023;
    }
    __device__ static constexpr double s(int k) {
        return k == 1 ? 0.9510565162951535721164 : 0.5877852522924731291687;
    }
};
template <> struct odd_consts<7> {
    __device__ static constexpr double c(int k) {
        return k == 1 ? 0.623489801858733530525 : (k == 2 ? -0.2225209339563144042889 : -0.9009688679024191262361);
    }
    __device__ static constexpr double s(int k) {
        return k == 1 ? 0.7818314824680298087084 : (k == 2 ? 0.9749279121818236070181 : 0.4338837391175581204758);
    }
};
template <> struct odd_consts<11> {
    __device__ static constexpr double c(int k) {
        return k == 1 ? 0.8412535328311811688618 : k == 2 ? 0.4154150130018864255293 :
               k == 3 ? -0.1423148382732851404438 : k == 4 ? -0.6548607339452850640569 : -0.9594929736144973898904;
    }
    __device__ static constexpr double s(int k) {
        return k == 1 ? 0.5406408174555975821076 : k == 2 ? 0.9096319953545183714117 :
               k == 3 ? 0.9898214418809327323761 : k == 4 ? 0.755749574354258283774 : 0.2817325568414296977114;
    }
};
template <> struct odd_consts<13> {
    __device__ static constexpr double c(int k) {
        return k == 1 ? 0.8854560256532098959004 : k == 2 ? 0.5680647467311558025118 :
               k == 3 ? 0.1205366802553230533491 : k == 4 ? -0.3546048870425356259696 :
               k == 5 ? -0.7485107481711010986346 : -0.970941817426052027157;
    }
    __device__ static constexpr double s(int k) {
        return k == 1 ? 0.464723172043768545656 : k == 2 ? 0.8229838658936563945796 :
               k == 3 ? 0.9927088740980539928008 : k == 4 ? 0.9350162426854148234398 :
               k == 5 ? 0.6631226582407952023768 : 0.2393156642875577671488;
    }
};

template <int R> struct dft_odd {
    template <typename C> __device__ __forceinline__ static void run(C *x) {
        using T = decltype(C::x);
        constexpr int H = (R - 1) / 2;
        C sp[H], dm[H];
        C y0 = x[0];
#pragma unroll
        for (int p = 1; p <= H; p++) {
            sp[p - 1] = cadd(x[p], x[R - p]);
            dm[p - 1] = csub(x[p], x[R - p]);
            y0 = cadd(y0, sp[p - 1]);
        }
        C y[R];
        y[0] = y0;
#pragma unroll
        for (int q = 1; q <= H; q++) {
            C re = x[0];
            T im_x = 0, im_y = 0;  // accumulates sum sin * dm (as a complex)
#pragma unroll
            for (int p = 1; p <= H; p++) {
                const int e = (p * q) % R;
                // cos/sin of 2 pi e / R with e folded into [1, H] by symmetry
                const int f = e <= H ? e : R - e;
                const T c = (T)odd_consts<R>::c(f);
                const T s = (T)(e <= H ? odd_consts<R>::s(f) : -odd_consts<R>::s(f));
                re.x += c * sp[p - 1].x;
                re.y += c * sp[p - 1].y;
                im_x += s * dm[p - 1].x;
                im_y += s * dm[p - 1].y;
            }
            // y_q = re - i * (im), y_{R-q} = re + i * (im)
            y[q] = cmk<C>(re.x + im_y, re.y - im_x);
            y[R - q] = cmk<C>(re.x - im_y, re.y + im_x);
        }
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = y[q];
    }
};
template <> struct dft<3> : dft_odd<3> {};
template <> struct dft<5> : dft_odd<5> {};
template <> struct dft<7> : dft_odd<7> {};
template <> struct dft<11> : dft_odd<11> {};
template <> struct dft<13> : dft_odd<13> {};

}  // namespace hmi
