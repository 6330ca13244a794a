// gfx950 FFT pass kernels.
//
//  * fft_uniform<...>: compile-time radix plan, every thread owns E elements of
//    one line for the whole transform; stages exchange data through a padded
//    LDS tile.  The thread->element mapping used for the global load and the
//    one used for the global store are chosen independently (walk the FFT axis
//    or walk the column axis), which is how a layout permutation is fused into
//    a pass with coalesced HBM traffic on both sides.
//  * fft_generic<R>: runtime radix plan (any length whose tile fits in LDS,
//    any prime factor), Stockham ping-pong in LDS.
//  * permute_kernel: heFFTe's in-node reshape (transpose_unpacker,
//    src/heffte_backend_rocm.hip.cpp:91-134) as an LDS-tiled permutation.
//
// Stockham autosort, stage s with radix R and Ls = product of earlier radices:
//   butterfly b in [0, n/R): inputs x[b + p n/R], twiddles w_{Ls R}^{p (b mod Ls)},
//   outputs to (b / Ls) Ls R + (b mod Ls) + q Ls.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "hmi_device.h"
#include "hmi_kernels.h"

namespace hmi {

template <int... Rs> struct radices {
    static constexpr int count = sizeof...(Rs);
    static constexpr int at(int i) {
        constexpr int a[] = {Rs...};
        return a[i];
    }
    static constexpr int ls(int s) {
        int p = 1;
        for (int i = 0; i < s; i++) p *= at(i);
        return p;
    }
    static constexpr int twoff(int s) {
        int o = 0;
        for (int i = 0; i < s; i++) o += ls(i) * (at(i) - 1);
        return o;
    }
    static constexpr int n() { return ls(count); }
};

template <int PS> __device__ __forceinline__ int padj(int j) { return j + (j >> PS); }

// LDS tile of the uniform pass: element j of line l.  Rows of n elements padded by one element
// per 2^ps (+1), or, for the 512-point complex<float> tiles (8 elements per thread), rows of
// n + 32 elements with the low 5 bits of j mapped linearly (j0 -> 0x11, j1..j3 -> themselves, j4 -> 0x12),
// XOR-ed with 0x1 * j5 and 0xC * j6 and with the line's low 4 bits.  Measured on gfx950 (a probe
// of strides and row patterns under rocprofv3's SQ_LDS_BANK_CONFLICT): ds_read_b64 serves 32
// lanes per pass over 64 four-byte banks, ds_write_b64 16 lanes per pass over 32 banks.  Under
// that rule every exchange of these tiles is conflict-free: stage writes of one line per wave
// (positions 8t + q and 64a + b + 8q), reads of positions t + 64e, and, in the mapping where
// consecutive lanes walk 16 lines, writes of 16 lines at one position and reads of 16 lines at
// two adjacent positions (the padded rows keep 1 conflict cycle per read there).
template <int N, int E, int PS, int BYTES> struct lds_map {
    static constexpr bool swz = (N == 512 && E == 8 && BYTES == 8);
    // (swizzled rows keep 32 more slots: the c2r load parks position N there)
    static constexpr int RS = swz ? N + 32 : N + (N >> PS) + 1;
    __device__ static __forceinline__ int idx(int l, int j) {
        if constexpr (swz) {
            const int lo = ((j & 0xE) ^ ((j & 1) * 0x11) ^ (((j >> 4) & 1) * 0x12)) ^ ((j >> 5) & 1) ^
                           (((j >> 6) & 1) * 0xC) ^ (l & 15);
            return l * RS + ((j & ~31) | lo);
        } else {
            return l * RS + padj<PS>(j);
        }
    }
};

// Row stride (complex elements of `bytes` bytes) of the LDS tile of an n-point line with one pad
// element per 2^ps, for tiles of t lines.
constexpr int tile_row_stride(int n, int ps, int t, int bytes) { return n + (n >> ps) + 1 + 0 * t * bytes; }

template <int S, typename RL, int E, int P, int RS, int PS, typename C>
__device__ __forceinline__ void run_stages(C (&v)[E], C *lds, const C *__restrict__ tw, int l, int t, int l2,
                                           int t2) {
    using LM = lds_map<RL::n(), E, PS, sizeof(C)>;
    constexpr int R = RL::at(S);
    constexpr int Ls = RL::ls(S);
    constexpr int NB = E / R;
    constexpr bool last = (S + 1 == RL::count);
    constexpr int TWO = RL::twoff(S);
    static_assert(E % R == 0, "every radix must divide the elements per thread");
#pragma unroll
    for (int r = 0; r < NB; r++) {
        const int b = t + P * r;
        const int k = b % Ls;
        C x[R];
#pragma unroll
        for (int p = 0; p < R; p++) x[p] = v[r + NB * p];
        if constexpr (Ls > 1) {
            const C *w = tw + TWO + k * (R - 1);
#pragma unroll
            for (int p = 1; p < R; p++) x[p] = cmul(x[p], w[p - 1]);
        }
        dft<R>::run(x);
        if constexpr (last) {
#pragma unroll
            for (int q = 0; q < R; q++) v[r + NB * q] = x[q];
        } else {
            const int base = (b / Ls) * Ls * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) lds[LM::idx(l, base + q * Ls)] = x[q];
        }
    }
    if constexpr (!last) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = lds[LM::idx(l2, t2 + P * e)];
        __syncthreads();
        run_stages<S + 1, RL, E, P, RS, PS>(v, lds, tw, l2, t2, l2, t2);
    }
}

// Offset of line L (column-fastest numbering of column x b1 x batch).
__device__ __forceinline__ long long line_offset(unsigned L, const pass_args &a, bool out) {
    const unsigned q = L / (unsigned)a.ncol;
    const unsigned c = L - q * (unsigned)a.ncol;
    const unsigned b2 = q / (unsigned)a.nb1;
    const unsigned b1 = q - b2 * (unsigned)a.nb1;
    return out ? (long long)c * a.out_sc + (long long)b1 * a.out_sb1 + (long long)b2 * a.out_sb2
               : (long long)c * a.in_sc + (long long)b1 * a.in_sb1 + (long long)b2 * a.in_sb2;
}

// Index of line L along dimension dim of its numbering (0 column, 1 b1, 2 batch).
__device__ __forceinline__ unsigned line_index(unsigned L, const pass_args &a, int dim) {
    const unsigned q = L / (unsigned)a.ncol;
    if (dim == 0) return L - q * (unsigned)a.ncol;
    const unsigned b2 = q / (unsigned)a.nb1;
    return dim == 1 ? q - b2 * (unsigned)a.nb1 : b2;
}

// One tile (T lines) of an n-point pass; the block has exactly (n / E) * T threads.
template <typename R, typename RL, int E, int T, int PS, int NTM = 0>
__device__ __forceinline__ void uniform_tile(const pass_args &a, unsigned tile, typename cvec<R>::type *lds) {
    using C = typename cvec<R>::type;
    constexpr int N = RL::n();
    constexpr int P = N / E;
    using LM = lds_map<N, E, PS, sizeof(C)>;
    constexpr int RS = LM::RS;
    constexpr bool NTL = (NTM & 2) != 0, NTS = (NTM & 1) != 0;

    const int tid = threadIdx.x;
    int l, t, l2, t2;
    if (a.load_colfast) { l = tid % T; t = tid / T; } else { l = tid / P; t = tid % P; }
    if (a.store_colfast) { l2 = tid % T; t2 = tid / T; } else { l2 = tid / P; t2 = tid % P; }
    // a block may carry more threads than P * T (batched kernel): the extra threads take
    // part in the barriers and work on a scratch LDS row T that the caller provides
    const bool active = tid < P * T;
    if (!active) { l = l2 = T; t = t2 = 0; }

    const R sg = a.conj ? R(-1) : R(1);
    const C *__restrict__ tw = static_cast<const C *>(a.tw);
    // post-twiddles w_{2N}^k of the real-data passes follow the stage twiddles
    const C *__restrict__ tw_half = tw + RL::twoff(RL::count);
    C v[E];
    {
        const unsigned L = tile * T + l;
        const bool valid = active && L < (unsigned)a.nlines;
        const long long base = valid ? line_offset(L, a, false) : 0;
        if (a.mode == kModeR2C) {
            // real line of 2N values read as N complex pairs z[m] = x[2m] + i x[2m+1]; a contiguous
            // line (even offsets: every other stride is a multiple of 2N) of an array aligned to
            // the pair size is read pairwise
            const R *__restrict__ src = static_cast<const R *>(a.in);
            if (a.in_sa == 1 && (reinterpret_cast<size_t>(a.in) & (sizeof(C) - 1)) == 0) {
#pragma unroll
                for (int e = 0; e < E; e++)
                    v[e] = valid ? ldg<NTL>(reinterpret_cast<const C *>(src + base) + (t + P * e)) : cmk<C>(R(0), R(0));
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const long long m2 = 2LL * (t + P * e);
                    v[e] = valid ? cmk<C>(ldg<NTL>(&src[base + m2 * a.in_sa]), ldg<NTL>(&src[base + (m2 + 1) * a.in_sa]))
                                 : cmk<C>(R(0), R(0));
                }
            }
        } else if (a.mode == kModeC2R) {
            // half spectrum X[0..N] of a real line of 2N values -> Z[k] = (X[k] + conj X[N-k])
            //   + i w^{-k} (X[k] - conj X[N-k]), the spectrum of z[m] = x[2m] + i x[2m+1];
            // the imaginary parts of X[0] and X[N] do not contribute to a real signal
            const C *__restrict__ src = static_cast<const C *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++)
                lds[LM::idx(l, t + P * e)] =
                    valid ? ldg<NTL>(&src[base + (long long)(t + P * e) * a.in_sa]) : cmk<C>(R(0), R(0));
            if (t == 0) lds[LM::idx(l, N)] = valid ? ldg<NTL>(&src[base + (long long)N * a.in_sa]) : cmk<C>(R(0), R(0));
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + P * e;
                C xk = lds[LM::idx(l, k)];
                C xm = lds[LM::idx(l, N - k)];
                if (k == 0) { xk.y = R(0); xm.y = R(0); }
                const C w = tw_half[k];  // exp(-i pi k / N); w^{-k} = conj(w)
                const C sum = cmk<C>(xk.x + xm.x, xk.y - xm.y);
                const C dif = cmk<C>(xk.x - xm.x, xk.y + xm.y);
                const C wd = cmk<C>(w.x * dif.x + w.y * dif.y, w.x * dif.y - w.y * dif.x);  // conj(w) * dif
                const C z = cmk<C>(sum.x - wd.y, sum.y + wd.x);                             // sum + i wd
                v[e] = cmk<C>(z.x, -z.y);  // inverse DFT through the forward stages: conjugate in and out
            }
            __syncthreads();
        } else if (a.mode == kModeDCT2 || a.mode == kModeDST2) {
            const bool sine = (a.mode == kModeDST2);
            const R *__restrict__ src = static_cast<const R *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int n = t + P * e;
                const int j = (n < (N + 1) / 2) ? 2 * n : 2 * (N - 1 - n) + 1;
                R x = valid ? ldg<NTL>(&src[base + (long long)j * a.in_sa]) : R(0);
                if (sine && (j & 1)) x = -x;
                v[e] = cmk<C>(x, R(0));
            }
        } else if (a.mode == kModeDCT3 || a.mode == kModeDST3) {
            const bool sine = (a.mode == kModeDST3);
            const R *__restrict__ src = static_cast<const R *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + P * e;
                const int jk = sine ? N - 1 - k : k;
                const int jm = sine ? k - 1 : N - k;
                const R xk = valid ? ldg<NTL>(&src[base + (long long)jk * a.in_sa]) : R(0);
                const R xm = (valid && k > 0) ? ldg<NTL>(&src[base + (long long)jm * a.in_sa]) : R(0);
                const C w = tw_half[k];  // exp(-i pi k / 2N)
                // W = conj(w) (xk - i xm); the inverse DFT runs as conj(DFT(conj W))
                const C W = cmk<C>(w.x * xk - w.y * xm, -w.x * xm - w.y * xk);
                v[e] = cmk<C>(W.x, -W.y);
            }
        } else if (a.mode == kModeDCT1 || a.mode == kModeDCT1B) {
            // even extension of N = F/2 + 1 reals to the F-point line
            const R *__restrict__ src = static_cast<const R *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int j = t + P * e;
                const int jj = (j <= N / 2) ? j : N - j;
                v[e] = cmk<C>(valid ? ldg<NTL>(&src[base + (long long)jj * a.in_sa]) : R(0), R(0));
            }
        } else if (a.in_real) {
            const R *__restrict__ src = static_cast<const R *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++)
                v[e] = valid ? cmk<C>(ldg<NTL>(&src[base + (long long)(t + P * e) * a.in_sa]), R(0)) : cmk<C>(R(0), R(0));
        } else if (a.herm_load) {
            const C *__restrict__ src = static_cast<const C *>(a.in);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int pos = t + P * e;
                const bool mirror = pos > N / 2;
                C x = valid ? ldg<NTL>(&src[base + (long long)(mirror ? N - pos : pos) * a.in_sa]) : cmk<C>(R(0), R(0));
                v[e] = cmk<C>(x.x, (mirror ? -sg : sg) * x.y);
            }
        } else {
            const C *__restrict__ src = static_cast<const C *>(a.in);
            if (a.mul == a.in) {
                // squared input (convolution of an array with itself): one load per element
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const long long off = base + (long long)(t + P * e) * a.in_sa;
                    C y = valid ? ldg<NTL>(&src[off]) : cmk<C>(R(0), R(0));
                    C x = cmul(y, y);
                    v[e] = cmk<C>(x.x, sg * x.y);
                }
            } else if (a.mul) {
                // pointwise product with a second array of the same layout (fused convolution)
                const C *__restrict__ mul = static_cast<const C *>(a.mul);
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const long long off = base + (long long)(t + P * e) * a.in_sa;
                    C x = valid ? cmul(ldg<NTL>(&src[off]), ldg<NTL>(&mul[off])) : cmk<C>(R(0), R(0));
                    v[e] = cmk<C>(x.x, sg * x.y);
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) {
                    C x = valid ? ldg<NTL>(&src[base + (long long)(t + P * e) * a.in_sa]) : cmk<C>(R(0), R(0));
                    v[e] = cmk<C>(x.x, sg * x.y);
                }
            }
        }
    }

    run_stages<0, RL, E, P, RS, PS>(v, lds, tw, l, t, l2, t2);
    if constexpr (RL::count == 1) {
        if (a.load_colfast != a.store_colfast) {
#pragma unroll
            for (int e = 0; e < E; e++) lds[LM::idx(l, t + P * e)] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = lds[LM::idx(l2, t2 + P * e)];
            __syncthreads();
        }
    }

    const unsigned L2 = tile * T + l2;
    const bool svalid = active && L2 < (unsigned)a.nlines;
    const R s = (R)a.scale;
    const long long obase = svalid ? line_offset(L2, a, true) : 0;
    if (a.mode == kModeR2C) {
        // Z = DFT_N(z); X[k] = E[k] + w^k O[k], E = (Z[k] + conj Z[N-k]) / 2, O = -i (Z[k] - conj Z[N-k]) / 2
#pragma unroll
        for (int e = 0; e < E; e++) lds[LM::idx(l2, t2 + P * e)] = v[e];
        __syncthreads();
        C *__restrict__ dst = static_cast<C *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t2 + P * e;
            const C zk = v[e];
            const C zm = lds[LM::idx(l2, k == 0 ? 0 : N - k)];
            const C ev = cmk<C>(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y - zm.y));
            const C od = cmk<C>(R(0.5) * (zk.y + zm.y), R(-0.5) * (zk.x - zm.x));
            const C w = tw_half[k];
            const C x = cadd(ev, cmul(w, od));
            if (svalid) stg<NTS>(&dst[obase + (long long)k * a.out_sa], cmk<C>(x.x * s, x.y * s));
            if (k == 0 && svalid) stg<NTS>(&dst[obase + (long long)N * a.out_sa], cmk<C>((ev.x - od.x) * s, (ev.y - od.y) * s));
        }
    } else if (!svalid) {
        // column beyond the box: nothing to store
    } else if (a.mode == kModeDCT2 || a.mode == kModeDST2) {
        const bool sine = (a.mode == kModeDST2);
        R *__restrict__ dst = static_cast<R *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t2 + P * e;
            const C w = tw_half[k];
            const R y = R(2) * (w.x * v[e].x - w.y * v[e].y);
            stg<NTS>(&dst[obase + (long long)(sine ? N - 1 - k : k) * a.out_sa], y * s);
        }
    } else if (a.mode == kModeDCT3 || a.mode == kModeDST3) {
        const bool sine = (a.mode == kModeDST3);
        R *__restrict__ dst = static_cast<R *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int m = t2 + P * e;
            const int j = (m < (N + 1) / 2) ? 2 * m : 2 * (N - 1 - m) + 1;
            R y = R(2) * v[e].x;
            if (sine && (j & 1)) y = -y;
            stg<NTS>(&dst[obase + (long long)j * a.out_sa], y * s);
        }
    } else if (a.mode == kModeDCT1 || a.mode == kModeDCT1B) {
        R *__restrict__ dst = static_cast<R *>(a.out);
        const R f = (a.mode == kModeDCT1B) ? R(2) * s : s;
#pragma unroll
        for (int e = 0; e < E; e++)
            if (t2 + P * e < a.store_len) stg<NTS>(&dst[obase + (long long)(t2 + P * e) * a.out_sa], v[e].x * f);
    } else if (a.mode == kModeC2R) {
        R *__restrict__ dst = static_cast<R *>(a.out);
        if (a.out_sa == 1 && (reinterpret_cast<size_t>(a.out) & (sizeof(C) - 1)) == 0) {
#pragma unroll
            for (int e = 0; e < E; e++)
                stg<NTS>(reinterpret_cast<C *>(dst + obase) + (t2 + P * e), cmk<C>(v[e].x * s, -v[e].y * s));
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const long long m2 = 2LL * (t2 + P * e);
                stg<NTS>(&dst[obase + m2 * a.out_sa], v[e].x * s);
                stg<NTS>(&dst[obase + (m2 + 1) * a.out_sa], -v[e].y * s);
            }
        }
    } else if (a.out_real) {
        R *__restrict__ dst = static_cast<R *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++) stg<NTS>(&dst[obase + (long long)(t2 + P * e) * a.out_sa], v[e].x * s);
    } else if (a.store_len < N) {
        C *__restrict__ dst = static_cast<C *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++)
            if (t2 + P * e < a.store_len)
                stg<NTS>(&dst[obase + (long long)(t2 + P * e) * a.out_sa], cmk<C>(v[e].x * s, sg * v[e].y * s));
    } else if (a.tw4) {
        // first step of a split axis: twiddle w_n4^(j k) before the (conjugating) store
        C *__restrict__ dst = static_cast<C *>(a.out);
        const C *__restrict__ w4 = static_cast<const C *>(a.tw4);
        const unsigned j = line_index(L2, a, a.tw4_dim);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const unsigned k = t2 + P * e;
            const C y = cmul(v[e], w4[j * k]);
            stg<NTS>(&dst[obase + (long long)k * a.out_sa], cmk<C>(y.x * s, sg * y.y * s));
        }
    } else {
        C *__restrict__ dst = static_cast<C *>(a.out);
#pragma unroll
        for (int e = 0; e < E; e++)
            stg<NTS>(&dst[obase + (long long)(t2 + P * e) * a.out_sa], cmk<C>(v[e].x * s, sg * v[e].y * s));
    }
}

// NTM: bit 0 = non-temporal stores, bit 1 = non-temporal loads of the pass data
template <typename R, typename RL, int E, int T, int PS, int NTM>
__global__ void __launch_bounds__((RL::n() / E) * T) fft_uniform(pass_args a) {
    using C = typename cvec<R>::type;
    constexpr int RS = lds_map<RL::n(), E, PS, sizeof(C)>::RS;
    constexpr int CHUNK = T * (int)sizeof(C);
    __shared__ C lds[T * RS];
    unsigned tile = blockIdx.x;
    if (a.load_colfast || a.store_colfast) {
        if constexpr (CHUNK >= 128 || CHUNK <= 32) {
            tile = xcd_grouped_tile<(CHUNK >= 1024 ? 1 : (CHUNK >= 128 ? 1024 : 128) / CHUNK)>(blockIdx.x, gridDim.x);
        } else {
            const long long es = (long long)sizeof(C);
            const bool straddle =
                (a.load_colfast && ((a.in_sa * es) % 128 || (a.nb1 > 1 && (a.in_sb1 * es) % 128))) ||
                (a.store_colfast && ((a.out_sa * es) % 128 || (a.nb1 > 1 && (a.out_sb1 * es) % 128)));
            if (straddle) tile = xcd_grouped_tile<2>(blockIdx.x, gridDim.x);
        }
    } else if constexpr (sizeof(C) == 16) {
        // contiguous lines on both sides: complex<double> tiles in groups of 8 per XCD as well
        // (512^3 y-pass 748 -> 703 us; for complex<float> the same grouping was slower, 355 -> 361 us)
        tile = xcd_grouped_tile<8>(blockIdx.x, gridDim.x);
    }
    uniform_tile<R, RL, E, T, PS, NTM>(a, tile, lds);
}

// ---------------------------------------------------------------- generic pass

template <int R, typename C>
__device__ __forceinline__ void generic_bfly(const C *sp, int istride, C *dp, int ostride, const C *w, bool twiddle) {
    C x[R];
#pragma unroll
    for (int p = 0; p < R; p++) x[p] = sp[p * istride];
    if (twiddle) {
#pragma unroll
        for (int p = 1; p < R; p++) x[p] = cmul(x[p], w[p - 1]);
    }
    dft<R>::run(x);
#pragma unroll
    for (int q = 0; q < R; q++) dp[q * ostride] = x[q];
}

// One tile (g.lines lines) of an n-point pass with a runtime radix plan; any block size.
template <typename R>
__device__ void generic_tile(const pass_args &a, const generic_plan &g, unsigned tile,
                             typename cvec<R>::type *smem) {
    using C = typename cvec<R>::type;
    const int N = g.n, T = g.lines;
    C *buf0 = smem;
    C *buf1 = buf0 + (long long)T * N;

    const int tid = threadIdx.x, nth = blockDim.x;
    const R sg = a.conj ? R(-1) : R(1);

    for (int idx = tid; idx < T * N; idx += nth) {
        int l, j;
        if (a.load_colfast) { l = idx % T; j = idx / T; } else { l = idx / N; j = idx % N; }
        C x = cmk<C>(R(0), R(0));
        const unsigned L = tile * T + l;
        if (L < (unsigned)a.nlines) {
            const long long lbase = line_offset(L, a, false);
            const long long off = lbase + (long long)j * a.in_sa;
            const R *rsrc = static_cast<const R *>(a.in);
            if (a.mode == kModeDCT2 || a.mode == kModeDST2) {
                const int jj = (j < (N + 1) / 2) ? 2 * j : 2 * (N - 1 - j) + 1;
                x.x = rsrc[lbase + (long long)jj * a.in_sa];
                if (a.mode == kModeDST2 && (jj & 1)) x.x = -x.x;
            } else if (a.mode == kModeDCT3 || a.mode == kModeDST3) {
                const bool sine = (a.mode == kModeDST3);
                const R xk = rsrc[lbase + (long long)(sine ? N - 1 - j : j) * a.in_sa];
                const R xm = j > 0 ? rsrc[lbase + (long long)(sine ? j - 1 : N - j) * a.in_sa] : R(0);
                const C w = static_cast<const C *>(a.tw)[g.postoff + j];
                x = cmk<C>(w.x * xk - w.y * xm, w.x * xm + w.y * xk);  // conj of conj(w) (xk - i xm)
            } else if (a.mode == kModeDCT1 || a.mode == kModeDCT1B) {
                x.x = rsrc[lbase + (long long)(j <= N / 2 ? j : N - j) * a.in_sa];
            } else if (a.in_real) {
                x.x = static_cast<const R *>(a.in)[off];
            } else if (a.herm_load && j > N / 2) {
                x = static_cast<const C *>(a.in)[lbase + (long long)(N - j) * a.in_sa];
                x.y *= -sg;
            } else {
                x = static_cast<const C *>(a.in)[off];
                if (a.mul) x = cmul(x, a.mul == a.in ? x : static_cast<const C *>(a.mul)[off]);
                x.y *= sg;
            }
        }
        buf0[l * N + j] = x;
    }
    __syncthreads();

    const C *tw = static_cast<const C *>(a.tw);
    C *src = buf0, *dst = buf1;
    int Ls = 1;
    for (int s = 0; s < g.nstages; s++) {
        const int Rr = g.radix[s];
        const int nb = N / Rr;
        for (int idx = tid; idx < T * nb; idx += nth) {
            const int l = idx / nb, b = idx - (idx / nb) * nb;
            const int k = b % Ls;
            const C *sp = src + l * N + b;
            C *dp = dst + l * N + (b / Ls) * Ls * Rr + k;
            const C *w = tw + g.twoff[s] + k * (Rr - 1);
            const bool twd = (k != 0);
            switch (Rr) {
                case 2: generic_bfly<2>(sp, nb, dp, Ls, w, twd); break;
                case 3: generic_bfly<3>(sp, nb, dp, Ls, w, twd); break;
                case 4: generic_bfly<4>(sp, nb, dp, Ls, w, twd); break;
                case 5: generic_bfly<5>(sp, nb, dp, Ls, w, twd); break;
                case 7: generic_bfly<7>(sp, nb, dp, Ls, w, twd); break;
                case 8: generic_bfly<8>(sp, nb, dp, Ls, w, twd); break;
                case 11: generic_bfly<11>(sp, nb, dp, Ls, w, twd); break;
                case 13: generic_bfly<13>(sp, nb, dp, Ls, w, twd); break;
                case 16: generic_bfly<16>(sp, nb, dp, Ls, w, twd); break;
                default: {
                    // any other prime: direct DFT with the w_R^j table
                    const C *root = tw + g.rootoff[s];
                    for (int q = 0; q < Rr; q++) {
                        C acc = sp[0];
                        for (int p = 1; p < Rr; p++) {
                            C x = sp[p * nb];
                            if (twd) x = cmul(x, w[p - 1]);
                            acc = cadd(acc, cmul(x, root[(p * q) % Rr]));
                        }
                        dp[q * Ls] = acc;
                    }
                }
            }
        }
        __syncthreads();
        C *tmp = src; src = dst; dst = tmp;
        Ls *= Rr;
    }

    const R sc = (R)a.scale;
    for (int idx = tid; idx < T * N; idx += nth) {
        int l, j;
        if (a.store_colfast) { l = idx % T; j = idx / T; } else { l = idx / N; j = idx % N; }
        const unsigned L = tile * T + l;
        if (L >= (unsigned)a.nlines || j >= a.store_len) continue;
        const C y = src[l * N + j];
        const long long lobase = line_offset(L, a, true);
        const long long off = lobase + (long long)j * a.out_sa;
        if (a.mode == kModeDCT2 || a.mode == kModeDST2) {
            const C w = static_cast<const C *>(a.tw)[g.postoff + j];
            const int k = (a.mode == kModeDST2) ? N - 1 - j : j;
            static_cast<R *>(a.out)[lobase + (long long)k * a.out_sa] = R(2) * (w.x * y.x - w.y * y.y) * sc;
        } else if (a.mode == kModeDCT3 || a.mode == kModeDST3) {
            const int jj = (j < (N + 1) / 2) ? 2 * j : 2 * (N - 1 - j) + 1;
            R r = R(2) * y.x;
            if (a.mode == kModeDST3 && (jj & 1)) r = -r;
            static_cast<R *>(a.out)[lobase + (long long)jj * a.out_sa] = r * sc;
        } else if (a.mode == kModeDCT1B) {
            static_cast<R *>(a.out)[off] = R(2) * y.x * sc;
        } else if (a.out_real) static_cast<R *>(a.out)[off] = y.x * sc;
        else if (a.tw4) {
            const C z = cmul(y, static_cast<const C *>(a.tw4)[line_index(L, a, a.tw4_dim) * (unsigned)j]);
            static_cast<C *>(a.out)[off] = cmk<C>(z.x * sc, sg * z.y * sc);
        } else static_cast<C *>(a.out)[off] = cmk<C>(y.x * sc, sg * y.y * sc);
    }
}

// ------------------------------------------------------ staged compile-time tile

template <typename RL, int T, int NTH, int S, int END, typename C>
__device__ __forceinline__ void staged_stages(C *src, C *dst, const C *__restrict__ tw) {
    constexpr int Rr = RL::at(S), N = RL::n(), NB = N / Rr, Ls = RL::ls(S), TWO = RL::twoff(S);
    for (int idx = threadIdx.x; idx < T * NB; idx += NTH) {
        const int l = idx / NB, b = idx - l * NB;
        const int k = b % Ls;
        generic_bfly<Rr>(src + l * N + b, NB, dst + l * N + (b / Ls) * Ls * Rr + k, Ls, tw + TWO + k * (Rr - 1), k != 0);
    }
    __syncthreads();
    if constexpr (S + 1 < END) staged_stages<RL, T, NTH, S + 1, END>(dst, src, tw);
}

// One tile (T lines) of an n-point c2c pass of the batched kernel for a length whose radices
// have no small common multiple (105 = 3 * 5 * 7: a register tile would hold 105 elements per
// thread).  Every stage runs one butterfly per work unit (T * n / radix of them) over the
// NTH threads; the first stage reads its inputs from global memory and the last stage writes
// its outputs there (stage 0 has Ls = 1: outputs at b * r + q; the last stage has Ls = n / r:
// outputs at b + q * n / r), the stages between go from one LDS buffer to the other.  Radices,
// lines and strides are compile-time constants and each line's offsets are computed once.
// Same Stockham stages and twiddle table as the register tiles.  LDS: 2 * T * n values + 2 * T
// offsets.
template <typename R, typename RL, int T, int NTH>
__device__ __forceinline__ void staged_tile(const pass_args &a, unsigned tile, typename cvec<R>::type *smem) {
    using C = typename cvec<R>::type;
    constexpr int N = RL::n();
    constexpr int LAST = RL::count - 1;
    static_assert(LAST >= 1, "a staged tile has at least two stages");
    C *buf0 = smem;
    C *buf1 = buf0 + T * N;
    long long *base = reinterpret_cast<long long *>(buf1 + T * N);  // [0, T) loads, [T, 2T) stores
    const int tid = threadIdx.x;
    if (tid < T) {
        const unsigned L = tile * T + tid;
        const bool valid = L < (unsigned)a.nlines;
        base[tid] = valid ? line_offset(L, a, false) : -1;
        base[T + tid] = valid ? line_offset(L, a, true) : -1;
    }
    __syncthreads();
    const R sg = a.conj ? R(-1) : R(1);
    const C *__restrict__ tw = static_cast<const C *>(a.tw);
    {
        // stage 0 from global memory into buf1
        constexpr int R0 = RL::at(0), NB0 = N / R0;
        const C *__restrict__ src = static_cast<const C *>(a.in);
        for (int idx = tid; idx < T * NB0; idx += NTH) {
            int l, b;
            if (a.load_colfast) { l = idx % T; b = idx / T; } else { l = idx / NB0; b = idx % NB0; }
            const long long lb = base[l];
            C x[R0];
#pragma unroll
            for (int p = 0; p < R0; p++) {
                const C v = lb >= 0 ? src[lb + (long long)(b + p * NB0) * a.in_sa] : cmk<C>(R(0), R(0));
                x[p] = cmk<C>(v.x, sg * v.y);
            }
            dft<R0>::run(x);
#pragma unroll
            for (int q = 0; q < R0; q++) buf1[l * N + b * R0 + q] = x[q];
        }
        __syncthreads();
    }
    if constexpr (LAST > 1) staged_stages<RL, T, NTH, 1, LAST>(buf1, buf0, tw);
    {
        // last stage from LDS (the output of stage LAST - 1) into global memory
        constexpr int RR = RL::at(LAST), NB = N / RR, TWO = RL::twoff(LAST);
        static_assert(RL::ls(LAST) == NB, "the last stage spans the whole line");
        const C *res = ((LAST - 1) % 2 == 0) ? buf1 : buf0;
        const R sc = (R)a.scale;
        C *__restrict__ dst = static_cast<C *>(a.out);
        for (int idx = tid; idx < T * NB; idx += NTH) {
            int l, b;
            if (a.store_colfast) { l = idx % T; b = idx / T; } else { l = idx / NB; b = idx % NB; }
            const long long ob = base[T + l];
            if (ob < 0) continue;
            C x[RR];
#pragma unroll
            for (int p = 0; p < RR; p++) x[p] = res[l * N + b + p * NB];
            if (b != 0) {
                const C *w = tw + TWO + b * (RR - 1);
#pragma unroll
                for (int p = 1; p < RR; p++) x[p] = cmul(x[p], w[p - 1]);
            }
            dft<RR>::run(x);
#pragma unroll
            for (int q = 0; q < RR; q++)
                dst[ob + (long long)(b + q * NB) * a.out_sa] = cmk<C>(x[q].x * sc, sg * x[q].y * sc);
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(256) fft_generic(pass_args a, generic_plan g) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    generic_tile<R>(a, g, blockIdx.x, reinterpret_cast<typename cvec<R>::type *>(smem_raw));
}

// --------------------------------------------------------------- batched

// One phase (one axis) of a batch of independent c2c transforms of different
// sizes.  The blocks walk the device work list with a grid stride (launched with one
// block per item); each item is one tile of lines of one transform, run by the
// compile-time tile of its length or the runtime-plan tile.
// Compile-time tiles of the batched kernel: (n, E, lines T for complex<float>, pad shift,
// radices).  P = n / E threads per line; P * T <= 256 (complex<float>, 256-thread blocks) and
// P * (T / 2) <= 128 (complex<double>, 128-thread blocks); LDS = (T + 1) rows stays <= 40 KiB.
// E = 0 marks a staged tile (staged_tile: one butterfly per work unit and stage, 2 T rows of LDS).
// The pad shift of each tile minimises the LDS bank conflicts of its stage exchanges, in both the
// contiguous and the column-chunk mapping, under the gfx950 rule (ds_read_b64: 32 lanes per pass
// over 64 banks; ds_write_b64: 16 lanes over 32 banks), at no more LDS than one pad per 16.
#define HMI_BATCH_TABLE(X)            \
    X(0, 16, 16, 256, 4, 16)          \
    X(1, 20, 10, 128, 5, 2, 2, 5)     \
    X(2, 24, 12, 128, 5, 3, 4, 2)     \
    X(3, 27, 9, 84, 4, 3, 3, 3)       \
    X(4, 28, 14, 128, 5, 2, 2, 7)     \
    X(5, 32, 8, 64, 4, 4, 8)          \
    X(6, 40, 10, 64, 6, 2, 2, 2, 5)   \
    X(7, 45, 15, 84, 5, 3, 3, 5)      \
    X(8, 48, 12, 64, 3, 4, 4, 3)      \
    X(9, 56, 14, 64, 6, 2, 2, 2, 7)   \
    X(10, 63, 21, 64, 4, 3, 3, 7)     \
    X(11, 64, 8, 32, 3, 8, 8)         \
    X(12, 80, 10, 32, 7, 2, 2, 2, 2, 5) \
    X(13, 96, 12, 32, 4, 4, 4, 2, 3)  \
    X(14, 112, 14, 32, 7, 2, 2, 2, 2, 7) \
    X(15, 128, 16, 32, 4, 16, 8)       \
    X(16, 105, 0, 16, 0, 3, 5, 7)

constexpr size_t batch_lds_bytes(int n, int e, int t, int ps, size_t value) {
    return e == 0 ? (size_t)2 * t * n * value + (size_t)2 * t * 8 : (size_t)(t + 1) * tile_row_stride(n, ps, t, (int)value) * value;
}
#define HMI_BATCH_INFO(K, N, E, T, PS, ...) {N, T, T / 2, {__VA_ARGS__}, radices<__VA_ARGS__>::count, \
    batch_lds_bytes(N, E, T, PS, 8), batch_lds_bytes(N, E, T / 2, PS, 16)},
static const batch_kernel_info kBatch[] = {HMI_BATCH_TABLE(HMI_BATCH_INFO)};

// Register class of a compile-time tile with E elements per thread (the runtime-plan tile is in
// the last class), and the waves per SIMD each class's build is held to ({complex<float>,
// complex<double>}): the highest occupancy at which the largest tile of the class does not spill
// VGPRs.  A build carries every tile of its class, so its register count is its largest tile's.
// The staged tile (E = 0) keeps one butterfly in registers at a time: light class.
constexpr int batch_class_of(int e) { return e <= 12 ? 0 : (e <= 16 ? 1 : 2); }
constexpr int kBatchWaves[kBatchClasses][2] = {{5, 4}, {4, 3}, {3, 2}};
static_assert(batch_class_of(21) == kBatchClasses - 1, "the 21-element tile shares the runtime-plan tile's class");

#define HMI_BATCH_CLS(K, N, E, T, PS, ...) batch_class_of(E),
static const int kBatchClass[] = {HMI_BATCH_TABLE(HMI_BATCH_CLS)};

int batch_kernel_class(int kid) { return kid < 0 ? kBatchClasses - 1 : kBatchClass[kid]; }

const char *kernels_build_stamp() { return __DATE__ " " __TIME__; }

int find_batch_kernel(int n, batch_kernel_info *info) {
    for (int k = 0; k < (int)(sizeof(kBatch) / sizeof(kBatch[0])); k++)
        if (kBatch[k].n == n) {
            if (info) *info = kBatch[k];
            return k;
        }
    return -1;
}

template <typename R, typename RL, int E, int T, int PS>
__device__ __forceinline__ void batch_tile_of(const pass_args &a, unsigned tile, typename cvec<R>::type *smem) {
    if constexpr (E == 0) staged_tile<R, RL, T, sizeof(R) == 8 ? 128 : 256>(a, tile, smem);
    else uniform_tile<R, RL, E, T, PS>(a, tile, smem);
}

template <typename R, int CLS>
__device__ __forceinline__ void batch_tile(int kid, const pass_args &a, const generic_plan &g, unsigned tile,
                                           typename cvec<R>::type *smem) {
    constexpr bool D = sizeof(R) == 8;
#define HMI_BATCH_CASE(K, N, E, T, PS, ...)                                                        \
    case K:                                                                                        \
        if constexpr (batch_class_of(E) == CLS)                                                    \
            batch_tile_of<R, radices<__VA_ARGS__>, E, (D ? T / 2 : T), PS>(a, tile, smem);         \
        break;
    switch (kid) {
        HMI_BATCH_TABLE(HMI_BATCH_CASE)
        default:
            if constexpr (CLS == kBatchClasses - 1) generic_tile<R>(a, g, tile, smem);
    }
#undef HMI_BATCH_CASE
}

// One build per register class CLS (batch_kernel_class), held to kBatchWaves[CLS] waves per SIMD.
// Splitting the tiles into three classes (from one build at 3 / 2 waves per SIMD) took the
// 1000-transform batch from 11.4 to 6.7 ms per forward + backward pair; the staged tile of
// length 105 (in place of the runtime-plan tile) to 5.7 ms.
template <typename R, int CLS>
__global__ void __launch_bounds__(sizeof(R) == 8 ? 128 : 256)
__attribute__((amdgpu_waves_per_eu(kBatchWaves[CLS][sizeof(R) == 8])))
fft_batched(batch_args b) {
    using C = typename cvec<R>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *smem = reinterpret_cast<C *>(smem_raw);
    const int ax = b.axis;
    const int col = (ax == 0) ? 1 : 0;
    const int b1 = 3 - ax - col;
    for (int i = blockIdx.x; i < b.nitems; i += gridDim.x) {
        const batch_item it = b.items[i];
        const batch_xform x = b.xforms[it.xform];
        const long long st[3] = {1, x.n[0], (long long)x.n[0] * x.n[1]};
        pass_args a;
        a.in = b.from_in ? static_cast<const void *>(static_cast<const C *>(b.in) + x.in_off)
                         : static_cast<const void *>(static_cast<const C *>(b.out) + x.out_off);
        a.out = static_cast<C *>(b.out) + x.out_off;
        a.tw = static_cast<const C *>(b.tw) + b.twoff[x.size_id[ax]];
        a.in_sa = a.out_sa = st[ax];
        a.in_sc = a.out_sc = st[col];
        a.in_sb1 = a.out_sb1 = st[b1];
        a.in_sb2 = a.out_sb2 = 0;
        a.ncol = x.n[col];
        a.nb1 = x.n[b1];
        a.nbatch = 1;
        a.nlines = x.n[col] * x.n[b1];
        // the x axis is contiguous: the y and z phases walk it with consecutive threads
        a.load_colfast = a.store_colfast = (ax != 0 && x.n[0] > 1) ? 1 : 0;
        a.conj = b.conj;
        a.in_real = a.out_real = 0;
        a.store_len = x.n[ax];
        a.herm_load = 0;
        a.mode = kModeC2C;
        a.stream = 0;
        a.mul = nullptr;
        a.tw4 = nullptr;
        a.tw4_dim = 0;
        const double ntot = (double)x.n[0] * (double)x.n[1] * (double)x.n[2];
        a.scale = (b.scaling == 0) ? 1.0 : (b.scaling == 1 ? 1.0 / ntot : sqrt(1.0 / ntot));
        batch_tile<R, CLS>(b.kid[x.size_id[ax]], a, b.plans[x.size_id[ax]], it.tile, smem);
        __syncthreads();
    }
}

using batched_fn = void (*)(batch_args);

template <int... C>
static const batched_fn *batched_builds(bool dbl, std::integer_sequence<int, C...>) {
    static const batched_fn f[2][sizeof...(C)] = {{&fft_batched<float, C>...}, {&fft_batched<double, C>...}};
    return f[dbl ? 1 : 0];
}

hipError_t launch_batched(bool dbl, int cls, const batch_args &b, int nblocks, size_t lds_bytes, hipStream_t s) {
    static const bool configured = [] {
        for (int d = 0; d < 2; d++)
            for (int c = 0; c < kBatchClasses; c++)
                (void)hipFuncSetAttribute(
                    reinterpret_cast<const void *>(batched_builds(d != 0, std::make_integer_sequence<int, kBatchClasses>())[c]),
                    hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
        return true;
    }();
    (void)configured;
    if (cls < 0 || cls >= kBatchClasses) return hipErrorInvalidValue;
    const batched_fn fn = batched_builds(dbl, std::make_integer_sequence<int, kBatchClasses>())[cls];
    fn<<<dim3((unsigned)nblocks), dim3(dbl ? 128 : 256), lds_bytes, s>>>(b);
    return hipGetLastError();
}

// ------------------------------------------------------ batched, fused x-y

// Where the column FFTs of the fused phases store their results: column c of block p of the
// lines at dst + p plane + c, its rows row elements apart; each value becomes (x sc, sg y sc).
template <typename C, typename R> struct col_store {
    C *dst;
    int row;
    long long plane;
    R sg, sc;
    __device__ __forceinline__ C value(C v) const { return cmk<C>(v.x * sc, sg * v.y * sc); }
};

// Stockham stages of one n-point line held in LDS at lb[j * es], E elements per thread (P = n / E
// threads per line); each exchange goes through the line's own slots.  Every thread of the block
// calls it (the barriers are block-wide); threads without a line (valid = false) only take part
// in the barriers.
template <int S, typename RL, int E, int P, typename C>
__device__ __forceinline__ void lds_line_stages(C (&v)[E], C *lb, int es, const C *__restrict__ tw, int t,
                                                bool valid) {
    constexpr int R = RL::at(S);
    constexpr int Ls = RL::ls(S);
    constexpr int NB = E / R;
    constexpr bool last = (S + 1 == RL::count);
    constexpr int TWO = RL::twoff(S);
    static_assert(E % R == 0, "every radix must divide the elements per thread");
#pragma unroll
    for (int r = 0; r < NB; r++) {
        const int b = t + P * r;
        const int k = b % Ls;
        C x[R];
#pragma unroll
        for (int p = 0; p < R; p++) x[p] = v[r + NB * p];
        if constexpr (Ls > 1) {
            const C *w = tw + TWO + k * (R - 1);
#pragma unroll
            for (int p = 1; p < R; p++) x[p] = cmul(x[p], w[p - 1]);
        }
        dft<R>::run(x);
        if constexpr (last) {
#pragma unroll
            for (int q = 0; q < R; q++) v[r + NB * q] = x[q];
        } else if (valid) {
            const int base = (b / Ls) * Ls * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) lb[(base + q * Ls) * es] = x[q];
        }
    }
    if constexpr (!last) {
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = lb[(t + P * e) * es];
        }
        __syncthreads();
        lds_line_stages<S + 1, RL, E, P>(v, lb, es, tw, t, valid);
    }
}

// n-point FFTs of the lines of the x-y planes held in LDS (the dynamic shared memory of the
// kernel): row r of plane p at (p n1 + r) rs, rs = xy_row_stride(n0) (odd), planes ps = n1 rs
// apart.  ACROSS = false: the rows (line l at l rs, element stride 1); ACROSS = true: the columns
// (line l = p n0 + c at p ps + c, element stride rs).  Consecutive lanes take consecutive lines
// at one position t of their lines: the odd row stride puts the lanes of a row pass on distinct
// LDS banks, consecutive columns are consecutive addresses.  NTH threads, rounds of NTH / P
// lines.  The columns' results go from registers straight to global memory (col_store:
// consecutive lanes store consecutive columns).
template <typename RL, int E, int NTH, bool ACROSS, typename C, typename R>
__device__ __forceinline__ void lds_lines_fft(int nlines, int n0, int rs, int ps, const C *__restrict__ tw,
                                              const col_store<C, R> &cs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *buf = reinterpret_cast<C *>(smem_raw);
    constexpr int N = RL::n(), P = N / E, LPI = NTH / P;
    static_assert(LPI >= 1, "a line needs at most the block's threads");
    const int es = ACROSS ? rs : 1;
    const int tid = threadIdx.x;
    const int l0 = tid % LPI;
    // with LPI a multiple of 64 a wave holds 64 consecutive lines at one position t: t, the
    // twiddles and every element offset are uniform over the wave (scalar registers and loads)
    int t = tid / LPI;
    if constexpr (LPI % 64 == 0) t = __builtin_amdgcn_readfirstlane(t);
    const bool inb = tid < LPI * P;
    // a column line = (plane q, column r) of the planes, walked by LPI per round
    int q = l0 / n0, r = l0 % n0;
    const int dq = LPI / n0, dr = LPI % n0;
    for (int first = 0; first < nlines; first += LPI) {
        const int line = first + l0;
        const bool valid = inb && line < nlines;
        C *lb = buf + (!valid ? 0 : (ACROSS ? q * ps + r : line * rs));
        C v[E];
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = lb[(t + P * e) * es];
        }
        __syncthreads();
        lds_line_stages<0, RL, E, P>(v, lb, es, tw, t, valid);
        if (valid) {
            if constexpr (ACROSS) {
                C *gb = cs.dst + q * cs.plane + r;
#pragma unroll
                for (int e = 0; e < E; e++) gb[(t + P * e) * cs.row] = cs.value(v[e]);
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) lb[(t + P * e) * es] = v[e];
            }
        }
        __syncthreads();
        q += dq;
        r += dr;
        if (r >= n0) r -= n0, q++;
    }
}

// The stages of one line as lds_line_stages, for lengths whose radices do not all divide one
// elements-per-thread count (105 = 3 5 7): each stage reads its butterflies' inputs from the
// line's LDS slots (thread t takes butterflies t, t + P, ... < n / R), a barrier, then writes
// their outputs; the last stage of a column (ACROSS) writes to global memory instead (gb: the
// column's first element, rows cs.row apart).
template <int S, typename RL, int P, bool ACROSS, typename C, typename R>
__device__ __forceinline__ void staged_line_stages(C *lb, int es, const C *__restrict__ tw, int t, bool valid, C *gb,
                                                   const col_store<C, R> &cs) {
    constexpr int Rr = RL::at(S), Ls = RL::ls(S), N = RL::n(), NBT = N / Rr, NBS = (NBT + P - 1) / P;
    constexpr bool last = (S + 1 == RL::count);
    constexpr int TWO = RL::twoff(S);
    C x[NBS][Rr];
#pragma unroll
    for (int j = 0; j < NBS; j++) {
        const int b = t + P * j;
        if (valid && b < NBT) {
#pragma unroll
            for (int p = 0; p < Rr; p++) x[j][p] = lb[(b + NBT * p) * es];
            if constexpr (Ls > 1) {
                const C *w = tw + TWO + (b % Ls) * (Rr - 1);
#pragma unroll
                for (int p = 1; p < Rr; p++) x[j][p] = cmul(x[j][p], w[p - 1]);
            }
            dft<Rr>::run(x[j]);
        }
    }
    __syncthreads();  // every read of the stage before its writes
#pragma unroll
    for (int j = 0; j < NBS; j++) {
        const int b = t + P * j;
        if (valid && b < NBT) {
            const int base = (b / Ls) * Ls * Rr + b % Ls;
            if constexpr (last && ACROSS) {
#pragma unroll
                for (int q = 0; q < Rr; q++) gb[(base + q * Ls) * cs.row] = cs.value(x[j][q]);
            } else {
#pragma unroll
                for (int q = 0; q < Rr; q++) lb[(base + q * Ls) * es] = x[j][q];
            }
        }
    }
    if constexpr (!last) {
        __syncthreads();
        staged_line_stages<S + 1, RL, P, ACROSS>(lb, es, tw, t, valid, gb, cs);
    }
}

// lds_lines_fft for those lengths: P threads per line, lines and planes as there.
template <typename RL, int P, int NTH, bool ACROSS, typename C, typename R>
__device__ __forceinline__ void lds_lines_fft_staged(int nlines, int n0, int rs, int ps, const C *__restrict__ tw,
                                                     const col_store<C, R> &cs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *buf = reinterpret_cast<C *>(smem_raw);
    constexpr int LPI = NTH / P;
    static_assert(LPI >= 1, "a line needs at most the block's threads");
    const int es = ACROSS ? rs : 1;
    const int tid = threadIdx.x;
    const int l0 = tid % LPI;
    int t = tid / LPI;
    if constexpr (LPI % 64 == 0) t = __builtin_amdgcn_readfirstlane(t);
    const bool inb = tid < LPI * P;
    // a column line = (plane q, column r) of the planes, walked by LPI per round
    int q = l0 / n0, r = l0 % n0;
    const int dq = LPI / n0, dr = LPI % n0;
    for (int first = 0; first < nlines; first += LPI) {
        const int line = first + l0;
        const bool valid = inb && line < nlines;
        C *lb = buf + (!valid ? 0 : (ACROSS ? q * ps + r : line * rs));
        C *gb = (ACROSS && valid) ? cs.dst + q * cs.plane + r : cs.dst;
        staged_line_stages<0, RL, P, ACROSS>(lb, es, tw, t, valid, gb, cs);
        __syncthreads();
        q += dq;
        r += dr;
        if (r >= n0) r -= n0, q++;
    }
}

// threads per line of the staged lengths in the fused phase
constexpr int kXyStagedThreads = 8;

// Tiles the fused phase can run: every compile-time tile (register tiles through
// lds_lines_fft, the staged tile through lds_lines_fft_staged)
#define HMI_XY_OK(K, N, E, T, PS, ...) true,
static const bool kXyOk[] = {HMI_BATCH_TABLE(HMI_XY_OK)};
bool batch_xy_supported(int kid) { return kid >= 0 && kXyOk[kid]; }


// the lines of one axis of the planes in LDS with the tile of length id kid (class <= CLS)
template <typename R, int CLS, int NTH, bool ACROSS>
__device__ __forceinline__ void xy_axis(int kid, int nlines, int n0, int rs, int ps, const typename cvec<R>::type *tw,
                                        const col_store<typename cvec<R>::type, R> &cs) {
#define HMI_XY_CASE(K, N, E, T, PS, ...)                                                        \
    case K:                                                                                     \
        if constexpr (E > 0 && batch_class_of(E) <= CLS && NTH / (N / (E > 0 ? E : 1)) >= 1)   \
            lds_lines_fft<radices<__VA_ARGS__>, (E > 0 ? E : 1), NTH, ACROSS>(nlines, n0, rs, ps, tw, cs); \
        else if constexpr (E == 0)                                                              \
            lds_lines_fft_staged<radices<__VA_ARGS__>, kXyStagedThreads, NTH, ACROSS>(nlines, n0, rs, ps, tw, cs); \
        break;
    switch (kid) {
        HMI_BATCH_TABLE(HMI_XY_CASE)
        default: break;
    }
#undef HMI_XY_CASE
}

// Fused x-y phase: work item = (transform, group of G consecutive z-planes), G = max(1, budget /
// (n1 rs)), rs = xy_row_stride(n0).  The planes are contiguous in memory: one coalesced load into
// the padded rows in LDS (conjugated for the backward transform), the n0-point FFTs of the G n1
// rows, the n1-point FFTs of the G n0 columns, whose results go from registers to global memory
// (conjugated back; consecutive lanes write consecutive columns).  No scale: the z phase applies
// it.
template <typename R, int CLS, int NTH>
__device__ __forceinline__ void xy_item(const batch_args &b, const batch_item &it, typename cvec<R>::type *buf) {
    using C = typename cvec<R>::type;
    const batch_xform x = b.xforms[it.xform];
    const int n0 = x.n[0], n1 = x.n[1], n2 = x.n[2];
    const int rs = xy_row_stride(n0), ps = n1 * rs;
    const int G = b.xy_budget / ps > 1 ? b.xy_budget / ps : 1;
    const int z0 = it.tile * G;
    const int g = (n2 - z0) < G ? (n2 - z0) : G;
    const int cnt = g * n1 * n0;
    const long long zoff = (long long)z0 * n1 * n0;
    const C *__restrict__ src = (b.from_in ? static_cast<const C *>(b.in) + x.in_off
                                           : static_cast<const C *>(b.out) + x.out_off) + zoff;
    C *__restrict__ dst = static_cast<C *>(b.out) + x.out_off + zoff;
    const R sg = b.conj ? R(-1) : R(1);
    // element i of the group sits in LDS row i / n0 (the g n1 rows back to back), column i % n0;
    // walking i by NTH moves the pair by (NTH / n0, NTH % n0).  Loads go out U at a time before
    // their LDS stores, so that U loads per thread are in flight.
    constexpr int U = 24;
    const int dr = NTH / n0, dc = NTH % n0;
    const int r0 = (int)threadIdx.x / n0, c0 = (int)threadIdx.x % n0;
    {
        int r = r0, c = c0;
        for (int i0 = threadIdx.x; i0 < cnt; i0 += U * NTH) {
            C v[U];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (i0 + u * NTH < cnt) v[u] = src[i0 + u * NTH];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (i0 + u * NTH < cnt) buf[r * rs + c] = cmk<C>(v[u].x, sg * v[u].y);
                r += dr;
                c += dc;
                if (c >= n0) c -= n0, r++;
            }
        }
    }
    __syncthreads();
    const C *__restrict__ tw = static_cast<const C *>(b.tw);
    // x: g n1 rows of n0 points; y: g n0 columns of n1 points, stored from registers
    const col_store<C, R> cs{dst, n0, (long long)n0 * n1, sg, R(1)};
    xy_axis<R, CLS, NTH, false>(b.kid[x.size_id[0]], g * n1, n0, rs, ps, tw + b.twoff[x.size_id[0]], cs);
    xy_axis<R, CLS, NTH, true>(b.kid[x.size_id[1]], g * n0, n0, rs, ps, tw + b.twoff[x.size_id[1]], cs);
}

// Every build at 2 waves per SIMD (up to 256 VGPRs): two 256-thread blocks per CU with 80 KiB of
// LDS each, or one 512-thread block with up to 160 KiB.  The launchers use only the build of
// class kXyMaxClass, which carries every tile.
constexpr int kXyWaves[kBatchClasses][2] = {{2, 2}, {2, 2}, {2, 2}};

template <typename R, int CLS>
__global__ void __launch_bounds__(kXySmallThreads)
__attribute__((amdgpu_waves_per_eu(kXyWaves[CLS][sizeof(R) == 8])))
fft_batched_xy(batch_args b) {
    using C = typename cvec<R>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    for (int i = blockIdx.x; i < b.nitems; i += gridDim.x)
        xy_item<R, CLS, kXySmallThreads>(b, b.items[i], reinterpret_cast<C *>(smem_raw));
}

__global__ void __launch_bounds__(kXyLargeThreads) __attribute__((amdgpu_waves_per_eu(2)))
fft_batched_xy_large(batch_args b) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    for (int i = blockIdx.x; i < b.nitems; i += gridDim.x)
        xy_item<float, kXyLargeMaxClass[0], kXyLargeThreads>(b, b.items[i], reinterpret_cast<float2 *>(smem_raw));
}
static_assert(kXyLargeMaxClass[1] < 0, "the 1024-thread build is complex<float> only");

// Column-block z phase: work item = (transform, block of w <= W = min(n0 n1, budget / n2)
// consecutive (x, y) positions), the n2-point z lines of those positions: one load of the n2
// rows of w contiguous elements (rows n0 n1 apart in memory) into LDS rows of w, the z FFTs with
// consecutive lanes on consecutive positions (as the column FFTs of the x-y phase), results
// stored from registers, scaled and conjugated back for the backward transform.
template <typename R, int CLS, int NTH>
__device__ __forceinline__ void zcol_item(const batch_args &b, const batch_item &it, typename cvec<R>::type *buf) {
    using C = typename cvec<R>::type;
    const batch_xform x = b.xforms[it.xform];
    const int plane = x.n[0] * x.n[1], n2 = x.n[2];
    const int W = zcol_width(n2, b.xy_budget, plane);
    const int c0 = it.tile * W;
    const int w = (plane - c0) < W ? (plane - c0) : W;
    const int cnt = n2 * w;
    const C *__restrict__ src = (b.from_in ? static_cast<const C *>(b.in) + x.in_off
                                           : static_cast<const C *>(b.out) + x.out_off) + c0;
    C *__restrict__ dst = static_cast<C *>(b.out) + x.out_off + c0;
    const R sg = b.conj ? R(-1) : R(1);
    const double ntot = (double)plane * (double)n2;
    const R sc = (R)((b.scaling == 0) ? 1.0 : (b.scaling == 1 ? 1.0 / ntot : sqrt(1.0 / ntot)));
    // element i of the block (row z = i / w, column c = i % w) sits at LDS i; walking i by NTH
    // moves (z, c) by (NTH / w, NTH % w).  Loads go out U at a time before their LDS stores.
    constexpr int U = 24;
    const int dz = NTH / w, dc = NTH % w;
    int z = (int)threadIdx.x / w, c = (int)threadIdx.x % w;
    for (int i0 = threadIdx.x; i0 < cnt; i0 += U * NTH) {
        C v[U];
        int zz = z, cc = c;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (i0 + u * NTH < cnt) v[u] = src[(long long)zz * plane + cc];
            zz += dz;
            cc += dc;
            if (cc >= w) cc -= w, zz++;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (i0 + u * NTH < cnt) buf[i0 + u * NTH] = cmk<C>(v[u].x, sg * v[u].y);
        z = zz;
        c = cc;
    }
    __syncthreads();
    const col_store<C, R> cs{dst, plane, 0, sg, sc};
    xy_axis<R, CLS, NTH, true>(b.kid[x.size_id[2]], w, w, w, 0,
                               static_cast<const C *>(b.tw) + b.twoff[x.size_id[2]], cs);
}

template <typename R, int CLS>
__global__ void __launch_bounds__(kXySmallThreads)
__attribute__((amdgpu_waves_per_eu(kXyWaves[CLS][sizeof(R) == 8])))
fft_batched_zcol(batch_args b) {
    using C = typename cvec<R>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    for (int i = blockIdx.x; i < b.nitems; i += gridDim.x)
        zcol_item<R, CLS, kXySmallThreads>(b, b.items[i], reinterpret_cast<C *>(smem_raw));
}

template <typename R, int CLS>
constexpr batched_fn zcol_build() {
    if constexpr (CLS == kXyMaxClass[sizeof(R) == 8]) return &fft_batched_zcol<R, CLS>;
    else return nullptr;
}
template <int... CS>
static const batched_fn *batched_zcol_builds(bool dbl, std::integer_sequence<int, CS...>) {
    static const batched_fn f[2][sizeof...(CS)] = {{zcol_build<float, CS>()...}, {zcol_build<double, CS>()...}};
    return f[dbl ? 1 : 0];
}

hipError_t launch_batched_zcol(bool dbl, int cls, const batch_args &b, int nblocks, size_t lds_bytes, hipStream_t s) {
    static const bool configured = [] {
        for (int d = 0; d < 2; d++)
            for (int c = 0; c < kBatchClasses; c++) {
                const batched_fn fn = batched_zcol_builds(d != 0, std::make_integer_sequence<int, kBatchClasses>())[c];
                if (fn)
                    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
            }
        return true;
    }();
    (void)configured;
    if (cls < 0 || cls >= kBatchClasses || cls > kXyMaxClass[dbl ? 1 : 0]) return hipErrorInvalidValue;
    if (lds_bytes > (size_t)kGenericMaxLds) return hipErrorInvalidValue;
    const batched_fn fn = batched_zcol_builds(dbl, std::make_integer_sequence<int, kBatchClasses>())[cls];
    if (!fn) return hipErrorInvalidValue;
    fn<<<dim3((unsigned)nblocks), dim3(kXySmallThreads), lds_bytes, s>>>(b);
    return hipGetLastError();
}

template <typename R, int CLS>
constexpr batched_fn xy_build() {
    if constexpr (CLS == kXyMaxClass[sizeof(R) == 8]) return &fft_batched_xy<R, CLS>;
    else return nullptr;
}
template <int... CS>
static const batched_fn *batched_xy_builds(bool dbl, std::integer_sequence<int, CS...>) {
    static const batched_fn f[2][sizeof...(CS)] = {{xy_build<float, CS>()...}, {xy_build<double, CS>()...}};
    return f[dbl ? 1 : 0];
}

hipError_t launch_batched_xy(bool dbl, bool large, int cls, const batch_args &b, int nblocks, size_t lds_bytes,
                             hipStream_t s) {
    static const bool configured = [] {
        for (int d = 0; d < 2; d++)
            for (int c = 0; c < kBatchClasses; c++) {
                const batched_fn fn = batched_xy_builds(d != 0, std::make_integer_sequence<int, kBatchClasses>())[c];
                if (fn)
                    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
            }
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&fft_batched_xy_large),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
        return true;
    }();
    (void)configured;
    if (cls < 0 || cls >= kBatchClasses || cls > (large ? kXyLargeMaxClass : kXyMaxClass)[dbl ? 1 : 0])
        return hipErrorInvalidValue;
    if (!large && !batched_xy_builds(dbl, std::make_integer_sequence<int, kBatchClasses>())[cls])
        return hipErrorInvalidValue;
    if (lds_bytes > (size_t)kGenericMaxLds) return hipErrorInvalidValue;
    if (large) {
        fft_batched_xy_large<<<dim3((unsigned)nblocks), dim3(kXyLargeThreads), lds_bytes, s>>>(b);
    } else {
        const batched_fn fn = batched_xy_builds(dbl, std::make_integer_sequence<int, kBatchClasses>())[cls];
        fn<<<dim3((unsigned)nblocks), dim3(kXySmallThreads), lds_bytes, s>>>(b);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------- permutation

// dst[i so_f + j so_m + k so_s] = src[i si_f + j si_m + k si_s] over a
// (n_f, n_m, n_s) index box, batched with strides bs_in / bs_out.  The grid
// tiles the (fi, fo) plane of the source-fastest and destination-fastest axes
// with 32x32 LDS tiles so that both the read and the write are coalesced.
template <typename E>
__global__ void __launch_bounds__(256) permute_kernel(const E *__restrict__ src, E *__restrict__ dst, permute_args p) {
    __shared__ E tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    long long blk = blockIdx.x;
    const long long ta = blk % p.ntile_a; blk /= p.ntile_a;
    const long long tb = blk % p.ntile_b; blk /= p.ntile_b;
    const long long c = blk % p.n_c;
    const long long bt = blk / p.n_c;
    const long long a0 = ta * 32, b0 = tb * 32;
    const long long sbase = bt * p.bs_in + c * p.si_c;
    const long long dbase = bt * p.bs_out + c * p.so_c;
    // read: a is the source-fastest axis
    for (int r = ty; r < 32; r += 8) {
        const long long ia = a0 + tx, ib = b0 + r;
        if (ia < p.n_a && ib < p.n_b) tile[r][tx] = src[sbase + ia * p.si_a + ib * p.si_b];
    }
    __syncthreads();
    if (p.same_fast) {
        // both sides are fastest along a: write rows back in the read order
        for (int r = ty; r < 32; r += 8) {
            const long long ia = a0 + tx, ib = b0 + r;
            if (ia < p.n_a && ib < p.n_b) dst[dbase + ia * p.so_a + ib * p.so_b] = tile[r][tx];
        }
    } else {
        // b is the destination-fastest axis
        for (int r = ty; r < 32; r += 8) {
            const long long ia = a0 + r, ib = b0 + tx;
            if (ia < p.n_a && ib < p.n_b) dst[dbase + ia * p.so_a + ib * p.so_b] = tile[tx][r];
        }
    }
}

// ----------------------------------------------------------- fast registry

#define HMI_FAST(TYPE, N, E, T, PS, ...)                                                                  \
    static hipError_t launch_##TYPE##_##N##_##T(const pass_args &a, long long nblocks, hipStream_t s) {    \
        if (a.stream == kStreamStores)                                                                     \
            fft_uniform<TYPE, radices<__VA_ARGS__>, E, T, PS, kStreamStores>                               \
                <<<dim3((unsigned)nblocks), dim3((N / E) * T), 0, s>>>(a);                                 \
        else if (a.stream)                                                                                 \
            fft_uniform<TYPE, radices<__VA_ARGS__>, E, T, PS, kStreamLoadsStores>                          \
                <<<dim3((unsigned)nblocks), dim3((N / E) * T), 0, s>>>(a);                                 \
        else                                                                                               \
            fft_uniform<TYPE, radices<__VA_ARGS__>, E, T, PS, 0>                                           \
                <<<dim3((unsigned)nblocks), dim3((N / E) * T), 0, s>>>(a);                                 \
        return hipGetLastError();                                                                          \
    }

#define HMI_FAST_TABLE(X)                        \
    X(float, 8, 8, 256, 3, 8)                    \
    X(float, 16, 16, 128, 4, 16)                 \
    X(float, 32, 8, 64, 3, 4, 8)                 \
    X(float, 64, 8, 32, 3, 8, 8)                 \
    X(float, 128, 16, 32, 4, 16, 8)              \
    X(float, 256, 16, 16, 4, 16, 16)             \
    X(float, 512, 8, 16, 3, 8, 8, 8)             \
    X(float, 512, 8, 4, 3, 8, 8, 8)              \
    X(float, 1024, 16, 16, 4, 16, 8, 8)          \
    X(float, 1024, 16, 8, 4, 16, 8, 8)           \
    X(float, 1024, 16, 4, 4, 16, 8, 8)           \
    X(float, 2048, 16, 8, 4, 8, 16, 16)          \
    X(float, 2048, 16, 4, 4, 8, 16, 16)          \
    X(float, 2048, 16, 2, 4, 8, 16, 16)          \
    X(float, 4096, 16, 4, 4, 16, 16, 16)         \
    X(float, 4096, 16, 1, 4, 16, 16, 16)         \
    X(float, 12, 12, 128, 4, 4, 3)               \
    X(float, 20, 20, 64, 4, 4, 5)                \
    X(float, 24, 12, 64, 4, 3, 4, 2)             \
    X(float, 27, 9, 64, 4, 3, 3, 3)              \
    X(float, 28, 28, 64, 4, 4, 7)                \
    X(float, 40, 20, 64, 4, 2, 4, 5)             \
    X(float, 45, 15, 64, 4, 3, 3, 5)             \
    X(float, 48, 12, 32, 4, 4, 4, 3)             \
    X(float, 56, 28, 64, 4, 2, 4, 7)             \
    X(float, 63, 21, 64, 4, 3, 3, 7)             \
    X(float, 80, 20, 32, 4, 4, 4, 5)             \
    X(float, 96, 12, 16, 4, 4, 4, 2, 3)          \
    X(float, 112, 28, 32, 4, 4, 4, 7)            \
    X(double, 8, 8, 128, 3, 8)                   \
    X(double, 16, 16, 64, 4, 16)                 \
    X(double, 32, 8, 32, 3, 4, 8)                \
    X(double, 64, 8, 16, 3, 8, 8)                \
    X(double, 128, 16, 16, 4, 16, 8)             \
    X(double, 256, 8, 16, 3, 8, 8, 4)            \
    X(double, 256, 8, 8, 3, 8, 8, 4)             \
    X(double, 256, 8, 4, 3, 8, 8, 4)             \
    X(double, 512, 8, 8, 3, 8, 8, 8)             \
    X(double, 512, 8, 4, 3, 8, 8, 8)             \
    X(double, 1024, 16, 8, 4, 16, 8, 8)          \
    X(double, 1024, 16, 4, 4, 16, 8, 8)          \
    X(double, 1024, 16, 2, 4, 16, 8, 8)          \
    X(double, 2048, 16, 4, 4, 8, 16, 16)         \
    X(double, 2048, 16, 2, 4, 8, 16, 16)         \
    X(double, 4096, 16, 2, 4, 16, 16, 16)        \
    X(double, 4096, 16, 1, 4, 16, 16, 16)

HMI_FAST_TABLE(HMI_FAST)

#define HMI_FAST_ENTRY(TYPE, N, E, T, PS, ...) \
    {sizeof(TYPE) == 8, N, E, T, PS, {__VA_ARGS__}, radices<__VA_ARGS__>::count, &launch_##TYPE##_##N##_##T},

static const fast_kernel_info kFast[] = {HMI_FAST_TABLE(HMI_FAST_ENTRY)};

const fast_kernel_info *find_fast_kernel(int n, bool dbl, bool wide) {
    const fast_kernel_info *best = nullptr;
    for (const auto &k : kFast) {
        if (k.n != n || k.dbl != dbl) continue;
        if (!best || (wide ? k.T > best->T : k.T < best->T)) best = &k;
    }
    return best;
}

hipError_t launch_generic(bool dbl, const pass_args &a, const generic_plan &g, long long nblocks, size_t lds_bytes,
                          hipStream_t s) {
    // once per process, thread-safe (static initialisation)
    static const bool configured = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&fft_generic<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&fft_generic<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kGenericMaxLds);
        return true;
    }();
    (void)configured;
    if (dbl) fft_generic<double><<<dim3((unsigned)nblocks), dim3(256), lds_bytes, s>>>(a, g);
    else fft_generic<float><<<dim3((unsigned)nblocks), dim3(256), lds_bytes, s>>>(a, g);
    return hipGetLastError();
}

hipError_t launch_permute(int elem_bytes, const void *src, void *dst, const permute_args &p, hipStream_t s) {
    const long long nblocks = p.ntile_a * p.ntile_b * p.n_c * p.nbatch;
    if (nblocks <= 0) return hipSuccess;
    const dim3 grid((unsigned)nblocks), block(256);
    switch (elem_bytes) {
        case 4: permute_kernel<uint32_t><<<grid, block, 0, s>>>((const uint32_t *)src, (uint32_t *)dst, p); break;
        case 8: permute_kernel<uint2><<<grid, block, 0, s>>>((const uint2 *)src, (uint2 *)dst, p); break;
        case 16: permute_kernel<uint4><<<grid, block, 0, s>>>((const uint4 *)src, (uint4 *)dst, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hmi
