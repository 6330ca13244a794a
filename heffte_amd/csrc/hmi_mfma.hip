// 512-point c2c pass of complex<float> data computed on the matrix cores (gfx950 MFMA).
//
// n = 512 = 32 x 16 (four-step): with n = 16 n1 + n2 and k = k1 + 32 k2,
//   Y[k1][n2] = sum_n1 W32^{n1 k1} x[16 n1 + n2]            (step A, 32-point DFTs)
//   X[k1 + 32 k2] = sum_n2 W16^{n2 k2} (W512^{n2 k1} Y[k1][n2])   (twiddle, step C, 16-point DFTs)
// Steps A and C are real GEMMs on the block form [Re F, -Im F; Im F, Re F] x [Re v; Im v]
// with v_mfma_f32_16x16x32_{f16,bf16} (FP32 accumulation).  Accuracy comes from splitting
// every FP32 operand into low-precision terms:
//   kSplitF16:  x = hi + lo (two FP16 terms), products hi*hi + hi*lo + lo*hi
//   kSplitBF16: x = hi + mid + lo (three BF16 terms), products hh + hm + mh + hl + lh + mm
// and from a power-of-two scale per GEMM column (one 32- or 16-point DFT) that puts its
// maximum just below 1 so the terms stay in the normal range of the 16-bit format; the
// scale is undone in FP32 on the GEMM output.
//
// MFMA 16x16x32 lane layout: A[i][k] in lane i + 16 (k / 8), element k % 8; B[k][j] in lane
// j + 16 (k / 8), element k % 8; C[i][j] in lane j + 16 (i / 4), element i % 4.  Only the row
// and column assignment matters here: A and B fill the same (lane group, element) K slots, so
// the sum over k does not depend on how the hardware orders them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hmi_device.h"
#include "hmi_kernels.h"


namespace hmi {

namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int PREC> struct split_traits;

template <> struct split_traits<kSplitF16> {
    static constexpr int terms = 2;
    using vec = half8_t;
    __device__ static f32x4_t mma(vec a, vec b, f32x4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct split_traits<kSplitBF16> {
    static constexpr int terms = 3;
    using vec = bf8_t;
    __device__ static f32x4_t mma(vec a, vec b, f32x4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// BF16 three-term split of 8 values: a BF16 value is the upper half of an FP32 word, so each
// term is the current remainder with its low 16 bits cleared (exact: remainder - term is
// representable) and two terms pack into one word with a byte permute
__device__ __forceinline__ void split8_bf16(const float (&x)[8], bf8_t (&out)[3]) {
    unsigned t[3][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float ra = x[2 * q], rb = x[2 * q + 1];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const unsigned ua = __float_as_uint(ra) & 0xffff0000u, ub = __float_as_uint(rb) & 0xffff0000u;
            t[k][q] = __builtin_amdgcn_perm(ub, ua, 0x07060302u);  // (a, b) as two BF16 halves
            ra -= __uint_as_float(ua);
            rb -= __uint_as_float(ub);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) __builtin_memcpy(&out[k], t[k], sizeof(t[k]));
}

// FP16 hi/lo split of 8 values times a power-of-two scale: per pair, v_fma_mix forms
// hi = f16(x s) and lo = f16(x s - hi) with the scale folded into the multiply
__device__ __forceinline__ void split8_f16_scaled(const float (&x)[8], float s, half8_t (&out)[3]) {
    unsigned hi[4], lo[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
        asm volatile(
            "v_fma_mixlo_f16 %0, %2, %4, 0\n\t"
            "v_fma_mixhi_f16 %0, %3, %4, 0\n\t"
            "v_fma_mixlo_f16 %1, %2, %4, -%0 op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixhi_f16 %1, %3, %4, -%0 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
            : "=&v"(hi[q]), "=&v"(lo[q])
            : "v"(x[2 * q]), "v"(x[2 * q + 1]), "v"(s));
    __builtin_memcpy(&out[0], hi, sizeof(hi));
    __builtin_memcpy(&out[1], lo, sizeof(lo));
}

// max(m, |a|, |b|) in one instruction
__device__ __forceinline__ float absmax3(float m, float a, float b) {
    float r;
    asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}

// Maximum of a non-negative value over lanes l, l ^ 16, l ^ 32 and l ^ 48 (the four K groups
// of a GEMM column) with the gfx950 row swaps; non-negative floats order like their bits.
__device__ __forceinline__ float max_over_groups(float m) {
    const unsigned u = __float_as_uint(m);
    const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    const unsigned v = max(r[0], r[1]);
    const auto t = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return __uint_as_float(max(t[0], t[1]));
}

// acc += A * B through the split products
template <int PREC>
__device__ __forceinline__ f32x4_t split_mma(const typename split_traits<PREC>::vec (&a)[3],
                                             const typename split_traits<PREC>::vec (&b)[3], f32x4_t acc) {
    using S = split_traits<PREC>;
    if constexpr (S::terms == 2) {
        acc = S::mma(a[0], b[0], acc);
        acc = S::mma(a[0], b[1], acc);
        acc = S::mma(a[1], b[0], acc);
    } else {
        acc = S::mma(a[0], b[0], acc);
        acc = S::mma(a[0], b[1], acc);
        acc = S::mma(a[1], b[0], acc);
        acc = S::mma(a[0], b[2], acc);
        acc = S::mma(a[2], b[0], acc);
        acc = S::mma(a[1], b[1], acc);
    }
    return acc;
}

__device__ __forceinline__ long long line_off(unsigned L, const pass_args &a, bool out) {
    const unsigned q = L / (unsigned)a.ncol;
    const unsigned c = L - q * (unsigned)a.ncol;
    const unsigned b2 = q / (unsigned)a.nb1;
    const unsigned b1 = q - b2 * (unsigned)a.nb1;
    return out ? (long long)c * a.out_sc + (long long)b1 * a.out_sb1 + (long long)b2 * a.out_sb2
               : (long long)c * a.in_sc + (long long)b1 * a.in_sb1 + (long long)b2 * a.in_sb2;
}

// Offsets of line tile * 16 of the input and output arrays when the column count is a
// multiple of 16 (AL16): the tile's lines then share (b1, batch), and line l of the tile
// lies l column strides further.  The divisions run once per tile on wave-uniform values.
struct tile_pos {
    long long in, out;
};
template <bool AL16> __device__ __forceinline__ tile_pos locate_tile(unsigned tile, const pass_args &a) {
    tile_pos t{0, 0};
    if constexpr (AL16) {
        const unsigned L0 = tile * 16;
        const unsigned q = L0 / (unsigned)a.ncol;
        const unsigned c0 = L0 - q * (unsigned)a.ncol;
        const unsigned b2 = q / (unsigned)a.nb1;
        const unsigned b1 = q - b2 * (unsigned)a.nb1;
        t.in = (long long)c0 * a.in_sc + (long long)b1 * a.in_sb1 + (long long)b2 * a.in_sb2;
        t.out = (long long)c0 * a.out_sc + (long long)b1 * a.out_sb1 + (long long)b2 * a.out_sb2;
    }
    return t;
}
template <bool AL16>
__device__ __forceinline__ long long tile_line_off(const tile_pos &t, unsigned tile, int l, const pass_args &a,
                                                   bool out) {
    if constexpr (AL16) return (out ? t.out : t.in) + (long long)l * (out ? a.out_sc : a.in_sc);
    else return line_off(tile * 16 + l, a, out);
}

// Y tile in LDS, unpadded, with a swizzle that spreads the accesses of a wave over the banks:
// element (line, k1, n2) at line * 512 + pr * 16 + (n2 ^ sw), pr = k1 with its low bit
// flipped by bit 2 of k1 and bit 3 of the line, sw = 2 ((k1 / 2 % 8) ^ (line % 8)) (even, so
// the pairs (2m, 2m + 1) read as one 16-byte access stay in order)
__device__ __forceinline__ int yswz(int line, int k1, int n2) {
    const int pr = k1 ^ (((k1 >> 2) ^ (line >> 3)) & 1);
    const int sw = (((k1 >> 1) & 7) ^ (line & 7)) << 1;
    return line * 512 + pr * 16 + (n2 ^ sw);
}
// W512 table [k1][n2] with the row swizzle of yswz for line 0
__device__ __forceinline__ int twswz(int k1, int n2) { return k1 * 16 + (n2 ^ (((k1 >> 1) & 7) << 1)); }

}  // namespace

// One tile = 16 lines x 512 points for one 1024-thread block (16 waves); lane = j + 16 q
// (j < 16, q < 4).
//
// Step A runs straight from the registers the tile was loaded into: thread (wave w, j, q)
// loads x[16 n1 + n2] for n1 = q + 4 e (e < 8) of one column (n2, line) -- n2 = j, line = w
// when the axis is contiguous in memory (each load then reads 512 contiguous bytes), n2 = w,
// line = j when the column axis is (16 lines per 128-byte chunk) -- which is its B fragment
// of v_mfma_f32_16x16x32 (B[k][j] in lane j + 16 (k / 8), element k % 8) with the K slot
// (q, e) standing for n1 = q + 4 e; the A operands use the same assignment.  Each column is
// scaled by its own power of two (maximum over the four lanes j, j + 16, j + 32, j + 48: two
// row swaps), split once, and multiplied by the 16 x 32 blocks Re F32 / Im F32 of output rows
// k1 < 16 and k1 >= 16 (real rows: Re F . re - Im F . im, the second product subtracted in
// FP32; imaginary rows: Im F . re + Re F . im), one operand matrix live at a time.  The
// unscaled Y[k1][n2] goes to LDS; after a barrier each wave takes 32 columns (k1, line) of
// step C -- k1 = 16 cb + j of line w, or k1 = 2 w + cb of line j when the store walks the
// column axis.  Its K index interleaves real and imaginary parts (element e of lane group q
// is part e & 1 of n2 = 4 q + e / 2), so each lane reads 4 consecutive n2 of Y and of the
// W512 table, forms the 4 twiddled complex values, scales the column by its own power of two,
// splits it and runs the two 16 x 32 row blocks of [Re F16, -Im F16; Im F16, Re F16] (columns
// permuted alike; row i of a block stands for k2 = 4 (i % 4) + i / 4).  Lane (j, q) then
// holds X[k1 + 32 k2] for k2 = 4 r + q (r < 4): per store, 128-byte chunks within 1 KiB.
//
// Occupancy: at most 64 VGPRs and 80 KB of LDS, so two blocks (8 waves per SIMD) share a CU
// and hide each other's memory latency; no tile is prefetched.  a.tw = W512^{n2 k1} as
// [k1][n2], then the step-A operands [rb][Re/Im][term][lane][8], then step C [row
// block][term][lane][8]; each block copies the W512 table, the step-C operands and (FP16) the
// step-A operands to LDS while its loads are in flight; BF16 reads its step-A operands from
// the cached table (no LDS left for three terms).
// LCF / SCF: the load / store side walks the column axis; AL16: the column count is a
// multiple of 16 (see tile_line_off).
template <int PREC, bool LCF, bool SCF, bool AL16>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
fft512_mfma(pass_args a) {
    using S = split_traits<PREC>;
    using vec = typename S::vec;
    constexpr int TERMS = S::terms;
    constexpr bool kOpALds = (TERMS == 2);
    constexpr size_t kY = 0, kTw = kY + sizeof(float2) * 16 * 512, kOpC = kTw + sizeof(float2) * 512;
    constexpr size_t kOpA = kOpC + sizeof(vec) * 2 * TERMS * 64;
    constexpr size_t kBytes = kOpA + (kOpALds ? sizeof(vec) * 4 * TERMS * 64 : 0);
    __shared__ __align__(16) unsigned char lds[kBytes];
    float2 *ybuf = reinterpret_cast<float2 *>(lds + kY);
    float2 *tws = reinterpret_cast<float2 *>(lds + kTw);
    vec *opCs = reinterpret_cast<vec *>(lds + kOpC);
    vec *opAs = reinterpret_cast<vec *>(lds + kOpA);

    // 16-line tiles: 128-byte column chunks, numbered in groups of 8 per XCD like the FP32 pass
    const unsigned tile = (LCF || SCF) ? xcd_grouped_tile<8>(blockIdx.x, gridDim.x) : blockIdx.x;
    const unsigned nlines = (unsigned)a.nlines;
    const float2 *__restrict__ w512 = static_cast<const float2 *>(a.tw);
    const vec *__restrict__ opA = reinterpret_cast<const vec *>(w512 + 512);
    const vec *__restrict__ opC = opA + 2 * 2 * TERMS * 64;
    const tile_pos tp = locate_tile<AL16>(tile, a);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, j = tid & 15, q = (tid >> 4) & 3;
    // the tile's loads first, then the operand tables to LDS
    float2 x[8];
    {
        const int line = LCF ? j : w, n2 = LCF ? w : j;
        const unsigned L = tile * 16 + line;
        const float2 *p = static_cast<const float2 *>(a.in) +
                          (L < nlines ? tile_line_off<AL16>(tp, tile, line, a, false) : 0) +
                          (long long)(16 * q + n2) * a.in_sa;
        const long long step = 64 * a.in_sa;
#pragma unroll
        for (int e = 0; e < 8; e++) x[e] = ld_stream(&p[e * step]);
    }
    if (tid < 256) {
        const int k1 = tid >> 3, n2 = 2 * (tid & 7);
        const float4 v = reinterpret_cast<const float4 *>(w512)[tid];
        *reinterpret_cast<float4 *>(&tws[twswz(k1, n2)]) = v;
    } else if (tid < 256 + 2 * TERMS * 64) {
        opCs[tid - 256] = opC[tid - 256];
    } else if (kOpALds && tid < 256 + 6 * TERMS * 64) {
        opAs[tid - 256 - 2 * TERMS * 64] = opA[tid - 256 - 2 * TERMS * 64];
    }
    const vec *opAsrc = kOpALds ? opAs : opA;

    // column scale and split (before the barrier: no LDS involved)
    vec bre[3], bim[3];
    int exA = 0;
    {
        float m = 0.f;
#pragma unroll
        for (int e = 0; e < 8; e++) m = absmax3(m, x[e].x, x[e].y);
        m = max_over_groups(m);
        (void)frexpf(m, &exA);
        const float s = ldexpf(1.0f, -exA);
        float re[8], im[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            if constexpr (PREC == kSplitF16) {
                re[e] = x[e].x;
                im[e] = x[e].y;
            } else {
                re[e] = x[e].x * s;
                im[e] = x[e].y * s;
            }
        }
        if constexpr (PREC == kSplitF16) {
            split8_f16_scaled(re, s, bre);
            split8_f16_scaled(im, s, bim);
        } else {
            split8_bf16(re, bre);
            split8_bf16(im, bim);
        }
    }
    __syncthreads();  // operand tables in LDS
    {
        const int line = LCF ? j : w, n2 = LCF ? w : j;
        const float u = ldexpf(1.0f, exA);
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            // one operand matrix live at a time: Re F (real rows . re, imaginary rows . im),
            // then Im F
            vec f[3];
#pragma unroll
            for (int tt = 0; tt < TERMS; tt++) f[tt] = opAsrc[((rb * 2 + 0) * TERMS + tt) * 64 + lane];
            const f32x4_t pr = split_mma<PREC>(f, bre, f32x4_t{});
            f32x4_t yi = split_mma<PREC>(f, bim, f32x4_t{});
#pragma unroll
            for (int tt = 0; tt < TERMS; tt++) f[tt] = opAsrc[((rb * 2 + 1) * TERMS + tt) * 64 + lane];
            const f32x4_t pi = split_mma<PREC>(f, bim, f32x4_t{});
            yi = split_mma<PREC>(f, bre, yi);
#pragma unroll
            for (int r = 0; r < 4; r++)
                ybuf[yswz(line, 16 * rb + 4 * q + r, n2)] = make_float2((pr[r] - pi[r]) * u, yi[r] * u);
        }
    }
    __syncthreads();
    {
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const int line = SCF ? j : w, k1 = SCF ? 2 * w + cb : 16 * cb + j;
            float2 y[4], t[4];
            {
                const float4 y0 = *reinterpret_cast<const float4 *>(&ybuf[yswz(line, k1, 4 * q)]);
                const float4 y1 = *reinterpret_cast<const float4 *>(&ybuf[yswz(line, k1, 4 * q + 2)]);
                const float4 t0 = *reinterpret_cast<const float4 *>(&tws[twswz(k1, 4 * q)]);
                const float4 t1 = *reinterpret_cast<const float4 *>(&tws[twswz(k1, 4 * q + 2)]);
                y[0] = make_float2(y0.x, y0.y);
                y[1] = make_float2(y0.z, y0.w);
                y[2] = make_float2(y1.x, y1.y);
                y[3] = make_float2(y1.z, y1.w);
                t[0] = make_float2(t0.x, t0.y);
                t[1] = make_float2(t0.z, t0.w);
                t[2] = make_float2(t1.x, t1.y);
                t[3] = make_float2(t1.z, t1.w);
            }
            float z[8];
            float m = 0.f;
#pragma unroll
            for (int p = 0; p < 4; p++) {
                z[2 * p] = y[p].x * t[p].x - y[p].y * t[p].y;
                z[2 * p + 1] = y[p].x * t[p].y + y[p].y * t[p].x;
                m = absmax3(m, z[2 * p], z[2 * p + 1]);
            }
            m = max_over_groups(m);
            int ex = 0;
            (void)frexpf(m, &ex);
            const float s = ldexpf(1.0f, -ex);
            vec b[3];
            if constexpr (PREC == kSplitF16) {
                split8_f16_scaled(z, s, b);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) z[e] *= s;
                split8_bf16(z, b);
            }
            vec fc[3];
#pragma unroll
            for (int tt = 0; tt < TERMS; tt++) fc[tt] = opCs[(0 * TERMS + tt) * 64 + lane];
            const f32x4_t xr = split_mma<PREC>(fc, b, f32x4_t{});
#pragma unroll
            for (int tt = 0; tt < TERMS; tt++) fc[tt] = opCs[(1 * TERMS + tt) * 64 + lane];
            const f32x4_t xi = split_mma<PREC>(fc, b, f32x4_t{});
            const unsigned L = tile * 16 + line;
            if (L < nlines) {
                const float sc = ldexpf((float)a.scale, ex);
                float2 *__restrict__ p = static_cast<float2 *>(a.out) +
                                         tile_line_off<AL16>(tp, tile, line, a, true) +
                                         (long long)(k1 + 32 * q) * a.out_sa;
                const long long step = 128 * a.out_sa;
#pragma unroll
                for (int r = 0; r < 4; r++) st_stream(&p[r * step], make_float2(xr[r] * sc, xi[r] * sc));
            }
        }
    }
}

namespace {
long double cpi() { return 3.141592653589793238462643383279502884L; }
// exp(-2 pi i r / m) in long double, or its conjugate (conj: the backward transform's roots)
void wroot(long long r, long long m, long double &re, long double &im, bool conj) {
    const long double ang = 2.0L * cpi() * (long double)(r % m) / (long double)m;
    re = cosl(ang);
    im = conj ? sinl(ang) : -sinl(ang);
}
template <typename S> void split_store(float v, int terms, S *out, int stride) {
    float r = v;
    for (int t = 0; t < terms; t++) {
        const S p = (S)r;
        out[t * stride] = p;
        r -= (float)p;
    }
}
template <typename S> void fill_operands(std::vector<unsigned char> &bytes, int terms, bool conj) {
    // step A: row block rb (k1 = 16 rb + i), matrix mt (0: Re F32, 1: Im F32), lane i + 16 q,
    // element e: K index n1 = q + 4 e
    const size_t nA = 2 * 2 * terms * 64 * 8, nC = 2 * terms * 64 * 8;
    std::vector<S> A(nA), Cm(nC);
    for (int rb = 0; rb < 2; rb++)
        for (int mt = 0; mt < 2; mt++)
            for (int lane = 0; lane < 64; lane++)
                for (int e = 0; e < 8; e++) {
                    const int k1 = 16 * rb + (lane & 15), n1 = (lane >> 4) + 4 * e;
                    long double wr, wi;
                    wroot((long long)k1 * n1, 32, wr, wi, conj);
                    split_store<S>((float)(mt == 0 ? wr : wi), terms,
                                   &A[(((size_t)(rb * 2 + mt) * terms) * 64 + lane) * 8 + e], 64 * 8);
                }
    // step C: row block rc = 0 gives Re X[k1 + 32 k2], rc = 1 Im X, row i = lane & 15 holding
    // k2 = 4 (i % 4) + i / 4; element e of lane group q multiplies part e & 1 (0: real,
    // 1: imaginary) of z[n2 = 4 q + e / 2]
    for (int rc = 0; rc < 2; rc++)
        for (int lane = 0; lane < 64; lane++)
            for (int e = 0; e < 8; e++) {
                const int i = lane & 15, k2 = 4 * (i & 3) + (i >> 2), n2 = 4 * (lane >> 4) + (e >> 1), part = e & 1;
                long double wr, wi;
                wroot((long long)k2 * n2, 16, wr, wi, conj);
                const long double v = (rc == 0) ? (part == 0 ? wr : -wi) : (part == 0 ? wi : wr);
                split_store<S>((float)v, terms, &Cm[(((size_t)rc * terms) * 64 + lane) * 8 + e], 64 * 8);
            }
    const size_t off = bytes.size();
    bytes.resize(off + (nA + nC) * sizeof(S));
    std::memcpy(bytes.data() + off, A.data(), nA * sizeof(S));
    std::memcpy(bytes.data() + off + nA * sizeof(S), Cm.data(), nC * sizeof(S));
}
}  // namespace

std::vector<unsigned char> mfma512_operand_table(int prec, bool conj) {
    // W512^{n2 k1} stored as [k1][n2] (k1 < 32, n2 < 16), then the split DFT operands
    std::vector<unsigned char> bytes(512 * 2 * sizeof(float));
    float *w = reinterpret_cast<float *>(bytes.data());
    for (int k1 = 0; k1 < 32; k1++)
        for (int n2 = 0; n2 < 16; n2++) {
            long double re, im;
            wroot((long long)n2 * k1, 512, re, im, conj);
            w[2 * (16 * k1 + n2)] = (float)re;
            w[2 * (16 * k1 + n2) + 1] = (float)im;
        }
    if (prec == kSplitF16) fill_operands<_Float16>(bytes, 2, conj);
    else fill_operands<__bf16>(bytes, 3, conj);
    return bytes;
}

namespace {
template <int PREC, bool LCF, bool SCF> void launch_mfma512_al(const pass_args &a, long long ntiles, hipStream_t s) {
    if ((a.ncol % 16) == 0)
        fft512_mfma<PREC, LCF, SCF, true><<<dim3((unsigned)ntiles), dim3(1024), 0, s>>>(a);
    else
        fft512_mfma<PREC, LCF, SCF, false><<<dim3((unsigned)ntiles), dim3(1024), 0, s>>>(a);
}
template <int PREC> void launch_mfma512_prec(const pass_args &a, long long ntiles, hipStream_t s) {
    if (a.load_colfast) {
        if (a.store_colfast) launch_mfma512_al<PREC, true, true>(a, ntiles, s);
        else launch_mfma512_al<PREC, true, false>(a, ntiles, s);
    } else {
        if (a.store_colfast) launch_mfma512_al<PREC, false, true>(a, ntiles, s);
        else launch_mfma512_al<PREC, false, false>(a, ntiles, s);
    }
}
}  // namespace

hipError_t launch_mfma512(int prec, const pass_args &a, long long nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (nblocks > 0xffffffffLL) return hipErrorInvalidValue;
    if (prec == kSplitF16) launch_mfma512_prec<kSplitF16>(a, nblocks, s);
    else if (prec == kSplitBF16) launch_mfma512_prec<kSplitBF16>(a, nblocks, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace hmi
