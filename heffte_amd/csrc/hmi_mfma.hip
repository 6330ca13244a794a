// 512-point c2c pass of complex<float> data computed on the matrix cores (gfx950 MFMA).
//
// n = 512 = 32 x 16 (four-step): with n = 16 n1 + n2 and k = k1 + 32 k2,
//   Y[k1][n2] = sum_n1 W32^{n1 k1} x[16 n1 + n2]            (step A, 32-point DFTs)
//   X[k1 + 32 k2] = sum_n2 W16^{n2 k2} (W512^{n2 k1} Y[k1][n2])   (twiddle, step C, 16-point DFTs)
// Steps A and C are real GEMMs on the block form [Re F, -Im F; Im F, Re F] x [Re v; Im v]
// with v_mfma_f32_32x32x16_{f16,bf16} (FP32 accumulation).  Accuracy comes from splitting
// every FP32 operand into low-precision terms:
//   kSplitF16:  x = hi + lo (two FP16 terms), products hi*hi + hi*lo + lo*hi
//   kSplitBF16: x = hi + mid + lo (three BF16 terms), products hh + hm + mh + hl + lh + mm
// and from a power-of-two scale per line (pencil) that puts the line maximum just below 1 so
// both terms stay in the normal range of the 16-bit format; the scale is undone on store.
//
// MFMA 32x32x16 lane layout (gfx950): A[i][k] in lane i + 32 (k / 8), element k % 8;
// B[k][j] in lane j + 32 (k / 8), element k % 8; C[i][j] in lane j + 32 ((i / 4) % 2),
// element (i % 4) + 4 (i / 8).
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
typedef float f32x16_t __attribute__((ext_vector_type(16)));

template <int PREC> struct split_traits;

template <> struct split_traits<kSplitF16> {
    static constexpr int terms = 2;
    using vec = half8_t;
    using scalar = _Float16;
    __device__ static f32x16_t mma(vec a, vec b, f32x16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct split_traits<kSplitBF16> {
    static constexpr int terms = 3;
    using vec = bf8_t;
    using scalar = __bf16;
    __device__ static f32x16_t mma(vec a, vec b, f32x16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// split 8 FP32 values into TERMS low-precision vectors
template <int PREC>
__device__ __forceinline__ void split8(const float (&x)[8], typename split_traits<PREC>::vec (&out)[3]) {
    using S = split_traits<PREC>;
    using scal = typename S::scalar;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        float r = x[e];
#pragma unroll
        for (int t = 0; t < S::terms; t++) {
            const scal p = (scal)r;
            out[t][e] = p;
            r -= (float)p;
        }
    }
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

// acc += A * B through the split products
template <int PREC>
__device__ __forceinline__ f32x16_t split_mma(const typename split_traits<PREC>::vec (&a)[3],
                                              const typename split_traits<PREC>::vec (&b)[3], f32x16_t acc) {
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

// Offset of line tile * 16 + l.  When the column count is a multiple of 16 the tile's lines
// share (b1, batch) and the division runs once per tile on uniform values.
__device__ __forceinline__ long long tile_line_off(unsigned tile, int l, const pass_args &a, bool out) {
    const long long sc = out ? a.out_sc : a.in_sc;
    if ((a.ncol & 15) == 0) {
        const unsigned L0 = tile * 16;
        const unsigned q = L0 / (unsigned)a.ncol;
        const unsigned c0 = L0 - q * (unsigned)a.ncol;
        const unsigned b2 = q / (unsigned)a.nb1;
        const unsigned b1 = q - b2 * (unsigned)a.nb1;
        const long long base = out ? (long long)c0 * a.out_sc + (long long)b1 * a.out_sb1 + (long long)b2 * a.out_sb2
                                   : (long long)c0 * a.in_sc + (long long)b1 * a.in_sb1 + (long long)b2 * a.in_sb2;
        return base + (long long)l * sc;
    }
    return line_off(tile * 16 + l, a, out);
}

}  // namespace

// One tile = 16 lines x 512 points for 1024 threads (16 waves).  a.tw holds the W512^r table
// (512 complex<float>) followed by the A operands of both GEMMs, already split and laid out
// per lane (mfma512_operand_table): step A [m][ks][term][lane][8], step C [ks][term][lane][8].
// LDS holds the tile in the layout each GEMM reads with contiguous K: x[16 n1 + n2] at
// (line, n2, n1) for step A, Y[k1][n2] at (line, k1, n2) for step C (odd row pitches spread
// the rows over the banks).
//
// Persistent form: one 1024-thread block per CU walks the tiles blockIdx.x, blockIdx.x +
// gridDim.x, ...  The step-C operands and W512 are copied to LDS once; FP16 keeps each wave's
// step-A operands in registers, BF16 reads them from the cached table.  The global loads of the
// next tile are in flight while the current tile runs step A.  Both tile layouts, the tables
// and the per-line maxima (by tile parity: one partial per wave and line) take 147,344 B (FP16)
// or 149,392 B (BF16) of LDS: one block per CU.
// Per tile: issue loads(next) | step A | barrier | Y -> LDS, next -> LDS | barrier | step C, store.
template <int PREC>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
fft512_mfma_persistent(pass_args a, unsigned ntiles) {
    constexpr int T = 16, P = 64, E = 8;
    using S = split_traits<PREC>;
    using vec = typename S::vec;
    constexpr int TERMS = S::terms;
    // FP16 keeps each wave's step-A operands in registers; BF16 (three terms, too many
    // registers) reads them from the cached table every tile
    constexpr bool kOpARegs = (TERMS == 2);
    constexpr int RA = 33, RLA = 16 * RA + 1;
    constexpr int RC = 17, RLC = 32 * RC + 1;
    constexpr size_t kA = 0, kC = kA + sizeof(float2) * T * RLA, kMax = kC + sizeof(float2) * T * RLC;
    constexpr size_t kW = kMax + sizeof(float) * 2 * T * T;
    constexpr size_t kOpC = kW + sizeof(float2) * 466;
    constexpr size_t kBytes = kOpC + sizeof(vec) * 2 * TERMS * 64;
    __shared__ __align__(16) unsigned char lds[kBytes];
    float2 *ldsA = reinterpret_cast<float2 *>(lds + kA);
    float2 *ldsC = reinterpret_cast<float2 *>(lds + kC);
    float(*pmax)[T][T] = reinterpret_cast<float(*)[T][T]>(lds + kMax);
    float2 *w512s = reinterpret_cast<float2 *>(lds + kW);  // W512^r for r = n2 k1 <= 15 * 31
    vec *opCs = reinterpret_cast<vec *>(lds + kOpC);

    const unsigned nlines = (unsigned)a.nlines;
    const bool lcf = a.load_colfast != 0, scf = a.store_colfast != 0;
    const float2 *__restrict__ w512 = static_cast<const float2 *>(a.tw);
    const vec *__restrict__ opA = reinterpret_cast<const vec *>(w512 + 512);
    const vec *__restrict__ opC = opA + 2 * 4 * TERMS * 64;
    {
        const int tid = threadIdx.x;
        for (int i = tid; i < 2 * TERMS * 64; i += 1024) opCs[i] = opC[i];
        if (tid < 466) w512s[tid] = w512[tid];
    }
    float *ldsCf = reinterpret_cast<float *>(ldsC);
    const float2 *__restrict__ src = static_cast<const float2 *>(a.in);
    const long long in_step = (long long)P * a.in_sa;

    // The thread index is passed through an empty asm at each use so that the per-lane
    // addresses derived from it are recomputed in the loop instead of being kept live
    // across it (128 VGPRs at this occupancy).
    auto thread = [] {
        int v = threadIdx.x;
        asm volatile("" : "+v"(v));
        return v;
    };
    // all eight loads of a tile in flight (an invalid line reads line 0)
    auto issue = [&](unsigned tile, float2 (&x)[E]) {
        const int tid = thread();
        const int l = lcf ? (tid & 15) : (tid >> 6), t = lcf ? (tid >> 4) : (tid & 63);
        const unsigned L = tile * T + l;
        const float2 *p = src + (L < nlines ? tile_line_off(tile, l, a, false) : 0) + (long long)t * a.in_sa;
#pragma unroll
        for (int e = 0; e < E; e++) x[e] = ld_stream(&p[e * in_step]);
    };
    // sign, step-A layout (x[16 n1 + n2] at row n2, column n1; pos = t + 64 e has
    // n2 = t & 15, n1 = (t >> 4) + 4 e), per-line maximum: a full line per wave (along the axis)
    // or lanes l, l+16, l+32, l+48 of every wave (column-fast), one partial per wave
    auto deposit = [&](float2 (&x)[E], int par) {
        const int tid = thread();
        const int lane = tid & 63, wave = tid >> 6;
        const int l = lcf ? (tid & 15) : (tid >> 6), t = lcf ? (tid >> 4) : (tid & 63);
        // a line past the end holds line 0's values: its results are never stored
        float2 *row = &ldsA[l * RLA + (t & 15) * RA + (t >> 4)];
        float m = 0.f;
#pragma unroll
        for (int e = 0; e < E; e++) {
            m = fmaxf(m, fmaxf(fabsf(x[e].x), fabsf(x[e].y)));
            row[4 * e] = x[e];
        }
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        if (!lcf) {
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if (lane == 0) pmax[par][l][0] = m;
        } else if (lane < 16) {
            pmax[par][l][wave] = m;
        }
    };
    auto line_max = [&](int par, int line) {
        if (!lcf) return pmax[par][line][0];
        float m = 0.f;
#pragma unroll
        for (int w = 0; w < T; w++) m = fmaxf(m, pmax[par][line][w]);
        return m;
    };

    // FP16: the wave's step-A operands stay in registers for the whole launch
    vec ra[4][TERMS];
    if constexpr (kOpARegs) {
        const int lane = threadIdx.x & 63, m = (threadIdx.x >> 6) & 1;
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int tt = 0; tt < TERMS; tt++) ra[ks][tt] = opA[((m * 4 + ks) * TERMS + tt) * 64 + lane];
    }
    // one tile t: loads of tile t + G into xn | step A | barrier | Y -> LDS, tile t + G -> LDS |
    // barrier | step C, store
    auto run_tile = [&](unsigned tile, int par, float2 (&xn)[E]) {
        const unsigned next = tile + gridDim.x;
        const bool has_next = next < ntiles;
        if (has_next) issue(next, xn);

        // ---- step A: wave -> (column block b, real / imaginary rows m); column j = (n2, line)
        f32x16_t acc = {};
        {
            const int tid = thread();
            const int lane = tid & 63, wave = tid >> 6, j = lane & 31, h = lane >> 5;
            const int m = wave & 1, b = wave >> 1;
            const int n2 = j & 15, line = 2 * b + (j >> 4);
            int ex = 0;
            (void)frexpf(line_max(par, line), &ex);
            const float s = ldexpf(1.0f, -ex);
            const float2 *rowp = &ldsA[line * RLA + n2 * RA + 8 * h];
            const vec *opm = &opA[m * 4 * TERMS * 64 + lane];
#pragma unroll
            for (int half = 0; half < 2; half++) {
                float re[8], im[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const float2 x = rowp[16 * half + q];
                    if constexpr (PREC == kSplitF16) {
                        re[q] = x.x;
                        im[q] = x.y;
                    } else {
                        re[q] = x.x * s;
                        im[q] = x.y * s;
                    }
                }
#pragma unroll
                for (int part = 0; part < 2; part++) {
                    const int ks = half + 2 * part;
                    vec A[3], B[3];
#pragma unroll
                    for (int tt = 0; tt < TERMS; tt++) {
                        if constexpr (kOpARegs) A[tt] = ra[ks][tt];
                        else A[tt] = opm[(ks * TERMS + tt) * 64];
                    }
                    if constexpr (PREC == kSplitF16) split8_f16_scaled(part == 0 ? re : im, s, B);
                    else split8<PREC>(part == 0 ? re : im, B);
                    acc = split_mma<PREC>(A, B, acc);
                }
            }
        }
        __syncthreads();  // step A has read the tile; step C of the previous tile is done
        {
            const int tid = thread();
            const int lane = tid & 63, wave = tid >> 6, j = lane & 31, h = lane >> 5;
            const int m = wave & 1, b = wave >> 1;
            const int n2 = j & 15, line = 2 * b + (j >> 4);
            float *yp = &ldsCf[2 * (line * RLC + 4 * h * RC + n2) + m];
#pragma unroll
            for (int e = 0; e < 16; e++) yp[2 * RC * ((e & 3) + 8 * (e >> 2))] = acc[e];  // k1 = that + 4 h
        }
        if (has_next) deposit(xn, par ^ 1);
        __syncthreads();

        // ---- twiddle + step C: wave -> 32 columns (k1, line)
        {
            const int tid = thread();
            const int lane = tid & 63, wave = tid >> 6, j = lane & 31, h = lane >> 5;
            const int line = scf ? (j & 15) : wave, k1 = scf ? 2 * wave + (j >> 4) : j;
            acc = f32x16_t{};
            float zr[8], zi[8];
            const float2 *yp = &ldsC[line * RLC + k1 * RC + 8 * h];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float2 y = yp[e];
                // W512^{n2 k1}, n2 = 8 h + e
                const float2 w = w512s[(8 * h + e) * k1];
                zr[e] = y.x * w.x - y.y * w.y;
                zi[e] = y.x * w.y + y.y * w.x;
            }
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                vec A[3], B[3];
#pragma unroll
                for (int tt = 0; tt < TERMS; tt++)
                    A[tt] = opCs[(ks * TERMS + tt) * 64 + lane];
                if constexpr (PREC == kSplitF16) split8_f16_scaled(ks == 0 ? zr : zi, 1.0f, B);
                else split8<PREC>(ks == 0 ? zr : zi, B);
                acc = split_mma<PREC>(A, B, acc);
            }
            // store: lane holds X[k1 + 32 k2], k2 = (e & 3) + 8 (e >> 2) + 4 h, e < 8
            const unsigned L = tile * T + line;
            if (L < nlines) {
                int ex = 0;
                (void)frexpf(line_max(par, line), &ex);
                const float s = ldexpf((float)a.scale, ex);
                float2 *__restrict__ p = static_cast<float2 *>(a.out) + tile_line_off(tile, line, a, true) +
                                         (long long)(k1 + 128 * h) * a.out_sa;
                const long long step = 32 * a.out_sa;
#pragma unroll
                for (int e = 0; e < 8; e++)
                    st_stream(&p[((e & 3) + 8 * (e >> 2)) * step], make_float2(acc[e] * s, acc[e + 8] * s));
            }
        }
    };

    unsigned tile = blockIdx.x;
    {
        float2 x[E];
        issue(tile, x);
        deposit(x, 0);
    }
    __syncthreads();
    for (int it = 0; tile < ntiles; it++, tile += gridDim.x) {
        float2 xn[E];
        run_tile(tile, it & 1, xn);
    }
}

namespace {
long double cpi() { return 3.141592653589793238462643383279502884L; }
// exp(-2 pi i r / m) in long double
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
    // step A: A[i][k] of [Re F32, -Im F32; Im F32, Re F32], rows i = k1 (m = 0: real rows,
    // m = 1: imaginary rows), K-step ks covers k = 16 ks + 8 (lane / 32) + e
    const size_t nA = 2 * 4 * terms * 64 * 8, nC = 2 * terms * 64 * 8;
    std::vector<S> A(nA), Cm(nC);
    for (int m = 0; m < 2; m++)
        for (int ks = 0; ks < 4; ks++)
            for (int lane = 0; lane < 64; lane++)
                for (int e = 0; e < 8; e++) {
                    const int i = lane & 31, n1 = 16 * (ks & 1) + 8 * (lane >> 5) + e;
                    long double wr, wi;
                    wroot((long long)i * n1, 32, wr, wi, conj);
                    const long double v = (m == 0) ? (ks < 2 ? wr : -wi) : (ks < 2 ? wi : wr);
                    split_store<S>((float)v, terms, &A[(((size_t)(m * 4 + ks) * terms) * 64 + lane) * 8 + e], 64 * 8);
                }
    // step C: rows i < 16 real part of X[k1 + 32 i], i >= 16 imaginary part of k2 = i - 16;
    // K-step 0 multiplies Re z, K-step 1 Im z, n2 = 8 (lane / 32) + e
    for (int ks = 0; ks < 2; ks++)
        for (int lane = 0; lane < 64; lane++)
            for (int e = 0; e < 8; e++) {
                const int i = lane & 31, n2 = 8 * (lane >> 5) + e;
                long double wr, wi;
                wroot((long long)(i & 15) * n2, 16, wr, wi, conj);
                const long double v = (i < 16) ? (ks == 0 ? wr : -wi) : (ks == 0 ? wi : wr);
                split_store<S>((float)v, terms, &Cm[(((size_t)ks * terms) * 64 + lane) * 8 + e], 64 * 8);
            }
    const size_t off = bytes.size();
    bytes.resize(off + (nA + nC) * sizeof(S));
    std::memcpy(bytes.data() + off, A.data(), nA * sizeof(S));
    std::memcpy(bytes.data() + off + nA * sizeof(S), Cm.data(), nC * sizeof(S));
}
}  // namespace

std::vector<unsigned char> mfma512_operand_table(int prec, bool conj) {
    std::vector<unsigned char> bytes(512 * 2 * sizeof(float));
    float *w = reinterpret_cast<float *>(bytes.data());
    for (int r = 0; r < 512; r++) {
        long double re, im;
        wroot(r, 512, re, im, conj);
        w[2 * r] = (float)re;
        w[2 * r + 1] = (float)im;
    }
    if (prec == kSplitF16) fill_operands<_Float16>(bytes, 2, conj);
    else fill_operands<__bf16>(bytes, 3, conj);
    return bytes;
}

hipError_t launch_mfma512(int prec, const pass_args &a, long long nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (prec == kSplitF16) {
        // one block per CU, never more blocks than tiles
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        const long long grid = std::min<long long>(nblocks, std::max(1, cus));
        fft512_mfma_persistent<kSplitF16><<<dim3((unsigned)grid), dim3(1024), 0, s>>>(a, (unsigned)nblocks);
    } else if (prec == kSplitBF16) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        const long long grid = std::min<long long>(nblocks, std::max(1, cus));
        fft512_mfma_persistent<kSplitBF16><<<dim3((unsigned)grid), dim3(1024), 0, s>>>(a, (unsigned)nblocks);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hmi
