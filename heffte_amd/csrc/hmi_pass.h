// Host/device description of one FFT pass: a batched 1D c2c transform of
// length n along one axis of a 3D array, reading one layout and writing another.
//
// A pass sees its data as lines (the FFT axis, stride sa) indexed by a column
// (stride sc), a third axis b1 (stride sb1) and the batch b2 (stride sb2), with
// independent strides on the input and on the output side.  The layout
// permutation heFFTe performs with reshape3d_transpose
// (include/heffte_reshape3d.h:443-487) is therefore folded into the strides of
// the load or of the store of a pass, and the scale of apply_scale
// (include/heffte_fft3d.h:662-678) into its store.
#pragma once
#include <cstdint>

namespace hmi {

enum : int { kMaxStages = 24 };

// Pass modes of the uniform kernel.  The real-data modes run a complex FFT of
// half the line length: kModeR2C reads 2N reals as N complex pairs and unpacks
// the half spectrum (N + 1 bins) after the FFT; kModeC2R packs the N + 1 bins
// into N complex values before an inverse FFT and writes 2N reals.
//
// The real-to-real modes compute one axis of rtransform<cos/sin> through an
// N-point complex FFT (Makhoul's reordering):
//   kModeDCT2 (forward cos): v[n] = x[2n], v[N-1-n] = x[2n+1]; Y[k] = 2 Re(e^{-i pi k/2N} V[k])
//   kModeDCT3 (backward cos): W[k] = e^{i pi k/2N} (x[k] - i x[N-k]), w = IDFT(W);
//                             y[2m] = 2 Re w[m], y[2m+1] = 2 Re w[N-1-m]
//   kModeDST2 / kModeDST3: the same with the sine identities
//     DST-II(x)[k] = DCT-II((-1)^n x[n])[N-1-k],  DST-III(x)[n] = (-1)^n DCT-III(x[N-1-.])[n]
//
// DCT-I (rtransform<cos1>) runs the kernel's FFT length F = 2(N-1) on the even extension
// E[j] = x[j] (j <= N-1), x[F-j] (j > N-1), imaginary parts zero, and stores y[k] = Re X[k]
// for k < N (store_len = N): kModeDCT1 forward (REDFT00), kModeDCT1B backward (2 REDFT00).
enum : int {
    kModeC2C = 0, kModeR2C = 1, kModeC2R = 2, kModeDCT2 = 3, kModeDCT3 = 4, kModeDST2 = 5, kModeDST3 = 6,
    kModeDCT1 = 7, kModeDCT1B = 8
};

struct pass_args {
    const void *in;
    void *out;
    const void *tw;  // twiddle table of the pass (real2 of the pass precision)
    long long in_sa, in_sc, in_sb1, in_sb2;
    long long out_sa, out_sc, out_sb1, out_sb2;
    int ncol;      // number of columns
    int nb1;       // extent of the third axis
    int nbatch;    // number of transforms in the batch
    int nlines;    // ncol * nb1 * nbatch: lines are numbered column-fastest, then b1, then batch
    int load_colfast;   // 1: consecutive threads walk the column axis on load
    int store_colfast;  // 1: ... on store
    int conj;      // 1: backward transform (conjugate on load and on store)
    int in_real;   // 1: input elements are real scalars (imaginary part 0)
    int out_real;  // 1: store only the real part
    int store_len; // positions >= store_len along the FFT axis are not stored (r2c half spectrum)
    int herm_load; // 1: positions > n/2 are read as conj(x[n - pos]) (c2r Hermitian extension)
    int mode;      // kModeC2C, kModeR2C or kModeC2R (uniform kernel only)
    int stream;    // non-temporal data accesses (outputs larger than the MALL): 0 none,
                   // kStreamStores (1) stores only, kStreamLoadsStores (3) loads and stores
    const void *mul;  // non-null: every loaded element is first multiplied by the element of this
                      // array at the same offset (same layout and type as in; plain c2c loads only)
    double scale;  // multiplies every stored value (cast to the pass precision)
    // First step of an axis split in two passes (n4 = n * n2 points, plain c2c store only):
    // non-null tw4 = table of w_n4^m (m < n4); the value stored at position k of a line is first
    // multiplied by w_n4^(j k), j = the line's index along dimension tw4_dim (0 column, 1 b1,
    // 2 batch), which runs over the n2 interleaved sub-sequences.
    const void *tw4;
    int tw4_dim;
};

// Runtime radix plan used by the generic (any length) pass.
struct generic_plan {
    int n;
    int nstages;
    int lines;                 // lines per block
    int radix[kMaxStages];
    int twoff[kMaxStages];     // offset (complex units) of stage twiddles w_{Ls R}^{p k}
    int rootoff[kMaxStages];   // offset of w_R^j, j < R (used by radices without a hard-coded butterfly)
    int postoff;               // offset of the real-data twiddles (modes other than kModeC2C)
};

// Batched mixed-size transforms (one phase = one axis of every transform).
// Every transform is a c2c box in order (0, 1, 2) at its own element offset.
struct batch_xform {
    long long in_off, out_off;  // element offsets from the batch's input / output base
    int n[3];
    int size_id[3];             // index of the radix plan of each axis length
};

struct batch_item {
    int xform;  // transform
    int tile;   // tile of lines along the phase axis
};

struct batch_args {
    const void *in;
    void *out;
    const void *tw;               // concatenated twiddle tables of every length
    const batch_xform *xforms;
    const batch_item *items;
    const generic_plan *plans;    // one per distinct length
    const int *twoff;             // table offset (complex units) of each length
    const int *kid;               // compile-time tile of each length (-1: runtime radix plan)
    int nitems;
    int axis;       // 0, 1 or 2
    int from_in;    // 1: read the input arrays (first phase), 0: read the output arrays in place
    int conj;       // backward transform
    int scaling;    // 0 none, 1 full (1/N of each transform), 2 symmetric; applied by the last phase
    int xy_budget;  // fused x-y phase: elements of the planes one work item holds in LDS (a work
                    // item of a transform takes max(1, xy_budget / (n0 n1)) consecutive z-planes)
};

}  // namespace hmi
