// Host-side entry points of the gfx950 pass kernels (hmi_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "hmi_pass.h"

namespace hmi {

// Largest dynamic LDS the generic pass may request (gfx950: 160 KiB per workgroup).
constexpr int kGenericMaxLds = 160 * 1024;
// pass outputs above this size use non-temporal stores (larger than the MALL), and their loads
// too when a side of the pass walks the column axis (pass_args::stream)
constexpr long long kStreamBytes = 256LL << 20;
enum : int { kStreamStores = 1, kStreamLoadsStores = 3 };

using fast_launcher = hipError_t (*)(const pass_args &, long long nblocks, hipStream_t);

// A compiled uniform pass: length n, E elements per thread, T lines per block.
struct fast_kernel_info {
    bool dbl;
    int n, E, T, PS;
    int radix[8];
    int nstages;
    fast_launcher launch;
};

// wide = the pass walks the column axis on load or store (needs 128-byte column
// chunks); otherwise the variant with the fewest lines per block is returned.
const fast_kernel_info *find_fast_kernel(int n, bool dbl, bool wide);

hipError_t launch_generic(bool dbl, const pass_args &a, const generic_plan &g, long long nblocks, size_t lds_bytes,
                          hipStream_t s);

// Compile-time tile of the batched kernel for one length: lines per tile and LDS bytes
// for each precision, and the complex-unit size of its stage twiddle table.
struct batch_kernel_info {
    int n, lines_f32, lines_f64;
    int radix[8];
    int nstages;
    size_t lds_f32, lds_f64;
};
// Returns the tile id for length n (or -1: the runtime-plan tile runs it).
int find_batch_kernel(int n, batch_kernel_info *info);
// Register class of a tile id: 0 for the compile-time tiles with at most 12 elements per
// thread and the staged tiles, 1 for 14 to 16, 2 for the others and the runtime-plan tile (-1).  Each class has its
// own build of the batched kernel, so light tiles are not held to heavy tiles' register count.
constexpr int kBatchClasses = 3;
int batch_kernel_class(int kid);
// Compile date and time of the kernels (stamps the autotuner's cache: choices timed with other
// kernels are not reused).
const char *kernels_build_stamp();

// One phase (one axis) of a batch of mixed-size c2c transforms (see batch_args), for the
// work items of one register class.
hipError_t launch_batched(bool dbl, int cls, const batch_args &b, int nblocks, size_t lds_bytes, hipStream_t s);

// Fused x-y phase of a batch: each work item loads whole x-y planes (contiguous) of one
// transform into LDS (n1 rows of xy_row_stride(n0) elements per plane), runs the x-axis and
// y-axis FFTs there and stores the planes back, so the x and y axes cost one read and one write
// of the array instead of two each.  Transforms whose x and y lengths both have compile-time
// tiles (batch_xy_supported) of a register class up to kXyMaxClass take it.  Builds: 256 threads
// with planes of at most kXySmallBudget elements (one build, cls = kXyMaxClass, carrying every
// tile of the classes up to it), and for complex<float> 512 threads with planes up to
// kXyLargeBudget elements (tiles of the classes up to kXyLargeMaxClass; -1: no such build).
constexpr int kXySmallThreads = 256, kXyLargeThreads = 512;
constexpr int kXySmallBudget[2] = {10240, 5120};   // elements: 80 KiB of LDS (complex<float>, <double>)
constexpr int kXyLargeBudget[2] = {20480, 0};      // elements: 160 KiB of LDS
constexpr int kXyMaxClass[2] = {2, 0};             // register classes of the 256-thread build
constexpr int kXyLargeMaxClass[2] = {2, -1};       // register classes the 512-thread build runs
// Row stride (elements) of an x-y plane of rows of n0 elements in the LDS of the fused phase: odd,
// so that 32 lanes reading one position of 32 consecutive rows hit distinct banks.
constexpr int xy_row_stride(int n0) { return n0 | 1; }
bool batch_xy_supported(int kid);
hipError_t launch_batched_xy(bool dbl, bool large, int cls, const batch_args &b, int nblocks, size_t lds_bytes,
                             hipStream_t s);

// Column-block z phase of a batch (256 threads, one build: cls = kXyMaxClass, carrying every
// tile up to it): items (transform, block of zcol_width(n2, budget, n0 n1) consecutive (x, y)
// positions); b.xy_budget = LDS elements per item.
constexpr int zcol_width(int n2, int budget, int plane) {
    return budget / n2 < 1 ? 1 : (budget / n2 < plane ? budget / n2 : plane);
}
hipError_t launch_batched_zcol(bool dbl, int cls, const batch_args &b, int nblocks, size_t lds_bytes, hipStream_t s);

// Precision forms of the c2c passes: plain FP32/FP64 vector ALU, or the 512-point
// complex<float> pass on the matrix cores with split-precision operands.
enum : int { kPrecPlain = 0, kSplitF16 = 1, kSplitBF16 = 2 };
// 512-point complex<float> c2c pass on MFMA (nblocks = number of 16-line tiles, one
// 1024-thread block each); a.tw = the bytes of
// mfma512_operand_table(prec, conj): the W512 table and the split DFT-matrix operands of the
// forward transform, or their conjugates for the backward one (the kernel conjugates nothing).
hipError_t launch_mfma512(int prec, const pass_args &a, long long nblocks, hipStream_t s);
std::vector<unsigned char> mfma512_operand_table(int prec, bool conj);

// Permutation of a box between two layouts (heFFTe's in-node reshape).
// a = source-fastest axis, b = destination-fastest axis (or any other axis
// when both sides are fastest along a, same_fast = 1), c = remaining axis.
struct permute_args {
    long long n_a, n_b, n_c, nbatch;
    long long si_a, si_b, si_c, bs_in;
    long long so_a, so_b, so_c, bs_out;
    long long ntile_a, ntile_b;
    int same_fast;
};

hipError_t launch_permute(int elem_bytes, const void *src, void *dst, const permute_args &p, hipStream_t s);

// Static device workspace of the no-allocation entry point (hmi_scratch.hip): kScratchBytes per
// GPU, held until the process ends.  scratch_address() returns its device address on the current
// device (resolved with hipGetSymbolAddress on the first call per device; throws on failure).
constexpr size_t kScratchBytes = 64ull << 20;
void *scratch_address();

}  // namespace hmi
