// Host-side entry points of the gfx950 pass kernels (hmi_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hmi_pass.h"

namespace hmi {

// Largest dynamic LDS the generic pass may request (gfx950: 160 KiB per workgroup).
constexpr int kGenericMaxLds = 160 * 1024;

using fast_launcher = hipError_t (*)(const pass_args &, long long nblocks, hipStream_t);

// A compiled uniform pass: length n, E elements per thread, T lines per block.
struct fast_kernel_info {
    bool dbl;
    int n, E, T;
    int radix[8];
    int nstages;
    fast_launcher launch;
};

// wide = the pass walks the column axis on load or store (needs 128-byte column
// chunks); otherwise the variant with the fewest lines per block is returned.
const fast_kernel_info *find_fast_kernel(int n, bool dbl, bool wide);

hipError_t launch_generic(bool dbl, const pass_args &a, const generic_plan &g, long long nblocks, size_t lds_bytes,
                          hipStream_t s);

// One phase of a batch of mixed-size c2c transforms (see batch_args).
hipError_t launch_batched(bool dbl, const batch_args &b, int nblocks, size_t lds_bytes, hipStream_t s);

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

}  // namespace hmi
