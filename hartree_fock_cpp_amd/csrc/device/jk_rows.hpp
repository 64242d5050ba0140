// jk_rows.hpp -- row-block decomposition of the stored unique ERI array for
// the one-pass J/K kernel (k_jk_rows in jk_stored.hip), shared with the CPU
// emulation used by the tests.
//
// Rows of the unique array are function pairs IJ = i(i+1)/2 + j, j <= i
// (include/indexing.hpp:42-52 of the reference); row IJ holds the integrals
// (ij|kl) for KL = 0..IJ.  A row block is up to R consecutive rows of one i,
// j = j0 .. min(i, j0 + R - 1), so row i has i/R + 1 blocks; blocks are
// numbered i-major, ascending.
#pragma once
#include <stdint.h>

#include "md_core.hpp"

namespace hfx {

// Blocks of rows i' < i: sum_{x < i} (x / R + 1).
__host__ __device__ inline long long jkr_blocks_before(int i, int R) {
    const long long q = i / R;
    return (long long)i + (long long)R * q * (q - 1) / 2 + q * ((long long)i - q * R);
}

// Block b (ascending numbering) -> (i, j0).
__host__ __device__ inline void jkr_block(long long b, int K, int R, int& i, int& j0) {
    int lo = 0, hi = K - 1;   // the largest i with jkr_blocks_before(i) <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jkr_blocks_before(mid, R) <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    i = lo;
    j0 = (int)(b - jkr_blocks_before(lo, R)) * R;
}

// Degeneracy weight of one stored integral (ij|kl) among the 8 images of
// its index quartet: 1/2 for each of i == j, k == l, IJ == KL.
__host__ __device__ inline double jkr_weight(int i, int j, int k, int l) {
    return (i == j ? 0.5 : 1.0) * (k == l ? 0.5 : 1.0) * ((i == k && j == l) ? 0.5 : 1.0);
}

}  // namespace hfx
