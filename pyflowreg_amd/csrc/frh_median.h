// Median of 25 samples as a pruned sorting network (shared by the solver's
// median kernels and the differentiable path).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace frh {

// The 5x5 median of scipy.ndimage.median_filter (core/optical_flow.py:
// 607-609) as a selection network: the median is element 12 of
// Batcher's odd-even merge sort of the 25 window samples; only the 113
// comparators that can reach output 12 are kept, and of those only the
// min or max that is used later (202 min/max operations).  Selection is
// exact, so the result equals scipy's.
struct MedCmp {
    int a, b, lo, hi;
};
static constexpr MedCmp kMed25[] = {
    {0, 1, 1, 1},
    {2, 3, 1, 1},
    {4, 5, 1, 1},
    {6, 7, 1, 1},
    {8, 9, 1, 1},
    {10, 11, 1, 1},
    {12, 13, 1, 1},
    {14, 15, 1, 1},
    {16, 17, 1, 1},
    {18, 19, 1, 1},
    {20, 21, 1, 1},
    {22, 23, 1, 1},
    {0, 2, 1, 1},
    {1, 3, 1, 1},
    {4, 6, 1, 1},
    {5, 7, 1, 1},
    {8, 10, 1, 1},
    {9, 11, 1, 1},
    {12, 14, 1, 1},
    {13, 15, 1, 1},
    {16, 18, 1, 1},
    {17, 19, 1, 1},
    {20, 22, 1, 1},
    {21, 23, 1, 1},
    {1, 2, 1, 1},
    {5, 6, 1, 1},
    {9, 10, 1, 1},
    {13, 14, 1, 1},
    {17, 18, 1, 1},
    {21, 22, 1, 1},
    {0, 4, 1, 1},
    {1, 5, 1, 1},
    {2, 6, 1, 1},
    {3, 7, 1, 1},
    {8, 12, 1, 1},
    {9, 13, 1, 1},
    {10, 14, 1, 1},
    {11, 15, 1, 1},
    {16, 20, 1, 1},
    {17, 21, 1, 1},
    {18, 22, 1, 1},
    {19, 23, 1, 1},
    {2, 4, 1, 1},
    {3, 5, 1, 1},
    {10, 12, 1, 1},
    {11, 13, 1, 1},
    {18, 20, 1, 1},
    {19, 21, 1, 1},
    {1, 2, 1, 1},
    {3, 4, 1, 1},
    {5, 6, 1, 1},
    {9, 10, 1, 1},
    {11, 12, 1, 1},
    {13, 14, 1, 1},
    {17, 18, 1, 1},
    {19, 20, 1, 1},
    {21, 22, 1, 1},
    {0, 8, 1, 1},
    {1, 9, 1, 1},
    {2, 10, 1, 1},
    {3, 11, 1, 1},
    {4, 12, 1, 1},
    {5, 13, 1, 1},
    {6, 14, 1, 1},
    {7, 15, 1, 0},
    {16, 24, 1, 1},
    {4, 8, 1, 1},
    {5, 9, 1, 1},
    {6, 10, 1, 1},
    {7, 11, 1, 1},
    {20, 24, 1, 1},
    {2, 4, 1, 1},
    {3, 5, 1, 1},
    {6, 8, 1, 1},
    {7, 9, 1, 1},
    {10, 12, 1, 1},
    {11, 13, 1, 1},
    {18, 20, 1, 1},
    {19, 21, 1, 1},
    {22, 24, 1, 1},
    {1, 2, 1, 1},
    {3, 4, 1, 1},
    {5, 6, 1, 1},
    {7, 8, 1, 1},
    {9, 10, 1, 1},
    {11, 12, 1, 1},
    {13, 14, 1, 0},
    {17, 18, 1, 1},
    {19, 20, 1, 1},
    {21, 22, 1, 1},
    {23, 24, 1, 1},
    {0, 16, 0, 1},
    {1, 17, 0, 1},
    {2, 18, 0, 1},
    {3, 19, 0, 1},
    {4, 20, 0, 1},
    {5, 21, 0, 1},
    {6, 22, 1, 0},
    {7, 23, 1, 0},
    {8, 24, 1, 0},
    {8, 16, 0, 1},
    {9, 17, 0, 1},
    {10, 18, 1, 0},
    {11, 19, 1, 0},
    {12, 20, 1, 0},
    {13, 21, 1, 0},
    {6, 10, 0, 1},
    {7, 11, 0, 1},
    {12, 16, 1, 0},
    {13, 17, 1, 0},
    {10, 12, 0, 1},
    {11, 13, 1, 0},
    {11, 12, 0, 1}
};

template <int I>
__device__ __forceinline__ void med25_step(double (&v)[25])
{
    constexpr MedCmp c = kMed25[I];
    double x = v[c.a], y = v[c.b];
    if constexpr (c.lo != 0) v[c.a] = __builtin_fmin(x, y);
    if constexpr (c.hi != 0) v[c.b] = __builtin_fmax(x, y);
}

template <int... I>
__device__ __forceinline__ double median25_net(double (&v)[25], std::integer_sequence<int, I...>)
{
    (med25_step<I>(v), ...);
    return v[12];
}

__device__ inline double median25(double (&a)[25])
{
    return median25_net(a, std::make_integer_sequence<int, sizeof(kMed25) / sizeof(MedCmp)>{});
}

}  // namespace frh
