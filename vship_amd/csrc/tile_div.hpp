// tile_div.hpp - exact division by a small divisor through a round-up
// reciprocal, for the front-end's tile index -> (x, y) (plain C++: the host
// computes the reciprocal, the kernels apply it with scalar instructions).
#pragma once

#include <stdint.h>

namespace vship {

// magic = ceil(2^(31 + l) / d) with 2^(l-1) < d <= 2^l (l >= 0) lies below
// 2^32, and magic * d = 2^(31 + l) + e with 0 <= e < d <= 2^l. Then
// n * magic / 2^(31 + l) = n / d + n e / (d 2^(31 + l)), whose second term is
// below 1 / d for every n below 2^31: it cannot reach the next multiple of
// 1 / d, so (n * magic) >> (31 + l) is floor(n / d). d lies in [1, 2^30].
inline void tile_div_magic(int d, uint32_t* magic, int32_t* shift)
{
    int l = 0;
    while ((1u << l) < (uint32_t)d) l++;
    *magic = (uint32_t)((((uint64_t)1 << (31 + l)) + (uint64_t)d - 1) / (uint64_t)d);
    *shift = 31 + l;
}

inline uint32_t tile_div_apply(uint32_t n, uint32_t magic, int32_t shift)
{
    return (uint32_t)(((uint64_t)n * magic) >> shift);
}

}  // namespace vship
