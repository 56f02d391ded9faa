// bounds.h -- host-side predicates on operand shapes, kept free of HIP
// headers so that the host tests compile them with g++ alone.
#pragma once

#include <cstddef>

namespace swift {

// B can be read through a buffer resource with 24-bit row offsets: fewer than
// 2^24 rows, fewer than 2^31 bytes, and a row stride below 2^24 bytes (the
// offset is __umul24(row, stride), which keeps only 24 bits of each factor).
inline bool buffer_ok(long long nrows, long long ldb, size_t elem) {
    return nrows < (1LL << 24) && ldb * (long long)elem < (1LL << 24) &&
           nrows * ldb * (long long)elem < (1LL << 31);
}

}  // namespace swift
