// rowview.h -- device building blocks of the row-major Swift kernels (gfx950),
// shared by the single-matrix launches (spmm.hip) and the batched persistent
// kernel (batched.hip).
#pragma once

#include "bounds.h"
#include "swift_internal.h"

#include <type_traits>
#include <utility>
#include <vector>

namespace swift {

// ---------------------------------------------------------- type helpers --

// Elements per lane per B load: 16-byte loads for FP32/FP64; 8-byte loads of
// four 16-bit values, which keeps the lane shape (and so the per-entry
// instruction count) of FP32 while moving half the bytes (SWIFT_NARROW_VEC=8
// builds 16-byte loads of eight).
#ifndef SWIFT_NARROW_VEC
#define SWIFT_NARROW_VEC 4
#endif
template <class T> struct Vec16;
template <> struct Vec16<float> { static constexpr int N = 4; };
template <> struct Vec16<double> { static constexpr int N = 2; };
template <> struct Vec16<__hip_bfloat16> { static constexpr int N = SWIFT_NARROW_VEC; };
template <> struct Vec16<__half> { static constexpr int N = SWIFT_NARROW_VEC; };

__device__ __forceinline__ float to_acc(float x) { return x; }
__device__ __forceinline__ double to_acc(double x) { return x; }
__device__ __forceinline__ float to_acc(__hip_bfloat16 x) { return __bfloat162float(x); }
__device__ __forceinline__ float to_acc(__half x) { return __half2float(x); }

template <class TA> struct AccOf { using type = float; };
template <> struct AccOf<double> { using type = double; };

// Load VEC consecutive elements of B starting at p (16-byte aligned when
// VEC > 1) and widen them to the accumulator type.
template <class TB, int VEC, class TC>
__device__ __forceinline__ void load_b(const TB *__restrict__ p, TC (&o)[VEC]) {
    if constexpr (VEC == 1) {
        o[0] = to_acc(p[0]);
    } else if constexpr (std::is_same<TB, float>::value) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if constexpr (std::is_same<TB, double>::value) {
        double2 v = *reinterpret_cast<const double2 *>(p);
        o[0] = v.x; o[1] = v.y;
    } else if constexpr (std::is_same<TB, __hip_bfloat16>::value && VEC == 4) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        o[0] = __uint_as_float(v.x << 16);
        o[1] = __uint_as_float(v.x & 0xFFFF0000u);
        o[2] = __uint_as_float(v.y << 16);
        o[3] = __uint_as_float(v.y & 0xFFFF0000u);
    } else if constexpr (std::is_same<TB, __half>::value && VEC == 4) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        __half2 h0 = *reinterpret_cast<__half2 *>(&v.x), h1 = *reinterpret_cast<__half2 *>(&v.y);
        o[0] = __low2float(h0);
        o[1] = __high2float(h0);
        o[2] = __low2float(h1);
        o[3] = __high2float(h1);
    } else if constexpr (std::is_same<TB, __hip_bfloat16>::value) {
        uint4 v = *reinterpret_cast<const uint4 *>(p);
        unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            o[2 * i] = __uint_as_float(w[i] << 16);
            o[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
        }
    } else {  // __half
        uint4 v = *reinterpret_cast<const uint4 *>(p);
        unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            __half2 h = *reinterpret_cast<__half2 *>(&w[i]);
            o[2 * i] = __low2float(h);
            o[2 * i + 1] = __high2float(h);
        }
    }
}

// Buffer-resource loads of B rows.  A raw buffer load whose byte offset is at
// or past the resource's range returns zeros without touching memory, so an
// unused slot takes the offset OOB_OFFSET instead of a branch around its load,
// and a row offset is one 24-bit multiply-add (B rows < 2^24, B < 2^31 bytes).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
constexpr unsigned OOB_OFFSET = 0x80000000u;
constexpr int GFX9_RSRC_WORD3 = 0x00020000;  // 32-bit data format, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             GFX9_RSRC_WORD3);
}

__device__ __forceinline__ float bf16_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(unsigned w) { return __uint_as_float(w & 0xFFFF0000u); }
__device__ __forceinline__ float f16_lo(unsigned w) {
    return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu)));
}
__device__ __forceinline__ float f16_hi(unsigned w) {
    return __half2float(__ushort_as_half((unsigned short)(w >> 16)));
}

template <class TB, class TC>
__device__ __forceinline__ void widen_words(const unsigned *w, int nw, TC *o) {
    if constexpr (std::is_same<TB, float>::value) {
        for (int i = 0; i < nw; i++) o[i] = __uint_as_float(w[i]);
    } else if constexpr (std::is_same<TB, double>::value) {
        for (int i = 0; i < nw / 2; i++)
            o[i] = __hiloint2double((int)w[2 * i + 1], (int)w[2 * i]);
    } else if constexpr (std::is_same<TB, __hip_bfloat16>::value) {
        for (int i = 0; i < nw; i++) {
            o[2 * i] = bf16_lo(w[i]);
            o[2 * i + 1] = bf16_hi(w[i]);
        }
    } else {
        for (int i = 0; i < nw; i++) {
            o[2 * i] = f16_lo(w[i]);
            o[2 * i + 1] = f16_hi(w[i]);
        }
    }
}

// VEC consecutive elements of B at byte offset off of the resource.
template <class TB, int VEC, class TC>
__device__ __forceinline__ void load_b_buf(__amdgpu_buffer_rsrc_t r, unsigned off,
                                           TC (&o)[VEC]) {
    constexpr int BYTES = (int)sizeof(TB) * VEC;
    if constexpr (BYTES == 16) {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        widen_words<TB, TC>(w, 4, o);
    } else if constexpr (BYTES == 8) {
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        const unsigned w[2] = {v.x, v.y};
        widen_words<TB, TC>(w, 2, o);
    } else if constexpr (BYTES == 4) {
        const unsigned w[1] = {__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0)};
        widen_words<TB, TC>(w, 1, o);
    } else {
        static_assert(BYTES == 2, "unsupported B load width");
        const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
        if constexpr (std::is_same<TB, __hip_bfloat16>::value) o[0] = bf16_lo(h);
        else o[0] = f16_lo(h);
    }
}

// B rows staged in LDS by a workgroup (window walk): rows [w0, w0 + wlen) of
// the launch's K slice, rb bytes apart from lds on.  wlen == 0: no window.
struct Window {
    unsigned char *lds;
    int w0, wlen;
    unsigned rb;
};

// VEC elements of a staged B row at LDS byte address p, widened.
template <class TB, int VEC, class TC>
__device__ __forceinline__ void load_b_lds(const unsigned char *p, TC (&o)[VEC]) {
    constexpr int BYTES = (int)sizeof(TB) * VEC;
    static_assert(BYTES == 16 || BYTES == 8, "window walk needs 8- or 16-byte lane loads");
    if constexpr (BYTES == 16) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        widen_words<TB, TC>(w, 4, o);
    } else {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        const unsigned w[2] = {v.x, v.y};
        widen_words<TB, TC>(w, 2, o);
    }
}

// Streaming accesses.  The row view's (column, value, row end) arrays and the
// finished output rows are touched once per launch; loading and storing them
// with the non-temporal bit (nt) puts their lines first in line for eviction,
// so L2 keeps the gathered B rows, which are reused.
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if constexpr (sizeof(T) == 2) {
        const unsigned short u =
            __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
        T r;
        __builtin_memcpy(&r, &u, 2);
        return r;
    } else {
        return __builtin_nontemporal_load(p);
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void st_stream4(float *p, float x, float y, float z, float w) {
    const f32x4 v = {x, y, z, w};
    __builtin_nontemporal_store(v, reinterpret_cast<f32x4 *>(p));
}
__device__ __forceinline__ void st_stream2(double *p, double x, double y) {
    const f64x2 v = {x, y};
    __builtin_nontemporal_store(v, reinterpret_cast<f64x2 *>(p));
}

__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------- row-major kernels --
//
// Lanes run along K in groups of LPE lanes (one entry stream per group, EPS =
// 64/LPE groups per wave), VEC elements per lane per chunk, Q chunks.  Two
// element layouts are used:
//   strided     k = q*LPE*VEC + v*LPE + sub   (atomics: for a fixed (q, v) the
//               LPE lanes hit LPE consecutive elements, i.e. whole lines)
//   contiguous  k = q*LPE*VEC + sub*VEC + v   (16-byte loads and stores)

struct RowArgs {
    // tile-order kernels
    int nwork;
    const int2 *work;           // (first tile, tile count) per work item
    const int *tile_colidx;
    const int *srow;            // sortrowidx
    const void *sval;           // sortval (TV)
    // reside (irregular) kernels
    int nblk;                   // balance blocks (or reside columns)
    const unsigned *colorblk;   // nullptr => unbalanced, one block per column
    const int *blkstart, *blkstop;
    const int *rptr, *rrow;
    const void *rval;
    // row-view kernel
    int m, nseg;
    const int2 *seg;
    const int *rv_ptr, *rv_col;
    const void *rv_val;
    const int *rv_row;          // output row of each row-view entry
    const int *inv_perm;        // B in sorted order: B row of column c is inv_perm[c]
    int accumulate;
    // shared
    int n;                      // columns of A
    const int *perm;            // tile-order: nullptr when B is already sorted
    const void *B;
    void *C;
    int ldb, ldc;               // row strides (elements)
    int kbase, kcount;          // K slice handled by this launch
    const float *kscale;        // F16: per-k output scale (nullptr otherwise)
    float gscale;               // F16: A-value scale
    // buffer-resource path (row-view walk): bytes of B from its K-slice start,
    // and the row stride in bytes
    unsigned bbytes, bstride;
    // row-view walk's carry slots: a segment that
    // starts inside a row stores that row's piece into carry slot
    // seg_slot[sid - 1] (ldk values per slot, the K slice at kbase).  The
    // fix-up (split_fixup) adds the nfix cut rows' slots onto C: fix[i] =
    // (row, first slot, end slot), the first nlong rows (those with more
    // than FIX_LONG slots) by a workgroup each, with atomics where rflag
    // marks rows that reside blocks of the same launch add into.
    void *carry;
    const int *seg_slot;
    const int4 *fix;
    int nfix, nlong, ldk;
    const unsigned char *rflag;
    // row-view walk: slot rows per step chosen by the autotuner (0: the rule
    // of rule_slots)
    int slots;
};

struct RowTarget {
    int vec, q, lpe;
    bool mask;
};

// Splits [0, K) into slices handled by one kernel each: the largest
// 16-byte-vector shape that fits, then masked scalar slices of <= 64.  Shared
// by the forward and backward walks.  max_cols > 0 caps the columns of a
// full-vector slice (a narrower lane group per entry, one more pass over A per
// extra slice; each pass gathers from a narrower slice of B).
inline std::vector<std::pair<int, RowTarget>> plan_row_slices(int K, int vec, bool aligned,
                                                              int max_cols = 0) {
    std::vector<std::pair<int, RowTarget>> out;
    int k = 0;
    auto fits = [&](int cols) { return max_cols <= 0 || cols <= max_cols; };
    if (aligned) {
        static const int qs[] = {4, 2, 1};
        while (K - k >= vec * 4) {
            int rem = K - k;
            bool done = false;
            for (int q : qs) {
                if (rem >= vec * 64 * q && fits(vec * 64 * q)) {
                    out.push_back({k, {vec, q, 64, false}});
                    k += vec * 64 * q;
                    done = true;
                    break;
                }
            }
            if (done) continue;
            for (int lpe = 32; lpe >= 4; lpe >>= 1) {
                if (lpe == 4 && vec != 8) break;
                if (rem >= vec * lpe && fits(vec * lpe)) {
                    out.push_back({k, {vec, 1, lpe, false}});
                    k += vec * lpe;
                    done = true;
                    break;
                }
            }
            if (!done) break;
        }
    }
    while (k < K) {
        out.push_back({k, {1, 1, 64, true}});
        k += std::min(64, K - k);
    }
    return out;
}

template <class TC>
__device__ __forceinline__ TC out_scale(const RowArgs &a, int k) {
    return a.kscale ? (TC)(a.kscale[a.kbase + k] * a.gscale) : (TC)a.gscale;
}

template <class TC, int VEC, int Q, int LPE, bool MASK>
__device__ __forceinline__ void flush_strided(const RowArgs &a, int row, int sub,
                                              TC (&acc)[Q][VEC]) {
    TC *crow = static_cast<TC *>(a.C) + (size_t)row * a.ldc + a.kbase;
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            const int k = q * LPE * VEC + v * LPE + sub;
            if (!MASK || k < a.kcount) atomic_add(crow + k, acc[q][v] * out_scale<TC>(a, k));
            acc[q][v] = (TC)0;
        }
}

template <class TB, class TC, int VEC, int Q, int LPE, bool MASK>
__device__ __forceinline__ void load_strided(const RowArgs &a, int bcol, int sub,
                                             TC (&b)[Q][VEC]) {
    const TB *brow = static_cast<const TB *>(a.B) + (size_t)bcol * a.ldb + a.kbase;
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            const int k = q * LPE * VEC + v * LPE + sub;
            b[q][v] = (!MASK || k < a.kcount) ? to_acc(brow[k]) : (TC)0;
        }
}

// -------------------------------------------------------- row view kernel --

// Output rows are written once: 16-byte non-temporal stores.
template <class TC, int VEC>
__device__ __forceinline__ void store_vec(TC *p, const TC (&x)[VEC]) {
    if constexpr (std::is_same<TC, float>::value && VEC % 4 == 0) {
#pragma unroll
        for (int i = 0; i < VEC; i += 4) st_stream4(p + i, x[i], x[i + 1], x[i + 2], x[i + 3]);
    } else if constexpr (std::is_same<TC, double>::value && VEC % 2 == 0) {
#pragma unroll
        for (int i = 0; i < VEC; i += 2) st_stream2(p + i, x[i], x[i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < VEC; i++) p[i] = x[i];
    }
}

// Output scales of a lane's elements (F16 storage: per-k scale times the
// value scale), loaded once per walk: a lane's k positions never change.
template <class TC, int VEC, int Q, int LPE, bool MASK>
__device__ __forceinline__ void lane_scales(const RowArgs &a, int sub, TC (&osc)[Q][VEC]) {
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            const int k = q * LPE * VEC + sub * VEC + v;
            osc[q][v] = (!MASK || k < a.kcount) ? out_scale<TC>(a, k) : (TC)0;
        }
}

// Writes one finished row: plain (or read-modify-write when accumulating)
// vector stores, or atomics for a row split across segments.  SCALE (F16
// storage only): the output is scaled per k, by the lane's hoisted scales osc
// (lane_scales) or, without them, read per store.
template <class TC, int VEC, int Q, int LPE, bool MASK, bool SCALE>
__device__ __forceinline__ void put_row(const RowArgs &a, int row, int sub, bool split,
                                        TC (&acc)[Q][VEC], const TC (*osc)[VEC] = nullptr) {
    TC *crow = static_cast<TC *>(a.C) + (size_t)row * a.ldc + a.kbase;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k0 = q * LPE * VEC + sub * VEC;
        TC x[VEC];
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            if constexpr (SCALE) x[v] = acc[q][v] * (osc ? osc[q][v] : out_scale<TC>(a, k0 + v));
            else x[v] = acc[q][v];
        }
        if (MASK) {
#pragma unroll
            for (int v = 0; v < VEC; v++)
                if (k0 + v < a.kcount) {
                    if (split) atomic_add(crow + k0 + v, x[v]);
                    else crow[k0 + v] = a.accumulate ? crow[k0 + v] + x[v] : x[v];
                }
        } else if (split) {
#pragma unroll
            for (int v = 0; v < VEC; v++) atomic_add(crow + k0 + v, x[v]);
        } else {
            if (a.accumulate) {
                TC old[VEC];
                load_b<TC, VEC, TC>(crow + k0, old);
#pragma unroll
                for (int v = 0; v < VEC; v++) x[v] += old[v];
            }
            store_vec<TC, VEC>(crow + k0, x);
        }
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
    }
}

// Stores the piece of a cut row that a segment starts with into its carry
// slot (scaled like put_row; no accumulation: the fix-up adds it onto C) and
// clears the accumulators, as put_row does.
template <class TC, int VEC, int Q, int LPE, bool MASK, bool SCALE>
__device__ __forceinline__ void put_carry(const RowArgs &a, int slot, int sub,
                                          TC (&acc)[Q][VEC],
                                          const TC (*osc)[VEC] = nullptr) {
    TC *crow = static_cast<TC *>(a.carry) + (size_t)slot * a.ldk + a.kbase;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k0 = q * LPE * VEC + sub * VEC;
        TC x[VEC];
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            if constexpr (SCALE) x[v] = acc[q][v] * (osc ? osc[q][v] : out_scale<TC>(a, k0 + v));
            else x[v] = acc[q][v];
        }
        if (MASK) {
#pragma unroll
            for (int v = 0; v < VEC; v++)
                if (k0 + v < a.kcount) crow[k0 + v] = x[v];
        } else {
            store_vec<TC, VEC>(crow + k0, x);
        }
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
    }
}

// Fix-up of cut row i by NT threads (one wave, or a 256-thread workgroup for
// rows with many slots): C[row] += the row's carry slots, summed in slot
// order, so the result does not depend on which segment finished first.
// Threads take VW-element vectors of the row (16 bytes when K, C and the
// slots allow, else single elements); a row of TE < NT vectors leaves P =
// NT / TE thread groups, each summing every P-th slot, and the groups' sums
// are added in group order (shuffles inside a wave, LDS across a
// workgroup).  The row's values (its first piece) are read while the slots
// load.  Rows that reside blocks of the same launch add into take atomics.
// GROUPS 256-thread groups of one workgroup may fix up GROUPS rows at once:
// group g is threads [256 g, 256 g + 256) and has its own LDS share; a group
// without a row (active false) still takes part in every barrier.
template <class TC, int NT, int GROUPS = 1>
__device__ __forceinline__ void split_fixup(const RowArgs &a, int i, bool active = true) {
    static_assert(NT == 64 || NT == 256, "one wave or 256-thread workgroup groups");
    static_assert(GROUPS == 1 || NT == 256, "several groups need 256-thread groups");
    constexpr int V16 = 16 / (int)sizeof(TC);
    constexpr int U = 8;  // slots in flight per thread
    const int t = NT == 64 ? (int)(threadIdx.x & 63) : (int)(threadIdx.x & (NT - 1));
    const int g = GROUPS == 1 ? 0 : (int)(threadIdx.x / NT);
    const int4 f = active ? a.fix[i] : make_int4(0, 0, 0, 0);
    const int row = f.x, b0 = f.y, b1 = f.z;
    TC *crow = static_cast<TC *>(a.C) + (size_t)row * a.ldc;
    const TC *carry = static_cast<const TC *>(a.carry);
    const bool atom = active && a.rflag && a.rflag[row];
    const int K = a.ldk;
    const bool vec = (K % V16) == 0 && (a.ldc % V16) == 0 && ((uintptr_t)a.C % 16) == 0 &&
                     ((uintptr_t)a.carry % 16) == 0;
    const int VW = vec ? V16 : 1;
    const int T = K / VW;                 // vectors per row
    const int TE = T < NT ? T : NT;       // threads per group
    const int P = NT / TE;                // groups
    const int p = t / TE, j0 = t % TE;
    for (int jb = 0; jb < T; jb += TE) {
        const int j = jb + j0;
        const bool mine = active && p < P && j < T;
        TC cur[V16], acc[V16];
#pragma unroll
        for (int v = 0; v < V16; v++) cur[v] = acc[v] = (TC)0;
        if (mine && p == 0 && !atom) {
            if (vec) load_b<TC, V16, TC>(crow + (size_t)j * VW, cur);
            else cur[0] = crow[j];
        }
        if (mine) {
            int b = b0 + p;
            if (vec) {
                for (; b + (U - 1) * P < b1; b += U * P) {
                    TC x[U][V16];
#pragma unroll
                    for (int u = 0; u < U; u++)
                        load_b<TC, V16, TC>(carry + (size_t)(b + u * P) * K + (size_t)j * VW, x[u]);
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int v = 0; v < V16; v++) acc[v] += x[u][v];
                }
            }
            for (; b < b1; b += P) {
                const TC *src = carry + (size_t)b * K + (size_t)j * VW;
                if (vec) {
                    TC x[V16];
                    load_b<TC, V16, TC>(src, x);
#pragma unroll
                    for (int v = 0; v < V16; v++) acc[v] += x[v];
                } else {
                    acc[0] += src[0];
                }
            }
        }
        // the groups' sums in group order
        if constexpr (NT == 64) {
            for (int q = 1; q < P; q++)  // every lane takes part in the shuffles
#pragma unroll
                for (int v = 0; v < V16; v++) {
                    const TC y = __shfl(acc[v], t + q * TE, 64);
                    if (p == 0) acc[v] += y;
                }
        } else {
            __shared__ TC part[GROUPS * NT * V16];
            TC *gp = part + g * NT * V16;
            if (P > 1) {
#pragma unroll
                for (int v = 0; v < V16; v++) gp[t * V16 + v] = acc[v];
                __syncthreads();
                if (p == 0)
                    for (int q = 1; q < P; q++)
#pragma unroll
                        for (int v = 0; v < V16; v++) acc[v] += gp[(q * TE + j0) * V16 + v];
                __syncthreads();
            }
        }
        if (mine && p == 0) {
            TC *dst = crow + (size_t)j * VW;
            for (int v = 0; v < VW; v++) {
                if (atom) atomic_add(dst + v, acc[v]);
                else dst[v] = cur[v] + acc[v];
            }
        }
    }
}

// Block order -> logical block such that each XCD (blocks are dealt to the 8
// XCDs round-robin) gets one contiguous range of logical blocks: neighbouring
// segments cover neighbouring rows, whose B rows then meet in one L2.  Any
// other dispatch order only costs locality, never correctness.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks) {
    constexpr unsigned X = 8;
    const unsigned per = nblocks / X, rem = nblocks % X, x = b % X;
    return x * per + min(x, rem) + b / X;
}

// x + x of lane (lane ^ O), every lane active: the lane exchanges of gfx950
// stay in the VALU, where a shuffle takes an LDS round trip.  O = 8: DPP
// row_ror:8 inside each 16-lane row; O = 16 / 32: v_permlane16_swap /
// v_permlane32_swap of x with itself leave x and its partner in the two
// results (which is which depends on the lane's half), and their sum is the
// pair's sum.  IEEE addition commutes, so every lane gets exactly the value
// of x + __shfl_xor(x, O).  Other O take the shuffle.
template <int O>
__device__ __forceinline__ unsigned xor_partner_u32(unsigned x) {
    static_assert(O == 8, "DPP row rotation gives the lane ^ 8 partner only");
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xf, 0xf, false);
}
template <int O>
__device__ __forceinline__ float xor_add(float x) {
    if constexpr (O == 32 || O == 16) {
        const unsigned u = __float_as_uint(x);
        const auto r = O == 32 ? __builtin_amdgcn_permlane32_swap(u, u, false, false)
                               : __builtin_amdgcn_permlane16_swap(u, u, false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    } else if constexpr (O == 8) {
        return x + __uint_as_float(xor_partner_u32<8>(__float_as_uint(x)));
    } else {
        return x + __shfl_xor(x, O, 64);
    }
}
template <int O>
__device__ __forceinline__ double xor_add(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    auto join = [](unsigned l, unsigned h) {
        return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
    };
    if constexpr (O == 32 || O == 16) {
        const auto rl = O == 32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                                : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto rh = O == 32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                                : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        return join(rl[0], rh[0]) + join(rl[1], rh[1]);
    } else if constexpr (O == 8) {
        return x + join(xor_partner_u32<8>(lo), xor_partner_u32<8>(hi));
    } else {
        return x + __shfl_xor(x, O, 64);
    }
}

// The value of x held by lane (lane ^ o), every lane active, for o a power
// of two below 64 that is known once the caller's loops unroll: o = 16 / 32
// by the lane swaps (the partner's value is the result that is not the
// lane's own), 8 by DPP row_ror:8, 2 and 1 by DPP quad permutations, 4 by a
// shuffle.
__device__ __forceinline__ unsigned lane_xor_u32(unsigned u, int o) {
    if (o == 32 || o == 16) {
        const auto r = o == 32 ? __builtin_amdgcn_permlane32_swap(u, u, false, false)
                               : __builtin_amdgcn_permlane16_swap(u, u, false, false);
        return (threadIdx.x & (unsigned)o) ? r[0] : r[1];
    }
    if (o == 8) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x128, 0xf, 0xf, false);
    if (o == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x4E, 0xf, 0xf, false);
    if (o == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0xB1, 0xf, 0xf, false);
    return (unsigned)__shfl_xor((int)u, o, 64);
}
__device__ __forceinline__ float lane_xor(float x, int o) {
    return __uint_as_float(lane_xor_u32(__float_as_uint(x), o));
}
__device__ __forceinline__ double lane_xor(double x, int o) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = lane_xor_u32((unsigned)u, o), hi = lane_xor_u32((unsigned)(u >> 32), o);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The row-end butterfly: sums the EPS = 64 / LPE sub-groups' partial sums,
// every lane ending with the total of its sub-lane.  With the lane exchanges
// above in place of shuffles: on the 1M-row matrix FP32 K=32 0.258 -> 0.239
// ms (waves had waited on the LDS in 11% of their cycles), FP64 K=32 0.349 ->
// 0.345 ms, the 512-matrix batch 5.51 -> 5.45 ms; K=128 unchanged.
template <class TC, int VEC, int Q, int LPE, int O = LPE>
__device__ __forceinline__ void subgroup_sum(TC (&acc)[Q][VEC]) {
    if constexpr (O < 64) {
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int v = 0; v < VEC; v++) acc[q][v] = xor_add<O>(acc[q][v]);
        subgroup_sum<TC, VEC, Q, LPE, 2 * O>(acc);
    }
}

// Lane permute of a float / double value (ds_bpermute moves 32 bits).
__device__ __forceinline__ float bperm_val(int addr, float x) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
}
__device__ __forceinline__ double bperm_val(int addr, double x) {
    const long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)u);
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(u >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float dot2_bf16(unsigned a, unsigned b, float c) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, a),
                                           __builtin_bit_cast(bf16x2_t, b), c, false);
}

// N slot rows of one step of the buffer-load walk, issued without branches:
// slot u of lane group h takes the staged (column, value) of lane
// base/4 + u*EPS.  Only the last slot row can reach past the row's end (the
// caller picks N = the slot rows left, at most UNROLL); its unused slots read
// zeros through OOB_OFFSET.
//
// PAIR (BF16 A values and B, FP32 accumulation, 8-byte lane loads, one chunk
// per lane) takes slot rows u and u+1 together: both entries' B values for
// one k are packed into one register (v_perm) and multiplied by both A
// values in one v_dot2_f32_bf16, so C[i][k] += a0*B0[k] + a1*B1[k] is one
// instruction where widening and two FMAs took four.  Both products belong
// to the same output element, so non-finite inputs propagate as in the plain
// sum; an odd last slot row pairs with zeros.
template <class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool PAIR, int N>
__device__ __forceinline__ void gather_body(const RowArgs &a, __amdgpu_buffer_rsrc_t rsrc,
                                            int base, int j, int nb, int h, int sub, int mc,
                                            TC mv, TC (&acc)[Q][VEC]) {
    constexpr int EPS = 64 / LPE;
    TC vv[N];
    unsigned off[N][Q];
#pragma unroll
    for (int u = 0; u < N; u++) {
        // one base address per step, the slot in the instruction offset
        const int c = __builtin_amdgcn_ds_bpermute(base + ((u * EPS) << 2), mc);
        const TC vs = bperm_val(base + ((u * EPS) << 2), mv);
        const bool ok = u < N - 1 || j + u * EPS + h < nb;
        vv[u] = ok ? vs : (TC)0;
        const unsigned roff = __umul24((unsigned)c, a.bstride);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int k0 = q * LPE * VEC + sub * VEC;
            const bool in = ok && (!MASK || k0 < a.kcount);
            off[u][q] = in ? roff + (unsigned)(k0 * (int)sizeof(TB)) : OOB_OFFSET;
        }
    }
    if constexpr (PAIR) {
        static_assert(std::is_same<TB, __hip_bfloat16>::value && std::is_same<TC, float>::value &&
                          VEC == 4 && Q == 1,
                      "paired BF16 products need 8-byte lane loads of one chunk");
        u32x2 w[N];
#pragma unroll
        for (int u = 0; u < N; u++)
            w[u] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off[u][0], 0, 0);
#pragma unroll
        for (int u = 0; u < N; u += 2) {
            // BF16 A values widened to FP32: their bits are the high halves
            const unsigned a0 = __float_as_uint(vv[u]);
            const unsigned a1 = u + 1 < N ? __float_as_uint(vv[u + 1]) : 0u;
            const unsigned ap = __builtin_amdgcn_perm(a1, a0, 0x07060302u);
            const u32x2 w1 = u + 1 < N ? w[u + 1] : u32x2{0u, 0u};
            const unsigned x0 = w[u].x, x1 = w[u].y;
            acc[0][0] = dot2_bf16(ap, __builtin_amdgcn_perm(w1.x, x0, 0x05040100u), acc[0][0]);
            acc[0][1] = dot2_bf16(ap, __builtin_amdgcn_perm(w1.x, x0, 0x07060302u), acc[0][1]);
            acc[0][2] = dot2_bf16(ap, __builtin_amdgcn_perm(w1.y, x1, 0x05040100u), acc[0][2]);
            acc[0][3] = dot2_bf16(ap, __builtin_amdgcn_perm(w1.y, x1, 0x07060302u), acc[0][3]);
        }
    } else {
        TC b[N][Q][VEC];
#pragma unroll
        for (int u = 0; u < N; u++)
#pragma unroll
            for (int q = 0; q < Q; q++) load_b_buf<TB, VEC, TC>(rsrc, off[u][q], b[u][q]);
#pragma unroll
        for (int u = 0; u < N; u++)
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * b[u][q][v];
    }
}

// One step: the branch-free body for the nu = min(UNROLL, slot rows left)
// slot rows (wave-uniform).  One body per count, instead of a guard per slot
// row, leaves the loads and products without branches between them; on the
// 1M-row matrix this took FP32 K=128 from 0.864 to 0.854 ms (banded rows
// alone 0.440 -> 0.422) and BF16 K=128 from 0.534 to 0.512 ms.
template <class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool PAIR, int N>
__device__ __forceinline__ void gather_pick(int nu, const RowArgs &a,
                                            __amdgpu_buffer_rsrc_t rsrc, int base, int j,
                                            int nb, int h, int sub, int mc, TC mv,
                                            TC (&acc)[Q][VEC]) {
    if constexpr (N == 1) {
        gather_body<TB, TC, VEC, Q, LPE, MASK, PAIR, 1>(a, rsrc, base, j, nb, h, sub, mc, mv,
                                                        acc);
    } else {
        if (nu >= N)
            gather_body<TB, TC, VEC, Q, LPE, MASK, PAIR, N>(a, rsrc, base, j, nb, h, sub, mc,
                                                            mv, acc);
        else
            gather_pick<TB, TC, VEC, Q, LPE, MASK, PAIR, N - 1>(nu, a, rsrc, base, j, nb, h,
                                                                sub, mc, mv, acc);
    }
}

template <class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool PAIR, int UNROLL>
__device__ __forceinline__ void gather_step(const RowArgs &a, __amdgpu_buffer_rsrc_t rsrc,
                                            int base, int j, int nb, int h, int sub, int mc,
                                            TC mv, TC (&acc)[Q][VEC]) {
    constexpr int EPS = 64 / LPE;
    const int nu = min(UNROLL, (nb - j + EPS - 1) / EPS);
    gather_pick<TB, TC, VEC, Q, LPE, MASK, PAIR, UNROLL>(nu, a, rsrc, base, j, nb, h, sub, mc,
                                                         mv, acc);
}

// Merge-path walk of the row view: each wave owns one segment of W items
// (rows + entries), so every wave does the same amount of work whatever the
// row lengths.  The segment's entries are contiguous in the row view, so they
// are staged 64 at a time (column, value) together with 64 row ends, and rows
// are walked inside the staged batch: the only dependent loads are the B row
// gathers.  Lanes split into EPS sub-groups of LPE lanes; sub-group h takes
// entries h, h+EPS, ... with UNROLL 16-byte loads per lane in flight; at a
// row end the sub-groups' sums are combined with a butterfly and sub-group 0
// stores the row.  Control flow is uniform across the wave.  Rows cut by a
// segment boundary are split rows, stored in pieces (carry slots below).
#ifndef SWIFT_UNROLL_NARROW
#define SWIFT_UNROLL_NARROW 4
#endif
#ifndef SWIFT_UNROLL_WIDE
#define SWIFT_UNROLL_WIDE 4
#endif
// Lane groups of 32 or 64 lanes (one or two entries per slot, at most two
// chunks per lane) take twice the slots, so a wave keeps 8-16 entries in
// flight like the narrower shapes (not in the window walk, whose 64-register
// budget would spill them).  Measured on the 1M-row matrix: FP32 K=128 0.879
// -> 0.851 ms, K=512 4.01 -> 3.45 ms; FP64 K=128 1.72 -> 1.62 ms, K=256 3.81
// -> 3.39 ms; BF16 K=128 0.516 -> 0.496 ms; K=32 (8-lane groups) was slower.
#ifndef SWIFT_UNROLL_LONG
#define SWIFT_UNROLL_LONG 8
#endif
// One-chunk 64-lane groups of 4- and 8-byte elements (FP32 K=256, FP64
// K=128) take 16 slots: on the 1M-row matrix FP32 K=256 1.550 -> 1.519 ms,
// FP64 K=128 1.531 -> 1.523 ms; BF16 K=256 was slower (0.896 -> 0.904 ms)
// and keeps 8.  16-lane groups of 2-byte elements (BF16/F16 K=64) take 8:
// BF16 K=64 0.297 -> 0.288 ms on the 1M-row matrix, 0.139 -> 0.132 ms on the
// 200k-node graph; the 8-lane groups (K=32) keep 4 (0.265 -> 0.270 ms at 8).

//
// WIN: the workgroup has staged a window of B rows in LDS (win).  A step
// whose used slots all fall inside the window reads them from LDS; any other
// step reads B from memory as without a window.
//
// Carry slots: a segment that starts inside a row stores that row's piece
// into the carry slot of the boundary it starts at, for the fix-up pass
// (split_fixup); every other row, including the first piece of a row the
// segment end cuts, is stored into C once.
//
// SLOTS: slot rows (B-row loads in flight per lane group) per step; 0 takes
// rule_slots, other values come from the autotuner.
constexpr int rule_slots(int lpe, int q, int elem, bool win = false) {
    return (lpe == 64 && q == 1 && elem >= 4 && !win)      ? 16
           : (lpe >= 32 && q <= 2 && !win)                ? SWIFT_UNROLL_LONG
           : (lpe == 16 && q == 1 && elem == 2 && !win) ? 8
           : elem == 2                                  ? SWIFT_UNROLL_NARROW
                                                        : SWIFT_UNROLL_WIDE;
}

// Slot-row counts the autotuner may choose for a lane shape: one chunk per
// lane (Q == 1) and at most one staged batch of 64 entries per step.
constexpr bool slots_ok(int lpe, int q, int slots) {
    return q == 1 && (slots == 4 || slots == 8 || slots == 16) && (64 / lpe) * slots <= 64;
}

// Elements per lane load for B elements of elem bytes (Vec16 on the host).
inline int vec_of(size_t elem) { return elem == 8 ? 2 : elem == 4 ? 4 : SWIFT_NARROW_VEC; }

// Widest K slice of one launch of the default single product (0: no cap):
// B rows of 1 KiB or more of 4- and 8-byte elements run in slices of 512
// bytes, each a pass over A with the 32-lane (FP32) or 64-lane (FP64) shape
// of one 512-byte chunk.  A narrower slice holds more B rows of a segment
// window in L2 for the same bytes.  Measured on the 1M-row matrix (ms, full
// width -> 512-byte slices): FP32 K=256 1.513 -> 1.439, K=512 3.266 ->
// 2.874 (1 KiB slices 3.024); FP64 K=128 1.518 -> 1.477, K=256 3.279 ->
// 2.951; column-major FP32 K=256 2.290 -> 2.213; the 200k-node graph FP32
// K=256 1.180 -> 1.101.  BF16 rows are not cut: 256-byte slices of BF16
// K=256 / 512 took 0.894 -> 0.904 / 1.849 -> 2.003 ms on the 1M-row matrix
// (and 0.543 -> 0.491 ms on the 200k-node graph, where the autotuner finds
// them).  The segment table follows the bytes of the first slice.
inline int default_slice_cols(int K, size_t elem) {
    return (elem >= 4 && (size_t)K * elem >= 1024) ? (int)(512 / elem) : 0;
}

// Column caps worth timing for a product of K columns (0: the default -- for
// a single product the default slicing rules, which may cut K; for a batch
// no cap): half the width of the first uncapped full-vector slice, when that
// still is a full-vector shape of at least 8 lanes, and for a single product
// the uncapped width itself.
inline std::vector<int> width_candidates(int K, size_t elem, bool aligned, bool single = true) {
    std::vector<int> out{0};
    const int vec = vec_of(elem);
    for (auto &sl : plan_row_slices(K, vec, aligned)) {
        const RowTarget &t = sl.second;
        if (t.mask) break;
        const int cols = vec * t.lpe * t.q;
        if (single) out.push_back(cols);
        if (cols / 2 >= vec * 8) out.push_back(cols / 2);
        break;
    }
    return out;
}

// Slot-row counts worth timing for a product of K columns with elem-byte B
// elements: 0 (the rule) and every other count the first full-vector slice's
// lane shape can take.
inline std::vector<int> slot_candidates(int K, size_t elem, bool aligned) {
    std::vector<int> out{0};
    for (auto &sl : plan_row_slices(K, vec_of(elem), aligned)) {
        const RowTarget &t = sl.second;
        if (t.mask) continue;
        for (int n : {4, 8, 16})
            if (slots_ok(t.lpe, t.q, n) && n != rule_slots(t.lpe, t.q, (int)elem)) out.push_back(n);
        break;
    }
    return out;
}

template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool BUF,
          bool WIN = false, int SLOTS = 0>
__device__ __forceinline__ void rowview_walk(const RowArgs &a, int sid, Window win = Window{}) {
    static_assert(!WIN || (BUF && !MASK), "the window walk is a buffer-load, unmasked walk");
    constexpr int EPS = 64 / LPE;
    // B-row loads in flight per lane group and step
    constexpr int UNROLL = SLOTS ? SLOTS : rule_slots(LPE, Q, (int)sizeof(TB), WIN);
    static_assert(EPS * UNROLL <= 64, "a step takes at most one staged batch of entries");
    constexpr int STEP = EPS * UNROLL;  // entries per iteration
    // BF16 A and B: two entries' products per v_dot2_f32_bf16 (gather_body)
    constexpr bool PAIR = std::is_same<TV, __hip_bfloat16>::value &&
                          std::is_same<TB, __hip_bfloat16>::value &&
                          std::is_same<TC, float>::value && VEC == 4 && Q == 1;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const int2 st = a.seg[sid], en = a.seg[sid + 1];
    const int r1 = en.x, e1 = en.y;
    int row = st.x, e = st.y;
    // Everything the segment's first and last row need is requested at once,
    // beside the staging below: the carry slot of the boundary before it
    // (read whether or not that boundary cuts a row) and whether the segment
    // end cuts its last row.
    const int slot_before = sid > 0 ? a.seg_slot[sid - 1] : 0;
    const bool carried_in = row < a.m && e > a.rv_ptr[row];
    const bool cut_end = r1 < a.m && e1 > a.rv_ptr[r1];
    // carry slot of the piece this segment starts with
    const int cslot = carried_in ? slot_before : 0;
    const TV *val = static_cast<const TV *>(a.rv_val);
    const TB *B = static_cast<const TB *>(a.B);
    __amdgpu_buffer_rsrc_t rsrc;
    if constexpr (BUF) rsrc = make_rsrc(B + a.kbase, a.bbytes);
    // staged row ends: lane l holds the end of row rbase + l
    int rbase = row;
    int rend_l = (rbase + lane < a.m) ? ld_stream(a.rv_ptr + rbase + lane + 1) : e1;
    // staged entries: lane l holds entry ebase + l
    int ebase = e, mc = 0;
    TC mv = (TC)0;
    if (ebase + lane < e1) {
        const int c = ld_stream(a.rv_col + ebase + lane);
        mc = a.inv_perm ? a.inv_perm[c] : c;
        mv = to_acc(ld_stream(val + ebase + lane));
    }
    // the next batch (entries ebase + 64 ...), requested one batch early so
    // that restaging waits on no load; columns are mapped through inv_perm
    // when they are staged.  On the 1M-row matrix: FP32 K=32 0.277 -> 0.272
    // ms, BF16 K=128 0.417 -> 0.406 ms, FP64 K=32 0.374 -> 0.363 ms, FP32
    // K=128 0.718 -> 0.716 ms (a segment of 64 items seldom needs a restage).
    int nc = 0;
    TC nv = (TC)0;
    if (ebase + 64 < e1 && ebase + 64 + lane < e1) {
        nc = ld_stream(a.rv_col + ebase + 64 + lane);
        nv = to_acc(ld_stream(val + ebase + 64 + lane));
    }
    TC acc[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
    // F16 storage scales its output per k: hoisted out of the row stores
    constexpr bool HOIST = std::is_same<TB, __half>::value;
    TC osc[Q][VEC];
    if constexpr (HOIST) lane_scales<TC, VEC, Q, LPE, MASK>(a, sub, osc);
    while (true) {
        if (row - rbase >= 64) {
            rbase = row;
            rend_l = (rbase + lane < a.m) ? ld_stream(a.rv_ptr + rbase + lane + 1) : e1;
        }
        // wave-uniform: kept in scalar registers so the slot tests below branch
        const int rend = __builtin_amdgcn_readlane(rend_l, row - rbase);
        const int row_end = row < r1 ? rend : e1;
        while (e < row_end) {
            if (e - ebase >= 64) {
                ebase = e;
                mc = (a.inv_perm && ebase + lane < e1) ? a.inv_perm[nc] : nc;
                mv = nv;
                nc = 0;
                nv = (TC)0;
                if (ebase + 64 < e1 && ebase + 64 + lane < e1) {
                    nc = ld_stream(a.rv_col + ebase + 64 + lane);
                    nv = to_acc(ld_stream(val + ebase + 64 + lane));
                }
            }
            const int off = e - ebase;
            const int nb = min(64 - off, row_end - e);  // entries of this row in the batch
            for (int j = 0; j < nb; j += STEP) {
                TC b[UNROLL][Q][VEC];
                TC vv[UNROLL];
                if constexpr (WIN) {
                    const int base = (off + j + h) << 2;
                    int cu[UNROLL];
                    bool oku[UNROLL];
                    bool far = false;
#pragma unroll
                    for (int u = 0; u < UNROLL; u++) {
                        cu[u] = 0;
                        oku[u] = false;
                        vv[u] = (TC)0;
                        if (u > 0 && j + u * EPS >= nb) break;
                        const int c = __builtin_amdgcn_ds_bpermute(base + ((u * EPS) << 2), mc);
                        const TC vs = bperm_val(base + ((u * EPS) << 2), mv);
                        const bool ok = j + u * EPS + h < nb;
                        cu[u] = c;
                        oku[u] = ok;
                        vv[u] = ok ? vs : (TC)0;
                        far = far || (ok && (unsigned)(c - win.w0) >= (unsigned)win.wlen);
                    }
                    // wave-uniform: every used slot of the step is staged
                    const bool staged = win.wlen > 0 && __ballot(far) == 0;
                    if (staged) {
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) {
                            if (u > 0 && j + u * EPS >= nb) break;
                            const unsigned char *p =
                                win.lds + (unsigned)(cu[u] - win.w0) * win.rb;
#pragma unroll
                            for (int q = 0; q < Q; q++) {
                                const int k0 = q * LPE * VEC + sub * VEC;
                                if (oku[u]) {
                                    load_b_lds<TB, VEC, TC>(p + k0 * (int)sizeof(TB), b[u][q]);
                                } else {
#pragma unroll
                                    for (int v = 0; v < VEC; v++) b[u][q][v] = (TC)0;
                                }
                            }
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) {
                            if (u > 0 && j + u * EPS >= nb) break;
                            const unsigned roff = __umul24((unsigned)cu[u], a.bstride);
#pragma unroll
                            for (int q = 0; q < Q; q++) {
                                const int k0 = q * LPE * VEC + sub * VEC;
                                load_b_buf<TB, VEC, TC>(
                                    rsrc, oku[u] ? roff + (unsigned)(k0 * (int)sizeof(TB))
                                                 : OOB_OFFSET,
                                    b[u][q]);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UNROLL; u++) {
                        if (u > 0 && j + u * EPS >= nb) break;
#pragma unroll
                        for (int q = 0; q < Q; q++)
#pragma unroll
                            for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * b[u][q][v];
                    }
                    continue;
                }
                if constexpr (BUF) {
                    gather_step<TB, TC, VEC, Q, LPE, MASK, PAIR, UNROLL>(
                        a, rsrc, (off + j + h) << 2, j, nb, h, sub, mc, mv, acc);
                    continue;
                }
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    const int idx = j + u * EPS + h;
                    // every lane takes part in both shuffles (a lane reading a
                    // source lane that is masked off would get a stale value)
                    const int c = __shfl(mc, (off + idx) & 63, 64);
                    const TC vs = __shfl(mv, (off + idx) & 63, 64);
                    vv[u] = idx < nb ? vs : (TC)0;
                    const TB *brow = B + (size_t)c * a.ldb + a.kbase;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const int k0 = q * LPE * VEC + sub * VEC;
                        if (idx < nb && (!MASK || k0 < a.kcount))
                            load_b<TB, VEC, TC>(brow + k0, b[u][q]);
                        else
#pragma unroll
                            for (int v = 0; v < VEC; v++) b[u][q][v] = (TC)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < UNROLL; u++)
#pragma unroll
                    for (int q = 0; q < Q; q++)
#pragma unroll
                        for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * b[u][q][v];
            }
            e += nb;
        }
        const bool last = row >= r1;
        if (last && !cut_end) break;
        // combine the EPS sub-groups' partial sums of this row
        subgroup_sum<TC, VEC, Q, LPE>(acc);
        if (h == 0) {
            if (row == st.x && carried_in)
                put_carry<TC, VEC, Q, LPE, MASK, HOIST>(a, cslot, sub, acc,
                                                       HOIST ? osc : nullptr);
            else
                put_row<TC, VEC, Q, LPE, MASK, HOIST>(a, row, sub, false, acc,
                                                     HOIST ? osc : nullptr);
        } else
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
        if (last) break;
        row++;
    }
}

// Flat walk of the row view: every group of LPE lanes is an independent
// walker that owns one merge-path segment and takes its entries U at a time
// whatever the row boundaries, so U B-row loads per walker (EPS*U per wave)
// are in flight even where rows hold a single entry.  The segment's (column,
// value, row) triples are staged LPE at a time in the walker's lanes (the
// next batch is fetched one chunk early) and broadcast with shuffles; when
// the row changes between two entries the finished row is written by the
// walker alone (no cross-lane combine).  Only the loop over chunks is
// wave-uniform; the row changes are per walker.  The first row of a segment
// that started in an earlier segment and the last row if it continues into a
// later one are split rows (zeroed beforehand) and take atomics; rows with no
// regular entry are zeroed beforehand too.
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK, int U>
__device__ __forceinline__ void flat_walk(const RowArgs &a, int w) {
    static_assert(U <= LPE, "a chunk must fit in one staged batch");
    constexpr bool SCALE = std::is_same<TB, __half>::value;  // F16 storage scales its output
    constexpr int SPAN = (LPE / U) * U;  // staged entries consumed per batch
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const bool live = w < a.nseg;
    const int e0 = live ? a.seg[w].y : 0;
    const int n = live ? a.seg[w + 1].y - e0 : 0;
    const TV *val = static_cast<const TV *>(a.rv_val);
    const TB *B = static_cast<const TB *>(a.B);
    int mc = 0, mr = -1, nc = 0, nr = -1;
    TC mv = (TC)0, nv = (TC)0;
    auto stage = [&](int rel, int &c_, TC &v_, int &r_) {
        c_ = 0;
        v_ = (TC)0;
        r_ = -1;
        if (rel + sub < n) {
            const int e = e0 + rel + sub;
            const int c = ld_stream(a.rv_col + e);
            c_ = a.inv_perm ? a.inv_perm[c] : c;
            v_ = to_acc(ld_stream(val + e));
            r_ = ld_stream(a.rv_row + e);
        }
    };
    stage(0, mc, mv, mr);
    const int first = __shfl(mr, h * LPE, 64);  // row of the first entry (-1: none)
    bool carried = false, cont = false;
    if (n > 0) {
        carried = e0 > a.rv_ptr[first];
        const int lr = a.rv_row[e0 + n - 1];
        cont = e0 + n < a.rv_ptr[lr + 1];
    }
    TC acc[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
    int cur = first;
    for (int rel = 0; __ballot(rel < n) != 0; rel += U) {
        const int off = rel % SPAN;  // wave-uniform
        if (off == 0 && rel > 0) {
            mc = nc;
            mv = nv;
            mr = nr;
        }
        if (off + U == SPAN) stage(rel + U, nc, nv, nr);
        TC b[U][Q][VEC];
        TC vv[U];
        int rr[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // all lanes take part in the shuffles
            const int src = h * LPE + off + u;
            const int c = __shfl(mc, src, 64);
            const TC vs = __shfl(mv, src, 64);
            rr[u] = __shfl(mr, src, 64);
            const bool ok = rel + u < n;
            vv[u] = ok ? vs : (TC)0;
            const TB *brow = B + (size_t)c * a.ldb + a.kbase;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int k0 = q * LPE * VEC + sub * VEC;
                if (ok && (!MASK || k0 < a.kcount))
                    load_b<TB, VEC, TC>(brow + k0, b[u][q]);
                else
#pragma unroll
                    for (int v = 0; v < VEC; v++) b[u][q][v] = (TC)0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (rel + u < n) {
                if (rr[u] != cur) {
                    put_row<TC, VEC, Q, LPE, MASK, SCALE>(a, cur, sub, cur == first && carried,
                                                          acc);
                    cur = rr[u];
                }
#pragma unroll
                for (int q = 0; q < Q; q++)
#pragma unroll
                    for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * b[u][q][v];
            }
        }
    }
    if (n > 0)
        put_row<TC, VEC, Q, LPE, MASK, SCALE>(a, cur, sub, (cur == first && carried) || cont, acc);
}

// One balance block (or one reside column when unbalanced) per wave: the
// block's single B row is loaded once, each entry is one row flush.
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK>
__device__ __forceinline__ void reside_block(const RowArgs &a, int blk) {
    constexpr int EPS = 64 / LPE;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    int r, start, stop;
    if (a.colorblk) {
        const unsigned u = a.colorblk[blk];
        r = (int)(u & 0x7FFFFFFFu);
        const bool split = (u >> 31) != 0;
        start = split ? a.blkstart[blk] : a.rptr[r];
        stop = split ? a.blkstop[blk] : a.rptr[r + 1];
    } else {
        r = blk;
        start = a.rptr[r];
        stop = a.rptr[r + 1];
    }
    const int jj = a.n - 1 - r;
    const int bcol = a.perm ? a.perm[jj] : jj;
    TC b[Q][VEC];
    load_strided<TB, TC, VEC, Q, LPE, MASK>(a, bcol, sub, b);
    const TV *rval = static_cast<const TV *>(a.rval);
    // the block's (row, value) pairs are loaded 64 at a time, one per lane,
    // together with the B row, and broadcast with shuffles: the flushes that
    // follow wait on no load
    for (int e0 = start; e0 < stop; e0 += 64) {
        int rl = 0;
        TC vl = (TC)0;
        if (e0 + lane < stop) {
            rl = a.rrow[e0 + lane];
            vl = to_acc(rval[e0 + lane]);
        }
        const int cnt = min(64, stop - e0);
        for (int i0 = 0; i0 < cnt; i0 += EPS) {
            const int row = __shfl(rl, i0 + h, 64);
            const TC val = __shfl(vl, i0 + h, 64);
            if (i0 + h < cnt) {
                TC acc[Q][VEC];
#pragma unroll
                for (int q = 0; q < Q; q++)
#pragma unroll
                    for (int v = 0; v < VEC; v++) acc[q][v] = val * b[q][v];
                flush_strided<TC, VEC, Q, LPE, MASK>(a, row, sub, acc);
            }
        }
    }
}

}  // namespace swift
