// The reductions' tree (cft_reduce, cft_reduce_seg), a few levels a launch.
//
// The contract's tree (host/include/cft.h:1210-1283, host/src/softfloat.c:
// 880-925) splits a range so its left child is the largest power of two
// strictly inside it.  That is the tree of pairing neighbours level by
// level, an odd last node carried up unchanged: at every level the nodes
// are the aligned power-of-two ranges, and the last, shorter one is the same
// tree over its own length.  So is any grouping of aligned runs of 2^k
// consecutive nodes: a launch takes each segment's partials (its elements,
// the first time) in groups of sixteen - four groups of four, each reduced
// to a sub-tree, then the sub-trees as a group of four - writes one node a
// group, and the next launch does the same over those, until one node a
// segment remains.  The result is node for node the recursive tree's,
// whatever the segment length.
//
// The first launch also forms the leaves: the element (SUM, MAXALL), its
// rounded product with b (DOT) or itself (SUMSQ), or its magnitude (SUMABS);
// each thread issues its group's loads before using any of them.  For
// sumSquare / sumAbs it records, a segment, whether an infinity or a
// signalling NaN was seen (754-2019 9.4's override, applied by the host).
// Flags are ORed a segment.
#include "fx_arith.hpp"
#include "fx_misc.hpp"
#include "kernels.h"

using namespace fx;

namespace {

enum { LEAF_VALUE = 0, LEAF_MUL = 1, LEAF_ABS = 2, LEAF_SQ = 3 };
enum { NODE_ADD = 0, NODE_MAX = 1 };

template <int NL>
__device__ __forceinline__ void ld(const uint32_t *__restrict__ p, uint64_t i, uint32_t (&x)[NL])
{
    if constexpr (NL == 8) {
        const uint4 *q = reinterpret_cast<const uint4 *>(p + i * 8);
        const uint4 v0 = q[0], v1 = q[1];
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
        x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
    } else if constexpr (NL == 4) {
        const uint4 v0 = reinterpret_cast<const uint4 *>(p)[i];
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
    } else if constexpr (NL == 2) {
        const uint2 v0 = reinterpret_cast<const uint2 *>(p)[i];
        x[0] = v0.x; x[1] = v0.y;
    } else {
        x[0] = p[i];
    }
}

template <int NL>
__device__ __forceinline__ void st(uint32_t *__restrict__ p, uint64_t i, const uint32_t (&x)[NL])
{
    if constexpr (NL == 8) {
        uint4 *q = reinterpret_cast<uint4 *>(p + i * 8);
        q[0] = make_uint4(x[0], x[1], x[2], x[3]);
        q[1] = make_uint4(x[4], x[5], x[6], x[7]);
    } else if constexpr (NL == 4) {
        reinterpret_cast<uint4 *>(p)[i] = make_uint4(x[0], x[1], x[2], x[3]);
    } else if constexpr (NL == 2) {
        reinterpret_cast<uint2 *>(p)[i] = make_uint2(x[0], x[1]);
    } else {
        p[i] = x[0];
    }
}

// 9.4's classes of an encoding: 1 an infinity, 2 a signalling NaN.
template <int NL>
__device__ __forceinline__ uint32_t special_bits(const uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    const uint32_t top = x[NL - 1];
    if (((top >> Fm::TOPF) & Fm::EMASK) != Fm::EMASK)
        return 0;
    uint32_t frac = top & Fm::FMASK;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        frac |= x[i];
    if (!frac)
        return 1u;
    return ((top >> (Fm::TOPF - 1)) & 1u) ? 0u : 2u;
}

// Bits into a segment's word: one atomic a wavefront where its lanes share
// the segment, and only for bits the word does not hold yet.
__device__ __forceinline__ void seg_or(uint32_t *words, uint64_t s, uint32_t bits)
{
    const uint64_t s0 = __builtin_amdgcn_readfirstlane((uint32_t)s) |
                        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(s >> 32)) << 32);
    if (__ballot(s != s0) == 0) {
        uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < 5; b++)
            if (__ballot((bits >> b) & 1u))
                w |= 1u << b;
        const uint64_t live = __ballot(1);
        if (w && __lane_id() == (unsigned)__builtin_ctzll(live)) {
            const uint32_t have = __hip_atomic_load(words + s0, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            if (w & ~have)
                atomicOr(words + s0, w);
        }
    } else if (bits) {
        const uint32_t have = __hip_atomic_load(words + s, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
        if (bits & ~have)
            atomicOr(words + s, bits);
    }
}

template <int NL, int LEAF, int KIND>
// At least four waves a SIMD: the multiplying leaves (DOT, SUMSQ) need 129-131
// VGPRs on their own, three waves; held to 128 (two spilled registers, outside
// the element loop's critical path) they run 6-8% faster at 2^24 binary256.
__global__ __launch_bounds__(256, 4) void k_reduce(const uint32_t *__restrict__ src,
                                                const uint32_t *__restrict__ srcb,
                                                uint32_t *__restrict__ dst, uint64_t nseg,
                                                uint64_t m, uint64_t m2, uint64_t sq,
                                                uint64_t sr, int rnd, int per_seg,
                                                uint32_t *__restrict__ segflags,
                                                uint32_t *__restrict__ segspec)
{
    const uint64_t total = nseg * m2;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // (segment, group) of o, then walked by the stride's own (sq, sr): no
    // division in the loop.  o < 2^32, and m2 <= o below, so 32 bits divide.
    uint64_t s = 0, g = o;
    if (o >= m2) {
        s = (uint32_t)o / (uint32_t)m2;
        g = o - s * m2;
    }
    uint64_t base = s * m + 16 * g;
    const uint64_t dbase = sq * m + 16 * sr;
    const uint64_t wrap = m - 16 * m2;           // s + 1, g - m2 (mod 2^64)
    const bool by_seg = per_seg || segspec;
    // one word a segment (per_seg), else the workgroup's flag slot
    const uint64_t fslot = (uint64_t)(blockIdx.x & (CFTK_FLAG_SLOTS - 1)) * CFTK_SLOT_WORDS;
    uint64_t cur = s;
    uint32_t fl = 0, spec = 0;
    for (; o < total; o += stride) {
        if (by_seg && s != cur) {                // a new segment: the last one's bits out
            seg_or(segflags, per_seg ? cur : fslot, fl);
            if (segspec)
                seg_or(segspec, cur, spec);
            fl = spec = 0;
            cur = s;
        }
        const uint64_t left = m - 16 * g;
        const uint32_t cnt16 = left < 16 ? (uint32_t)left : 16u;
        // turns 0..3 reduce sub-group t (partials 4t .. 4t+3 of the group)
        // to one sub-tree, rotated into u3; turn 4 reduces the sub-trees
        uint32_t u0[NL] = {}, u1[NL] = {}, u2[NL] = {}, u3[NL] = {};
        uint32_t t0[NL], t1[NL];
#pragma unroll 1
        for (uint32_t turn = 0; turn < 5; turn++) {
            uint32_t cnt;
            uint32_t v0[NL] = {}, v1[NL] = {}, v2[NL] = {}, v3[NL] = {};
            if (turn < 4) {
                // the sub-group's partials, every load issued before the
                // first is used; absent ones stay +0 and are never combined
                cnt = cnt16 > 4 * turn ? min(cnt16 - 4 * turn, 4u) : 0u;
                const uint64_t b4 = base + 4 * turn;
                if (cnt > 0)
                    ld<NL>(src, b4, v0);
                if (cnt > 1)
                    ld<NL>(src, b4 + 1, v1);
                if (cnt > 2)
                    ld<NL>(src, b4 + 2, v2);
                if (cnt > 3)
                    ld<NL>(src, b4 + 3, v3);
                if (segspec)
                    spec |= special_bits<NL>(v0) | special_bits<NL>(v1) |
                            special_bits<NL>(v2) | special_bits<NL>(v3);
                if constexpr (LEAF == LEAF_ABS) {
                    v0[NL - 1] &= 0x7FFFFFFFu;                   // 5.5.1: no flags
                    v1[NL - 1] &= 0x7FFFFFFFu;
                    v2[NL - 1] &= 0x7FFFFFFFu;
                    v3[NL - 1] &= 0x7FFFFFFFu;
                }
                if constexpr (LEAF == LEAF_MUL || LEAF == LEAF_SQ) {
                    // the products in place: each step multiplies v0 and
                    // rotates it to the back (unrolled, the rotation is only
                    // register naming)
#pragma unroll
                    for (uint32_t k = 0; k < 4; k++) {
                        uint32_t r[NL];
                        if (k < cnt) {
                            uint32_t y[NL];
                            if constexpr (LEAF == LEAF_MUL) {
                                ld<NL>(srcb, b4 + k, y);
                            } else {
#pragma unroll
                                for (int i = 0; i < NL; i++)
                                    y[i] = v0[i];
                            }
                            fma_core<NL>(OP_MUL, v0, y, v0, rnd, r, fl);    // c unread
                        } else {
#pragma unroll
                            for (int i = 0; i < NL; i++)
                                r[i] = v0[i];
                        }
#pragma unroll
                        for (int i = 0; i < NL; i++) {
                            v0[i] = v1[i];
                            v1[i] = v2[i];
                            v2[i] = v3[i];
                            v3[i] = r[i];
                        }
                    }
                }
            } else {
                cnt = (cnt16 + 3) / 4;
#pragma unroll
                for (int i = 0; i < NL; i++) {
                    v0[i] = u0[i];
                    v1[i] = u1[i];
                    v2[i] = u2[i];
                    v3[i] = u3[i];
                }
            }
            // three nodes at most: (v0, v1), (v2, v3), then the two of
            // those; an odd last one goes up as it is
#pragma unroll
            for (int i = 0; i < NL; i++) {
                t0[i] = v0[i];
                t1[i] = v2[i];
            }
#pragma unroll 1
            for (int step = 0; step < 3; step++) {
                const bool on = step == 0 ? cnt >= 2 : step == 1 ? cnt == 4 : cnt >= 3;
                if (!on)
                    continue;
                uint32_t x[NL], y[NL], r[NL];
#pragma unroll
                for (int i = 0; i < NL; i++) {
                    x[i] = step == 0 ? v0[i] : step == 1 ? v2[i] : t0[i];
                    y[i] = step == 0 ? v1[i] : step == 1 ? v3[i] : t1[i];
                }
                if constexpr (KIND == NODE_ADD)
                    add_core<NL>(x, y, false, rnd, r, fl);        // x + y
                else
                    misc_op<NL>(8, x, y, y, r, fl);                // 754-2019 maximum
#pragma unroll
                for (int i = 0; i < NL; i++) {
                    t0[i] = step != 1 ? r[i] : t0[i];
                    t1[i] = step == 1 ? r[i] : t1[i];
                }
            }
#pragma unroll
            for (int i = 0; i < NL; i++) {
                u0[i] = u1[i];
                u1[i] = u2[i];
                u2[i] = u3[i];
                u3[i] = t0[i];
            }
        }
        st<NL>(dst, o, t0);
        s += sq;
        g += sr;
        base += dbase;
        if (g >= m2) {
            g -= m2;
            s++;
            base += wrap;
        }
    }
    seg_or(segflags, per_seg ? cur : fslot, fl);
    if (segspec)
        seg_or(segspec, cur, spec);
}

template <int NL, int LEAF, int KIND>
hipError_t launch_one(const void *src, const void *srcb, void *dst, uint64_t nseg, uint64_t m,
                      uint64_t m2, int rnd, int per_seg, uint32_t *segflags, uint32_t *segspec,
                      int grid, hipStream_t st)
{
    const uint64_t stride = (uint64_t)grid * 256;
    hipLaunchKernelGGL((k_reduce<NL, LEAF, KIND>), dim3(grid), dim3(256), 0, st,
                       (const uint32_t *)src, (const uint32_t *)srcb, (uint32_t *)dst, nseg, m,
                       m2, stride / m2, stride % m2, rnd, per_seg, segflags, segspec);
    return hipGetLastError();
}

template <int NL>
hipError_t launch(const void *src, const void *srcb, void *dst, uint64_t nseg, uint64_t m,
                  uint64_t m2, int leaf, int kind, int rnd, int per_seg, uint32_t *segflags,
                  uint32_t *segspec, int grid, hipStream_t st)
{
#define CFT_ONE_ARGS src, srcb, dst, nseg, m, m2, rnd, per_seg, segflags, segspec, grid, st
    if (kind == NODE_MAX)
        return launch_one<NL, LEAF_VALUE, NODE_MAX>(CFT_ONE_ARGS);
    switch (leaf) {
    case LEAF_MUL: return launch_one<NL, LEAF_MUL, NODE_ADD>(CFT_ONE_ARGS);
    case LEAF_SQ: return launch_one<NL, LEAF_SQ, NODE_ADD>(CFT_ONE_ARGS);
    case LEAF_ABS: return launch_one<NL, LEAF_ABS, NODE_ADD>(CFT_ONE_ARGS);
    default: return launch_one<NL, LEAF_VALUE, NODE_ADD>(CFT_ONE_ARGS);
    }
#undef CFT_ONE_ARGS
}

} // namespace

extern "C" hipError_t cftk_reduce(int fmt, const void *src, const void *srcb, void *dst,
                                  uint64_t nseg, uint64_t m, uint64_t m2, int leaf, int kind,
                                  int rnd, int per_seg, uint32_t *segflags, uint32_t *segspec,
                                  int grid, hipStream_t st)
{
#define CFT_REDUCE_ARGS src, srcb, dst, nseg, m, m2, leaf, kind, rnd, per_seg, segflags, segspec, \
                        grid, st
    switch (fmt) {
    case 0: return launch<1>(CFT_REDUCE_ARGS);
    case 1: return launch<2>(CFT_REDUCE_ARGS);
    case 2: return launch<4>(CFT_REDUCE_ARGS);
    default: return launch<8>(CFT_REDUCE_ARGS);
    }
#undef CFT_REDUCE_ARGS
}
