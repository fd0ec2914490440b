// The sequencer's interpreter: one program instance a lane, a wave of 64
// lanes walking the pre-decoded instruction stream together.
//
// Semantics follow the reference executor (host/src/seq_exec.h:412-685,
// seq_run_block, and the per-lane work around it at seq_exec.h:699-923) for
// a block of one lane per thread, which the reference's contract makes
// equivalent to any block size: the loader refuses ACTALL and HALT inside a
// loop, so a lane that leaves a loop early (its trips done, or inactive)
// stays out of it until the loop ends, and nothing a lane computes depends on
// another lane.
//
// Layout on the device:
//  - registers: LDS, per wave, [phys reg][limb group][lane] in 16-byte (or
//    8/4-byte for binary64/32) rows, so a wave's read of one operand is one
//    conflict-free ds_read per limb group;
//  - code and constant bank: staged into LDS by the workgroup at entry (the
//    code only when it fits; otherwise read through the scalar cache);
//  - scratch: an HBM pool, [slot][limb group][lane], lane fastest, so every
//    STL/LDL/STX/LDX of a wave is one coalesced access per limb group;
//  - the caller's streams, deposits and scratch blocks: dense little-endian
//    elements as the C ABI has them.
//
// Control is wave-uniform (the program counter, the loop nesting, the quiet
// depth); activity, loop trip counts, deposit counts and flags are per lane.
// A lane executes an instruction iff it is active and not parked, where a
// lane parks at loop level L when its own trip count for that loop runs out
// (or it is inactive at REPEAT/ENDREP) and unparks when the wave leaves the
// loop.
#include "fx_arith.hpp"
#include "fx_misc.hpp"
#include "kernels.h"
#include "seq_code.h"

using namespace fx;

namespace {

constexpr int kWave = 64;

template <int NL> struct Row;
template <> struct Row<1> { using T = uint32_t; static constexpr int NQ = 1; };
template <> struct Row<2> { using T = uint2; static constexpr int NQ = 1; };
template <> struct Row<4> { using T = uint4; static constexpr int NQ = 1; };
template <> struct Row<8> { using T = uint4; static constexpr int NQ = 2; };

template <int NL>
__device__ __forceinline__ void row_get(const typename Row<NL>::T *p, size_t stride,
                                        uint32_t (&x)[NL])
{
    if constexpr (NL == 1) {
        x[0] = p[0];
    } else if constexpr (NL == 2) {
        const uint2 v = p[0];
        x[0] = v.x; x[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < Row<NL>::NQ; q++) {
            const uint4 v = p[q * stride];
            x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
        }
    }
}

template <int NL>
__device__ __forceinline__ void row_put(typename Row<NL>::T *p, size_t stride,
                                        const uint32_t (&x)[NL])
{
    if constexpr (NL == 1) {
        p[0] = x[0];
    } else if constexpr (NL == 2) {
        p[0] = make_uint2(x[0], x[1]);
    } else {
#pragma unroll
        for (int q = 0; q < Row<NL>::NQ; q++)
            p[q * stride] = make_uint4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
    }
}

// A dense element of the caller's arrays (element i at i * NL limbs).
template <int NL>
__device__ __forceinline__ void el_get(const uint32_t *p, size_t i, uint32_t (&x)[NL])
{
    row_get<NL>(reinterpret_cast<const typename Row<NL>::T *>(p + i * NL), 1, x);
}

template <int NL>
__device__ __forceinline__ void el_put(uint32_t *p, size_t i, const uint32_t (&x)[NL])
{
    row_put<NL>(reinterpret_cast<typename Row<NL>::T *>(p + i * NL), 1, x);
}

template <int NL>
__device__ __forceinline__ void zero(uint32_t (&x)[NL])
{
#pragma unroll
    for (int k = 0; k < NL; k++)
        x[k] = 0;
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// bit length of an NL-limb integer, for the strict scratch range test
template <int NL>
__device__ __forceinline__ int bitlen_of(const uint32_t (&x)[NL])
{
    int b = 0;
#pragma unroll
    for (int k = 0; k < NL; k++)
        if (x[k])
            b = 32 * k + 32 - __clz(x[k]);
    return b;
}

template <int NL>
__global__ __launch_bounds__(256) void k_seq(SeqArgs A, uint32_t nregs, int phys_a, int phys_b,
                                             int phys_c, uint32_t trip_pc)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    extern __shared__ uint4 smem[];

    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;

    // ---- LDS carve-up: code | constants | per-wave register files | words
    uint4 *s_code = smem;
    const uint32_t code_rows = A.code_in_lds ? A.n_insns : 0;
    RT *s_konst = reinterpret_cast<RT *>(smem + code_rows);
    const uint32_t konst_rows = A.n_consts * NQ;
    RT *s_regs = s_konst + konst_rows;
    const size_t wave_rows = (size_t)nregs * NQ * kWave;
    RT *my_regs = s_regs + wave * wave_rows + lane;
    uint32_t *s_words = reinterpret_cast<uint32_t *>(s_regs + waves * wave_rows);

    for (uint32_t i = threadIdx.x; i < code_rows; i += blockDim.x) {
        const SeqInsn in = A.code[i];
        s_code[i] = make_uint4(in.w, in.src, in.imm, in.aux);
    }
    {
        const RT *g = reinterpret_cast<const RT *>(A.bank);
        for (uint32_t i = threadIdx.x; i < konst_rows; i += blockDim.x)
            s_konst[i] = g[i];
    }
    if (threadIdx.x < 2)
        s_words[threadIdx.x] = 0;
    __syncthreads();

    // ---- this lane
    const uint64_t li = A.lane0 + (uint64_t)(blockIdx.x * waves + wave) * kWave + lane;
    const uint64_t gi = li;                        // index into the caller's arrays
    const bool valid = li < A.n;
    const bool keep = valid && (!A.mask || ((A.mask[li >> 3] >> (li & 7)) & 1u));
    const uint64_t pl = li - A.lane0;              // index into the scratch pool
    RT *my_pool = reinterpret_cast<RT *>(A.pool) + pl;
    const size_t pstride = A.npool;

    // registers: +0, then r0/r1/r2 from the streams (seq_exec.h:794-806)
    {
        uint32_t z[NL];
        zero<NL>(z);
        for (uint32_t r = 0; r < nregs; r++)
            row_put<NL>(my_regs + (size_t)r * NQ * kWave, kWave, z);
        if (valid) {
            const uint32_t *src[3] = {A.a, A.b, A.c};
            const uint32_t *idx[3] = {A.idx_a, A.idx_b, A.idx_c};
            const int phys[3] = {phys_a, phys_b, phys_c};
#pragma unroll
            for (int s = 0; s < 3; s++) {
                if (phys[s] < 0 || !src[s])
                    continue;
                uint32_t x[NL];
                if (!idx[s]) {
                    el_get<NL>(src[s], gi, x);
                } else {
                    const uint32_t e = idx[s][gi];
                    if (e == 0xFFFFFFFFu) zero<NL>(x); else el_get<NL>(src[s], e, x);
                }
                row_put<NL>(my_regs + (size_t)phys[s] * NQ * kWave, kWave, x);
            }
        }
        // the scratch a run clears, then the scratch-in block
        if (keep && A.pool) {
            for (uint32_t s = 0; s < A.reach; s++)
                row_put<NL>(my_pool + (size_t)s * NQ * pstride, pstride, z);
            for (uint32_t s = 0; A.sin && s < A.n_in; s++) {
                uint32_t x[NL];
                const uint64_t e = gi * A.n_in + s;
                if (!A.idx_si) {
                    el_get<NL>(A.sin, e, x);
                } else {
                    const uint32_t t = A.idx_si[e];
                    if (t == 0xFFFFFFFFu) zero<NL>(x); else el_get<NL>(A.sin, t, x);
                }
                row_put<NL>(my_pool + (size_t)s * NQ * pstride, pstride, x);
            }
        }
        // deposits of every kept lane start at +0 (seq_exec.h:744-760)
        if (keep && A.dep)
            for (uint32_t k = 0; k < A.max_dep; k++)
                el_put<NL>(A.dep, gi * A.max_dep + k, z);
    }

    bool active = keep;
    int park = 0;                 // loop level this lane has left early, 0 for none
    uint32_t left0 = 0, left1 = 0, left2 = 0, left3 = 0;
    uint32_t count = 0, lfl = 0, fl = 0, status = 0;
    uint32_t body0 = 0, body1 = 0, body2 = 0, body3 = 0;
    int level = 0, qdepth = 0;

    uint32_t pc = 0;
    while (pc < A.n_insns) {
        pc = uni(pc);
        uint4 in;
        if (A.code_in_lds) {
            in = s_code[pc];
        } else {
            const SeqInsn g = A.code[pc];
            in = make_uint4(g.w, g.src, g.imm, g.aux);
        }
        const uint32_t w = uni(in.x), src = uni(in.y), imm = uni(in.z), aux = uni(in.w);
        const bool en = active && park == 0;

        if (!SEQ_W_CTRL(w)) {
            if (wave_any(en)) {
                uint32_t xa[NL], xb[NL], xc[NL], o[NL];
                const uint32_t ia = SEQ_SRC_A(src), ib = SEQ_SRC_B(src), ic = SEQ_SRC_C(src);
                if (SEQ_W_KA(w)) row_get<NL>(s_konst + ia * NQ, 1, xa);
                else row_get<NL>(my_regs + (size_t)ia * NQ * kWave, kWave, xa);
                if (SEQ_W_KB(w)) row_get<NL>(s_konst + ib * NQ, 1, xb);
                else row_get<NL>(my_regs + (size_t)ib * NQ * kWave, kWave, xb);
                if (SEQ_W_KC(w)) row_get<NL>(s_konst + ic * NQ, 1, xc);
                else row_get<NL>(my_regs + (size_t)ic * NQ * kWave, kWave, xc);
                const uint32_t op = SEQ_W_OP(w);
                uint32_t f = 0;
                if (op <= 3)
                    fma_core<NL>((int)op, xa, xb, xc, (int)SEQ_W_RND(w), o, f);
                else
                    misc_op<NL>((int)op, xa, xb, xc, o, f);
                if (en) {
                    row_put<NL>(my_regs + (size_t)SEQ_W_RD(w) * NQ * kWave, kWave, o);
                    if (!qdepth) {
                        fl |= f;
                        lfl |= f;
                    }
                }
            }
            pc++;
            continue;
        }

        switch (SEQ_W_OP(w)) {
        case SEQ_HALT:
            pc = A.n_insns;
            break;
        case SEQ_REPEAT: {
            uint32_t trips = imm;
            if (A.trips && pc == trip_pc)
                trips = valid ? A.trips[gi] : 0u;
            const bool enter = en && trips != 0;
            if (!wave_any(enter)) {
                pc = aux + 1;                 // nobody runs the body: skip the loop
                break;
            }
            if (en && !enter)
                park = level + 1;
            switch (level) {
            case 0: left0 = trips; body0 = pc + 1; break;
            case 1: left1 = trips; body1 = pc + 1; break;
            case 2: left2 = trips; body2 = pc + 1; break;
            default: left3 = trips; body3 = pc + 1; break;
            }
            level++;
            pc++;
            break;
        }
        case SEQ_ENDREP: {
            bool again = false;
            if (en) {
                uint32_t left;
                switch (level) {
                case 1: left = --left0; break;
                case 2: left = --left1; break;
                case 3: left = --left2; break;
                default: left = --left3; break;
                }
                again = left > 0;
                if (!again)
                    park = level;
            }
            if (wave_any(again)) {
                pc = level == 1 ? body0 : level == 2 ? body1 : level == 3 ? body2 : body3;
            } else {
                if (park == level)
                    park = 0;
                level--;
                pc++;
            }
            break;
        }
        case SEQ_DEPOSIT:
            if (en) {
                if (count >= A.max_dep) {
                    status |= SEQ_ST_DEPOSIT_OVERFLOW;
                    lfl |= SEQ_LF_DEPOSIT_OVERFLOW;
                } else {
                    uint32_t x[NL];
                    row_get<NL>(my_regs + (size_t)SEQ_SRC_A(src) * NQ * kWave, kWave, x);
                    el_put<NL>(A.dep, gi * A.max_dep + count, x);
                    count++;
                }
            }
            pc++;
            break;
        case SEQ_SETACT:
            if (en) {
                uint32_t x[NL];
                row_get<NL>(my_regs + (size_t)SEQ_SRC_A(src) * NQ * kWave, kWave, x);
                x[NL - 1] &= 0x7FFFFFFFu;
                bool nz = false;
#pragma unroll
                for (int k = 0; k < NL; k++)
                    nz |= x[k] != 0;
                active = nz;
            }
            pc++;
            break;
        case SEQ_ACTALL:
            active = keep;
            pc++;
            break;
        case SEQ_STL:
        case SEQ_LDL:
            if (en) {
                RT *cell = my_pool + (size_t)(imm & 0x00FFFFFFu) * NQ * pstride;
                uint32_t x[NL];
                if (SEQ_W_OP(w) == SEQ_STL) {
                    row_get<NL>(my_regs + (size_t)SEQ_SRC_A(src) * NQ * kWave, kWave, x);
                    row_put<NL>(cell, pstride, x);
                } else {
                    row_get<NL>(cell, pstride, x);
                    row_put<NL>(my_regs + (size_t)SEQ_W_RD(w) * NQ * kWave, kWave, x);
                }
            }
            pc++;
            break;
        case SEQ_STX:
        case SEQ_LDX:
            if (en) {
                const uint32_t rb = SEQ_SRC_B(src);
                RT *rbp = my_regs + (size_t)rb * NQ * kWave;
                uint32_t ix[NL], x[NL];
                row_get<NL>(rbp, kWave, ix);
                const bool is_st = SEQ_W_OP(w) == SEQ_STX;
                if (A.strict && bitlen_of<NL>(ix) > A.log2depth) {
                    status |= SEQ_ST_SCRATCH_RANGE;
                    lfl |= SEQ_LF_SCRATCH_RANGE;
                    if (!is_st) {
                        zero<NL>(x);
                        row_put<NL>(my_regs + (size_t)SEQ_W_RD(w) * NQ * kWave, kWave, x);
                    }
                } else {
                    RT *cell = my_pool + (size_t)(ix[0] & (A.depth - 1u)) * NQ * pstride;
                    if (is_st) {
                        row_get<NL>(my_regs + (size_t)SEQ_SRC_A(src) * NQ * kWave, kWave, x);
                        row_put<NL>(cell, pstride, x);
                    } else {
                        row_get<NL>(cell, pstride, x);
                        row_put<NL>(my_regs + (size_t)SEQ_W_RD(w) * NQ * kWave, kWave, x);
                    }
                }
                // the post-step (seq_exec.h:380-409): rb += step modulo 2^width,
                // except an LDX into rb itself
                int step = (int)(imm & 0xFFFu);
                if (step & 0x800)
                    step -= 0x1000;
                if (step && !(!is_st && SEQ_W_RD(w) == rb)) {
                    uint32_t cur[NL], mag[NL], zc[NL], o[NL], f = 0;
                    row_get<NL>(rbp, kWave, cur);
                    zero<NL>(mag);
                    zero<NL>(zc);
                    mag[0] = (uint32_t)(step < 0 ? -step : step);
                    misc_op<NL>(step > 0 ? 19 : 20, cur, mag, zc, o, f);
                    row_put<NL>(rbp, kWave, o);
                }
            }
            pc++;
            break;
        case SEQ_QUIET:
            qdepth++;
            pc++;
            break;
        case SEQ_ENDQUIET:
            qdepth--;
            pc++;
            break;
        case SEQ_RAISE:
            if (en) {
                uint32_t x[NL];
                row_get<NL>(my_regs + (size_t)SEQ_SRC_A(src) * NQ * kWave, kWave, x);
                const uint32_t low = x[0] & 0xFFu;
                if (!qdepth) {
                    fl |= low & 0x1Fu;
                    lfl |= low & 0x1Fu;
                }
                if (low & 0x80u) {
                    status |= SEQ_ST_MARKED;
                    lfl |= SEQ_LF_MARKED;
                }
            }
            pc++;
            break;
        default:
            // the loader refuses every other control code
            pc = A.n_insns;
            break;
        }
    }

    // ---- a kept lane's outputs (seq_exec.h:850-920)
    if (keep) {
        if (A.counts)
            A.counts[gi] = count;
        if (A.lflags)
            A.lflags[gi] = (uint8_t)lfl;
        for (uint32_t s = 0; A.sout && s < A.n_out; s++) {
            uint32_t x[NL];
            row_get<NL>(my_pool + (size_t)s * NQ * pstride, pstride, x);
            el_put<NL>(A.sout, gi * A.n_out + s, x);
        }
    }
    if (fl)
        atomicOr(&s_words[0], fl);
    if (status)
        atomicOr(&s_words[1], status);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_words[0])
            atomicOr(&A.words[0], s_words[0]);
        if (s_words[1])
            atomicOr(&A.words[1], s_words[1]);
    }
}

template <int NL>
hipError_t launch(const SeqArgs &A, uint32_t nregs, int pa, int pb, int pc, uint32_t trip_pc,
                  int waves, size_t lds, hipStream_t st)
{
    const uint64_t per_block = (uint64_t)waves * kWave;
    const uint64_t blocks = (A.n - A.lane0 + per_block - 1) / per_block;
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq<NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL((k_seq<NL>), dim3((uint32_t)blocks), dim3(waves * kWave), lds, st, A,
                       nregs, pa, pb, pc, trip_pc);
    return hipGetLastError();
}

} // namespace

extern "C" size_t cftk_seq_lds(int fmt, uint32_t n_insns_in_lds, uint32_t n_consts,
                               uint32_t nregs, int waves)
{
    const size_t nl = (size_t)1 << fmt;
    const size_t row = nl * 4;
    return (size_t)n_insns_in_lds * 16 + (size_t)n_consts * row +
           (size_t)waves * nregs * row * kWave + 16;
}

extern "C" hipError_t cftk_seq(int fmt, const SeqArgs *A, uint32_t nregs, int phys_a,
                               int phys_b, int phys_c, uint32_t trip_pc, int waves,
                               hipStream_t st)
{
    const size_t lds = cftk_seq_lds(fmt, A->code_in_lds ? A->n_insns : 0, A->n_consts, nregs,
                                    waves);
    switch (fmt) {
    case 0: return launch<1>(*A, nregs, phys_a, phys_b, phys_c, trip_pc, waves, lds, st);
    case 1: return launch<2>(*A, nregs, phys_a, phys_b, phys_c, trip_pc, waves, lds, st);
    case 2: return launch<4>(*A, nregs, phys_a, phys_b, phys_c, trip_pc, waves, lds, st);
    default: return launch<8>(*A, nregs, phys_a, phys_b, phys_c, trip_pc, waves, lds, st);
    }
}
