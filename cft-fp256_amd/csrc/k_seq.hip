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
#include "seq_wide.hpp"

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

// LDS layout of one workgroup: code | constants | per-wave register files |
// two words (the workgroup's flags and status).
template <int NL> struct Lds {
    using RT = typename Row<NL>::T;
    uint4 *code;
    RT *konst;
    RT *regs;          // this lane's register 0, rows kWave apart
    uint32_t *words;
};

template <int NL>
__device__ __forceinline__ Lds<NL> stage_lds(const SeqArgs &A, uint32_t nregs)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    extern __shared__ uint4 smem[];
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    Lds<NL> L;
    L.code = smem;
    const uint32_t code_rows = A.code_in_lds ? A.n_insns : 0;
    L.konst = reinterpret_cast<RT *>(smem + code_rows);
    const uint32_t konst_rows = A.n_consts * NQ;
    RT *s_regs = L.konst + konst_rows;
    const size_t wave_rows = (size_t)nregs * NQ * kWave;
    L.regs = s_regs + wave * wave_rows + lane;
    L.words = reinterpret_cast<uint32_t *>(s_regs + waves * wave_rows);
    for (uint32_t i = threadIdx.x; i < code_rows; i += blockDim.x) {
        const SeqInsn in = A.code[i];
        L.code[i] = make_uint4(in.w, in.src, in.imm, in.aux);
    }
    const RT *g = reinterpret_cast<const RT *>(A.bank);
    for (uint32_t i = threadIdx.x; i < konst_rows; i += blockDim.x)
        L.konst[i] = g[i];
    if (threadIdx.x < 2)
        L.words[threadIdx.x] = 0;
    __syncthreads();
    return L;
}

// One workgroup's flags and status into the run's words.
__device__ __forceinline__ void flush_words(uint32_t *s_words, uint32_t *g_words, uint32_t fl,
                                            uint32_t status)
{
    if (fl)
        atomicOr(&s_words[0], fl);
    if (status)
        atomicOr(&s_words[1], status);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_words[0])
            atomicOr(&g_words[0], s_words[0]);
        if (s_words[1])
            atomicOr(&g_words[1], s_words[1]);
    }
}

// A lane's registers to +0 and r0/r1/r2 from the streams (element `gi`, or
// through the index tables); then, for a lane that takes part, the scratch
// it clears and its scratch-in block (host/src/seq_exec.h:763-812).
template <int NL>
__device__ __forceinline__ void lane_in(const SeqArgs &A, typename Row<NL>::T *regs,
                                        typename Row<NL>::T *pool, size_t pstride,
                                        uint32_t nregs, const int (&phys)[3], bool valid,
                                        bool keep, uint64_t gi, const uint32_t *sin_row)
{
    constexpr int NQ = Row<NL>::NQ;
    uint32_t z[NL];
    zero<NL>(z);
    for (uint32_t r = 0; r < nregs; r++)
        row_put<NL>(regs + (size_t)r * NQ * kWave, kWave, z);
    if (valid) {
        const uint32_t *src[3] = {A.a, A.b, A.c};
        const uint32_t *idx[3] = {A.idx_a, A.idx_b, A.idx_c};
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
            row_put<NL>(regs + (size_t)phys[s] * NQ * kWave, kWave, x);
        }
    }
    if (keep && pool) {
        for (uint32_t s = 0; s < A.reach; s++)
            row_put<NL>(pool + (size_t)s * NQ * pstride, pstride, z);
        for (uint32_t s = 0; sin_row && s < A.n_in; s++) {
            uint32_t x[NL];
            if (!A.idx_si) {
                el_get<NL>(sin_row, s, x);
            } else {
                const uint32_t t = A.idx_si[gi * A.n_in + s];
                if (t == 0xFFFFFFFFu) zero<NL>(x); else el_get<NL>(A.sin, t, x);
            }
            row_put<NL>(pool + (size_t)s * NQ * pstride, pstride, x);
        }
    }
}

// One lane through the program (host/src/seq_exec.h:412-685 for a block of
// one lane, the wave sharing the program counter).  keep: the lane takes
// part; trips: its count for the REPEAT at trip_pc (trip_pc past the end
// when every REPEAT keeps the image's count).  Deposits go to element
// gi * max_dep + count.  ORs the lane's flags and status into fl / status
// and its lane byte into lfl; returns its deposit count.
template <int NL>
__device__ __forceinline__ uint32_t exec_program(const SeqArgs &A, const Lds<NL> &L,
                                                 typename Row<NL>::T *pool, size_t pstride,
                                                 bool keep, uint32_t my_trips, uint32_t trip_pc,
                                                 uint64_t gi, uint32_t &fl, uint32_t &status,
                                                 uint32_t &lfl)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    RT *const my_regs = L.regs;
    const RT *const s_konst = L.konst;
    bool active = keep;
    int park = 0;                 // loop level this lane has left early, 0 for none
    uint32_t left0 = 0, left1 = 0, left2 = 0, left3 = 0;
    uint32_t body0 = 0, body1 = 0, body2 = 0, body3 = 0;
    uint32_t count = 0;
    int level = 0, qdepth = 0;

    uint32_t pc = 0;
    while (pc < A.n_insns) {
        pc = uni(pc);
        uint4 in;
        if (A.code_in_lds) {
            in = L.code[pc];
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
            const uint32_t trips = pc == trip_pc ? my_trips : imm;
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
                RT *cell = pool + (size_t)(imm & 0x00FFFFFFu) * NQ * pstride;
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
                    RT *cell = pool + (size_t)(ix[0] & (A.depth - 1u)) * NQ * pstride;
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

    return count;
}

template <int NL>
__global__ __launch_bounds__(256) void k_seq(SeqArgs A, uint32_t nregs, int phys_a, int phys_b,
                                             int phys_c, uint32_t trip_pc)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    const Lds<NL> L = stage_lds<NL>(A, nregs);
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;

    const uint64_t li = A.lane0 + (uint64_t)(blockIdx.x * waves + wave) * kWave + lane;
    const uint64_t gi = li;                        // index into the caller's arrays
    const bool valid = li < A.n;
    const bool keep = valid && (!A.mask || ((A.mask[li >> 3] >> (li & 7)) & 1u));
    RT *my_pool = A.pool ? reinterpret_cast<RT *>(A.pool) + (li - A.lane0) : nullptr;
    const size_t pstride = A.npool;
    const int phys[3] = {phys_a, phys_b, phys_c};

    lane_in<NL>(A, L.regs, my_pool, pstride, nregs, phys, valid, keep, gi,
                A.sin && !A.idx_si ? A.sin + gi * A.n_in * NL : A.sin);
    if (keep && A.dep) {                           // deposits start at +0 (seq_exec.h:744-760)
        uint32_t z[NL];
        zero<NL>(z);
        for (uint32_t k = 0; k < A.max_dep; k++)
            el_put<NL>(A.dep, gi * A.max_dep + k, z);
    }
    const uint32_t my_trips = (A.trips && valid) ? A.trips[gi] : 0u;
    uint32_t fl = 0, status = 0, lfl = 0;
    const uint32_t count = exec_program<NL>(A, L, my_pool, pstride, keep, my_trips,
                                            A.trips ? trip_pc : 0xFFFFFFFFu, gi, fl, status,
                                            lfl);

    // a kept lane's outputs (seq_exec.h:850-920)
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
    flush_words(L.words, A.words, fl, status);
}

// The batched runner: persistent waves take tiles of lanes from a counter;
// every lane runs its steps as a chain of segments, each a full run of the
// program (fresh registers and scratch, its state as the scratch-in block,
// the scratch-out block back into its state) with the segment REPEAT at
// min(steps left, seg) trips.
// A persistent runner takes its tiles longest first, and the batch's time is
// its longest lanes' time: a wave whose tile holds a lane of at least
// hi_steps trips issues ahead of the other waves on its SIMD (they fill the
// cycles it leaves), so the longest lanes run near their lone speed while
// the rest of the batch shares the machine.  hi_steps 0 leaves every wave at
// the default priority.
__device__ __forceinline__ void tile_priority(uint32_t hi_steps, uint32_t left)
{
    if (hi_steps == 0)
        return;
    if (wave_any(left >= hi_steps))
        __builtin_amdgcn_s_setprio(2);
    else
        __builtin_amdgcn_s_setprio(0);
}

// A persistent wave's tiles: its first by its place in the grid, wave-major
// (wave w of every workgroup before wave w + 1 of any), the rest from the
// queue, after those.  Tiles come longest first; handed out by the queue
// alone, the first ones went to whichever waves reached it first, which
// could put several of the longest on one CU and SIMD, where they split its
// issue at the same priority.  Workgroups are spread over the CUs in order,
// so wave-major order spreads the longest tiles too.
__device__ __forceinline__ uint32_t first_tile(uint32_t wave)
{
    return uni(wave * gridDim.x + blockIdx.x);
}

__device__ __forceinline__ uint32_t later_tile(uint32_t *queue, uint32_t waves)
{
    uint32_t t = 0;
    if (__lane_id() == 0)
        t = atomicAdd(queue, 1u);
    return gridDim.x * waves + uni(__shfl(t, 0));
}

template <int NL>
__global__ __launch_bounds__(256) void k_seq_batch(SeqArgs A, SeqBatch B, uint32_t nregs,
                                                   int phys_a, int phys_b, int phys_c,
                                                   uint32_t trip_pc)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    const Lds<NL> L = stage_lds<NL>(A, nregs);
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t slot = (uint64_t)(blockIdx.x * waves + wave) * kWave + lane;
    RT *my_pool = A.pool ? reinterpret_cast<RT *>(A.pool) + slot : nullptr;
    const size_t pstride = A.npool;
    const int phys[3] = {phys_a, phys_b, phys_c};
    uint32_t fl = 0, status = 0;

    // tiles from the queue alone: the first-tile order (first_tile) takes
    // this kernel over 128 VGPRs (4 -> 3 waves a SIMD)
    for (;;) {
        uint32_t tile = 0;
        if (lane == 0)
            tile = atomicAdd(B.queue, 1u);
        tile = uni(__shfl(tile, 0));
        if (tile >= B.ntiles)
            break;
        const uint64_t idx = (uint64_t)tile * B.lpw + lane;
        const bool valid = lane < B.lpw && idx < B.n;
        const uint32_t id = valid ? B.order[idx] : 0u;
        uint32_t left = valid ? B.steps[id] : 0u;
        tile_priority(B.hi_steps, left);
        uint32_t *st = B.state + (size_t)id * A.n_in * NL;
        uint32_t lfl = 0;
        while (wave_any(left > 0)) {
            const bool keep = left > 0;
            const uint32_t trips = left < B.seg ? left : B.seg;
            lane_in<NL>(A, L.regs, my_pool, pstride, nregs, phys, valid, keep, id, st);
            uint32_t seg_lfl = 0;
            exec_program<NL>(A, L, my_pool, pstride, keep, trips, trip_pc, id, fl, status,
                             seg_lfl);
            if (keep) {
                for (uint32_t s = 0; s < A.n_out; s++) {
                    uint32_t x[NL];
                    row_get<NL>(my_pool + (size_t)s * NQ * pstride, pstride, x);
                    el_put<NL>(st, s, x);
                }
                lfl |= seg_lfl;
                left -= trips;
            }
        }
        if (valid && B.lflags)
            B.lflags[id] = (uint8_t)lfl;
    }
    flush_words(L.words, A.words, fl, status);
}

// ---- wide execution: an instance over W lanes --------------------------------
//
// A wave holds 64 / W instances; lane j of an instance executes record j of
// each step.  Instance registers live in LDS, [register][instance][limb
// group]; the scratch pool in HBM, [slot][limb group][instance slot], and
// slot s of an instance is only ever touched by its lane s mod W (the
// scheduler places loads and stores so), which keeps every scratch access in
// one thread's program order.

#define WIDE_FENCE() __builtin_amdgcn_wave_barrier()

// A step record, read through the address space it lives in (the launch
// says which, uniformly): ds_read from LDS, global_load from HBM.  A generic
// (flat) load would do for both, but the compiler can then only wait for it
// with every counter drained, which defeats fetching the next step's records
// while the current step computes.
typedef unsigned int RecV __attribute__((ext_vector_type(4)));
typedef const RecV __attribute__((address_space(3))) *LdsRec;
typedef const RecV __attribute__((address_space(1))) *HbmRec;
// (The typed base pointers are set once from the generic ones and indexed
// as typed; a cast at each load would be folded back to a flat load.)

// One instance of a wide schedule: its records, its register file in LDS and
// its scratch slots in HBM.
template <int NL>
struct WideInst {
    using RT = typename Row<NL>::T;
    LdsRec recs_lds;          // the records in LDS (wide_run<NL, true>)
    HbmRec recs_hbm;          // or in HBM (wide_run<NL, false>)
    uint32_t nsteps, W, j, top_step, nvregs, seg, reach, n_in, n_out;
    const RT *konst;          // constant k at konst + k * NQ
    RT *regs;                 // register r at regs + r * rstride
    size_t rstride;
    RT *pool;                 // scratch slot s at pool + s * NQ * pstride
    size_t pstride;
};

template <int NL, bool RECS_LDS>
__device__ __forceinline__ uint4 wide_rec(const WideInst<NL> &I, uint32_t k)
{
    RecV v;
    if constexpr (RECS_LDS)
        v = I.recs_lds[k];
    else
        v = I.recs_hbm[k];
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Every segment of one instance, its W lanes together: `left` trips of the
// segment loop as segments of `seg` trips and a last one of the remainder,
// each a fresh run (registers +0, r0..r2 from the streams, scratch cleared
// to reach, the state as the scratch-in block, the scratch-out block back
// into the state).  Returns this lane's flags.
template <int NL, bool RECS_LDS>
__device__ __forceinline__ uint32_t wide_run(const WideInst<NL> &I, const bool valid,
                                             uint32_t left, uint32_t *st,
                                             const uint32_t *const src[3], const int phys[3],
                                             const uint32_t id)
{
    constexpr int NQ = Row<NL>::NQ;
    const uint32_t W = I.W, j = I.j;
    const size_t rstride = I.rstride, pstride = I.pstride;
    uint32_t lfl = 0;
    while (wave_any(left > 0)) {
        const bool keep = left > 0;
        const uint32_t trips = left < I.seg ? left : I.seg;
        {
            uint32_t z[NL];
            zero<NL>(z);
            for (uint32_t r = j; r < I.nvregs; r += W)
                row_put<NL>(I.regs + r * rstride, 1, z);
            WIDE_FENCE();
            if (j == 0 && valid) {
                for (int k = 0; k < 3; k++)
                    if (src[k] && phys[k] >= 0) {
                        uint32_t x[NL];
                        el_get<NL>(src[k], id, x);
                        row_put<NL>(I.regs + (uint32_t)phys[k] * rstride, 1, x);
                    }
            }
            if (keep && I.pool) {
                for (uint32_t s = j; s < I.reach; s += W)
                    row_put<NL>(I.pool + (size_t)s * NQ * pstride, pstride, z);
                for (uint32_t s = j; s < I.n_in; s += W) {
                    uint32_t x[NL];
                    el_get<NL>(st, s, x);
                    row_put<NL>(I.pool + (size_t)s * NQ * pstride, pstride, x);
                }
            }
            WIDE_FENCE();
        }
        // the steps
        bool active = keep;
        int park = 0;
        uint32_t left0 = 0, left1 = 0, left2 = 0, left3 = 0;
        uint32_t body0 = 0, body1 = 0, body2 = 0, body3 = 0;
        int level = 0;
        uint32_t pc = 0;
        // step nx's records are fetched while step pc computes; a jump
        // (REPEAT skipped, ENDREP taken) refetches
        uint32_t nx = 0xFFFFFFFFu;
        uint4 n0 = make_uint4(0, 0, 0, 0), nr = n0;
        while (pc < I.nsteps) {
            pc = uni(pc);
            uint4 r0, rec;
            if (pc == nx) {
                r0 = n0;
                rec = nr;
            } else {
                r0 = wide_rec<NL, RECS_LDS>(I, pc * W);
                rec = wide_rec<NL, RECS_LDS>(I, pc * W + j);
            }
            nx = pc + 1 < I.nsteps ? pc + 1 : pc;
            n0 = wide_rec<NL, RECS_LDS>(I, nx * W);
            nr = wide_rec<NL, RECS_LDS>(I, nx * W + j);
            const uint32_t x0 = uni(r0.x);
            const bool en = active && park == 0;
            if (WR_KIND(x0) == WK_CTRL) {
                const uint32_t op = WR_OP(x0), aux = uni(r0.y), imm = uni(r0.w);
                if (op == SEQ_HALT) {
                    pc = I.nsteps;
                } else if (op == SEQ_REPEAT) {
                    const uint32_t t = pc == I.top_step ? trips : imm;
                    const bool enter = en && t != 0;
                    if (!wave_any(enter)) {
                        pc = aux + 1;
                        continue;
                    }
                    if (en && !enter)
                        park = level + 1;
                    switch (level) {
                    case 0: left0 = t; body0 = pc + 1; break;
                    case 1: left1 = t; body1 = pc + 1; break;
                    case 2: left2 = t; body2 = pc + 1; break;
                    default: left3 = t; body3 = pc + 1; break;
                    }
                    level++;
                    pc++;
                } else {                                  // ENDREP
                    bool again = false;
                    if (en) {
                        uint32_t l;
                        switch (level) {
                        case 1: l = --left0; break;
                        case 2: l = --left1; break;
                        case 3: l = --left2; break;
                        default: l = --left3; break;
                        }
                        again = l > 0;
                        if (!again)
                            park = level;
                    }
                    if (wave_any(again)) {
                        pc = level == 1 ? body0 : level == 2 ? body1 : level == 3 ? body2
                                                                               : body3;
                    } else {
                        if (park == level)
                            park = 0;
                        level--;
                        pc++;
                    }
                }
                continue;
            }
            const uint32_t kind = WR_KIND(rec.x);
            if (en && kind != WK_NOP) {
                if (kind == WK_ALU) {
                    uint32_t xa[NL], xb[NL], xc[NL], o[NL], f = 0;
                    const uint32_t ia = WR_A(rec.y), ib = WR_B(rec.y), ic = WR_C(rec.z);
                    // constants and registers are rows of one LDS array: an
                    // operand's row is an offset from the constants, chosen by
                    // a mask (no branch), and one load
                    const uint32_t rbase = (uint32_t)(I.regs - I.konst);
                    auto row = [&](uint32_t idx, uint32_t k) -> uint32_t {
                        const uint32_t ko = idx * (uint32_t)NQ;
                        const uint32_t ro = rbase + idx * (uint32_t)rstride;
                        return ro ^ ((ro ^ ko) & (0u - k));
                    };
                    row_get<NL>(I.konst + row(ia, WR_KA(rec.x)), 1, xa);
                    row_get<NL>(I.konst + row(ib, WR_KB(rec.x)), 1, xb);
                    row_get<NL>(I.konst + row(ic, WR_KC(rec.x)), 1, xc);
                    {
                        // fused ABS / NEG of an operand: clear, then flip, its sign
                        const uint32_t m = WR_MODS(rec.x);
                        xa[NL - 1] = (xa[NL - 1] & ~(m << 31)) ^ ((m >> 1) << 31);
                        xb[NL - 1] = (xb[NL - 1] & ~((m >> 2) << 31)) ^ ((m >> 3) << 31);
                        xc[NL - 1] = (xc[NL - 1] & ~((m >> 4) << 31)) ^ ((m >> 5) << 31);
                    }
                    const uint32_t op = WR_OP(rec.x);
                    const bool addsub = op == OP_ADD || op == OP_SUB;
                    if (__builtin_amdgcn_ballot_w64(!addsub) == 0) {
                        // every ALU lane of the bundle adds or subtracts (a ring's
                        // components side by side): the two-term path, same bits
                        add_core<NL>(xa, xc, op == OP_SUB, (int)WR_RND(rec.x), o, f);
                    } else if (op <= 3) {
                        // the lanes of a bundle mix FMA, ADD, SUB and MUL: steer each
                        // to the FMA it is defined as (host/src/softfloat.c sf_steer:
                        // ADD a*1+c, SUB a*1+(-c), MUL a*b+(+-0, the product's sign))
                        // so all of them run one path.  By selects, not branches: the
                        // lanes of a bundle take different ones, and the branches
                        // cost register moves under every exec mask
                        const bool mul = op == OP_MUL;
                        const uint32_t sg = (xa[NL - 1] ^ xb[NL - 1]) & 0x80000000u;
                        uint32_t one[NL];
                        enc_one<NL>(one);
#pragma unroll
                        for (int k = 0; k < NL; k++)
                            xb[k] = addsub ? one[k] : xb[k];
#pragma unroll
                        for (int k = 0; k < NL - 1; k++)
                            xc[k] = mul ? 0u : xc[k];
                        xc[NL - 1] = mul ? sg
                                         : xc[NL - 1] ^ (op == OP_SUB ? 0x80000000u : 0u);
                        fma_core<NL>(OP_FMA, xa, xb, xc, (int)WR_RND(rec.x), o, f);
                    } else {
                        misc_op<NL>((int)op, xa, xb, xc, o, f);
                    }
                    row_put<NL>(I.regs + WR_DST(rec.z) * rstride, 1, o);
                    lfl |= f;
                } else if (kind == WK_LOAD) {
                    uint32_t x[NL];
                    row_get<NL>(I.pool + (size_t)WR_A(rec.y) * NQ * pstride, pstride, x);
                    row_put<NL>(I.regs + WR_DST(rec.z) * rstride, 1, x);
                } else if (kind == WK_STORE) {
                    uint32_t x[NL];
                    row_get<NL>(I.regs + WR_A(rec.y) * rstride, 1, x);
                    row_put<NL>(I.pool + (size_t)WR_DST(rec.z) * NQ * pstride, pstride, x);
                } else {                                  // MOVE
                    uint32_t x[NL];
                    row_get<NL>(I.regs + WR_A(rec.y) * rstride, 1, x);
                    row_put<NL>(I.regs + WR_DST(rec.z) * rstride, 1, x);
                }
            }
            WIDE_FENCE();
            pc++;
        }
        // the scratch-out block back into the state (each slot by its owner lane)
        if (keep) {
            for (uint32_t s = j; s < I.n_out; s += W) {
                uint32_t x[NL];
                row_get<NL>(I.pool + (size_t)s * NQ * pstride, pstride, x);
                el_put<NL>(st, s, x);
            }
            left -= trips;
        }
        WIDE_FENCE();
    }
    return lfl;
}

// An instance's flag byte: the OR of its lanes' flags, through its word of
// the wave's LDS flag words.
__device__ __forceinline__ uint32_t instance_flags(uint32_t *s_lflags, uint32_t g, uint32_t j,
                                                   uint32_t lfl)
{
    if (j == 0)
        s_lflags[g] = 0;
    WIDE_FENCE();
    if (lfl)
        atomicOr(&s_lflags[g], lfl);
    WIDE_FENCE();
    const uint32_t v = s_lflags[g];
    WIDE_FENCE();
    return v;
}

template <int NL, bool RECS_LDS>
__global__ __launch_bounds__(256) void k_seq_wide(SeqArgs A, SeqBatch B, SeqWide Wd,
                                                  int phys_a, int phys_b, int phys_c)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    extern __shared__ uint4 smem[];
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t W = Wd.width, ipw = kWave >> Wd.log2w;
    const uint32_t g = lane >> Wd.log2w, j = lane & (W - 1);

    // LDS: records (when they fit) | constants | register files | lane-flag words
    uint4 *s_recs = smem;
    const uint32_t rec_rows = RECS_LDS ? Wd.nsteps * W : 0;
    RT *s_konst = reinterpret_cast<RT *>(s_recs + rec_rows);
    const uint32_t konst_rows = A.n_consts * NQ;
    RT *s_regs = s_konst + konst_rows;
    const size_t wave_rows = (size_t)Wd.nvregs * ipw * NQ;
    uint32_t *s_words = reinterpret_cast<uint32_t *>(s_regs + waves * wave_rows);
    uint32_t *s_lflags = s_words + 2 + wave * kWave;        // per instance

    {
        const RT *gk = reinterpret_cast<const RT *>(A.bank);
        for (uint32_t i = threadIdx.x; i < konst_rows; i += blockDim.x)
            s_konst[i] = gk[i];
        for (uint32_t i = threadIdx.x; i < rec_rows; i += blockDim.x)
            s_recs[i] = Wd.recs[i];
        if (threadIdx.x < 2)
            s_words[threadIdx.x] = 0;
        __syncthreads();
    }
    WideInst<NL> I;
    I.recs_lds = (LdsRec)s_recs;
    I.recs_hbm = (HbmRec)Wd.recs;
    I.nsteps = Wd.nsteps;
    I.W = W;
    I.j = j;
    I.top_step = Wd.top_step;
    I.nvregs = Wd.nvregs;
    I.seg = B.seg;
    I.reach = A.reach;
    I.n_in = A.n_in;
    I.n_out = A.n_out;
    I.konst = s_konst;
    I.regs = s_regs + wave * wave_rows + g * NQ;
    I.rstride = (size_t)ipw * NQ;
    const uint64_t islot = (uint64_t)(blockIdx.x * waves + wave) * ipw + g;
    I.pool = A.pool ? reinterpret_cast<RT *>(A.pool) + islot : nullptr;
    I.pstride = A.npool;
    const uint32_t *const src[3] = {A.a, A.b, A.c};
    const int phys[3] = {phys_a, phys_b, phys_c};
    uint32_t fl = 0, status = 0;

    for (uint32_t tile = first_tile(wave); tile < B.ntiles; tile = later_tile(B.queue, waves)) {
        const uint64_t idx = (uint64_t)tile * ipw + g;
        const bool valid = idx < B.n;
        const uint32_t id = valid ? B.order[idx] : 0u;
        const uint32_t left = valid ? B.steps[id] : 0u;
        tile_priority(B.hi_steps, left);
        uint32_t *st = B.state + (size_t)id * A.n_in * NL;
        const uint32_t lfl = wide_run<NL, RECS_LDS>(I, valid, left, st, src, phys, id);
        const uint32_t inst = instance_flags(s_lflags, g, j, lfl);
        if (j == 0 && valid && B.lflags)
            B.lflags[id] = (uint8_t)inst;
        fl |= lfl;
    }
    flush_words(s_words, A.words, fl, status);
}

// Instances of several wide programs of one format in one launch.  Each wave
// owns an LDS region of M.region_rows rows: the constants of the program of
// its current tile, then the register files of the tile's instances.  A
// wave's scratch slots are instance slots [gwave * 64, gwave * 64 + 64) of
// the pool, instance g of a tile taking slot g.
template <int NL>
__global__ __launch_bounds__(256) void k_seq_multi(SeqArgs A, SeqMulti M)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    extern __shared__ uint4 smem[];
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    RT *region = reinterpret_cast<RT *>(smem) + (size_t)wave * M.region_rows;
    uint32_t *s_words =
        reinterpret_cast<uint32_t *>(reinterpret_cast<RT *>(smem) + (size_t)waves * M.region_rows);
    uint32_t *s_lflags = s_words + 2 + wave * kWave;
    if (threadIdx.x < 2)
        s_words[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t gwave = (uint64_t)blockIdx.x * waves + wave;
    const uint32_t *const src[3] = {nullptr, nullptr, nullptr};
    const int phys[3] = {-1, -1, -1};
    uint32_t fl = 0, status = 0, staged = 0xFFFFFFFFu;

    for (uint32_t tile = first_tile(wave); tile < M.ntiles; tile = later_tile(M.queue, waves)) {
        const uint4 T = M.tiles[tile];
        const uint32_t pi = uni(T.x), first = uni(T.y), count = uni(T.z);
        const SeqMProg *P = M.progs + pi;
        const uint32_t log2w = uni(P->log2w), n_consts = uni(P->n_consts);
        const uint32_t W = 1u << log2w, ipw = kWave >> log2w;
        const uint32_t g = lane >> log2w, j = lane & (W - 1);
        if (pi != staged) {
            const RT *gk = reinterpret_cast<const RT *>(P->bank);
            for (uint32_t i = lane; i < n_consts * NQ; i += kWave)
                region[i] = gk[i];
            WIDE_FENCE();
            staged = pi;
        }
        WideInst<NL> I;
        I.recs_lds = nullptr;
        I.recs_hbm = (HbmRec)P->recs;
        I.nsteps = uni(P->nsteps);
        I.W = W;
        I.j = j;
        I.top_step = uni(P->top_step);
        I.nvregs = uni(P->nvregs);
        I.seg = uni(P->seg);
        I.reach = uni(P->reach);
        I.n_in = uni(P->n_in);
        I.n_out = I.n_in;
        I.konst = region;
        I.regs = region + (size_t)n_consts * NQ + g * NQ;
        I.rstride = (size_t)ipw * NQ;
        I.pool = A.pool ? reinterpret_cast<RT *>(A.pool) + gwave * kWave + g : nullptr;
        I.pstride = A.npool;
        const bool valid = g < count;
        const uint32_t pos = first + (valid ? g : 0u);
        uint32_t *st = valid ? reinterpret_cast<uint32_t *>(M.saddr[pos]) : nullptr;
        const uint32_t left = valid ? M.steps[pos] : 0u;
        const uint32_t lfl = wide_run<NL, false>(I, valid, left, st, src, phys, pos);
        const uint32_t inst = instance_flags(s_lflags, g, j, lfl);
        if (j == 0 && valid && M.lflags)
            M.lflags[pos] = (uint8_t)inst;
        fl |= lfl;
    }
    flush_words(s_words, A.words, fl, status);
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

template <int NL>
hipError_t launch_batch(const SeqArgs &A, const SeqBatch &B, uint32_t nregs, int pa, int pb,
                        int pc, uint32_t trip_pc, int waves, int blocks, size_t lds,
                        hipStream_t st)
{
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_batch<NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL((k_seq_batch<NL>), dim3((uint32_t)blocks), dim3(waves * kWave), lds, st,
                       A, B, nregs, pa, pb, pc, trip_pc);
    return hipGetLastError();
}

template <int NL>
int batch_occupancy(int waves, size_t lds)
{
    int nb = 0;
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_batch<NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_seq_batch<NL>, waves * kWave, lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

// ---- the plain form: one launch per instruction ---------------------------
//
// Registers and scratch live in HBM, [reg or slot][limb group][position],
// position fastest; position p holds lane order[p].  The host walks the
// instruction stream and launches k_seq_op once per instruction over the
// positions still running, which the order (longest first) makes a prefix.

template <int NL>
__global__ __launch_bounds__(256) void k_seq_op(SeqInsn in, typename Row<NL>::T *regs,
                                                typename Row<NL>::T *pool, uint64_t stride,
                                                uint64_t n, const uint32_t *bank, int strict,
                                                uint32_t depth, int log2depth,
                                                uint32_t *words)
{
    using RT = typename Row<NL>::T;
    constexpr int NQ = Row<NL>::NQ;
    __shared__ uint32_t s_words[2];
    if (threadIdx.x < 2)
        s_words[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t fl = 0, status = 0;
    if (p < n) {
        RT *my = regs + p;
        RT *cellbase = pool + p;
        const RT *konst = reinterpret_cast<const RT *>(bank);
        const uint32_t w = in.w, src = in.src;
        if (!SEQ_W_CTRL(w)) {
            uint32_t xa[NL], xb[NL], xc[NL], o[NL];
            const uint32_t ia = SEQ_SRC_A(src), ib = SEQ_SRC_B(src), ic = SEQ_SRC_C(src);
            if (SEQ_W_KA(w)) row_get<NL>(konst + ia * NQ, 1, xa);
            else row_get<NL>(my + (size_t)ia * NQ * stride, stride, xa);
            if (SEQ_W_KB(w)) row_get<NL>(konst + ib * NQ, 1, xb);
            else row_get<NL>(my + (size_t)ib * NQ * stride, stride, xb);
            if (SEQ_W_KC(w)) row_get<NL>(konst + ic * NQ, 1, xc);
            else row_get<NL>(my + (size_t)ic * NQ * stride, stride, xc);
            const uint32_t op = SEQ_W_OP(w);
            if (op <= 3)
                fma_core<NL>((int)op, xa, xb, xc, (int)SEQ_W_RND(w), o, fl);
            else
                misc_op<NL>((int)op, xa, xb, xc, o, fl);
            row_put<NL>(my + (size_t)SEQ_W_RD(w) * NQ * stride, stride, o);
        } else {
            const uint32_t op = SEQ_W_OP(w);
            uint32_t x[NL];
            if (op == SEQ_STL) {
                row_get<NL>(my + (size_t)SEQ_SRC_A(src) * NQ * stride, stride, x);
                row_put<NL>(cellbase + (size_t)(in.imm & 0x00FFFFFFu) * NQ * stride, stride, x);
            } else if (op == SEQ_LDL) {
                row_get<NL>(cellbase + (size_t)(in.imm & 0x00FFFFFFu) * NQ * stride, stride, x);
                row_put<NL>(my + (size_t)SEQ_W_RD(w) * NQ * stride, stride, x);
            } else if (op == SEQ_STX || op == SEQ_LDX) {
                const uint32_t rb = SEQ_SRC_B(src);
                RT *rbp = my + (size_t)rb * NQ * stride;
                uint32_t ix[NL];
                row_get<NL>(rbp, stride, ix);
                const bool is_st = op == SEQ_STX;
                if (strict && bitlen_of<NL>(ix) > log2depth) {
                    status |= SEQ_ST_SCRATCH_RANGE;
                    if (!is_st) {
                        zero<NL>(x);
                        row_put<NL>(my + (size_t)SEQ_W_RD(w) * NQ * stride, stride, x);
                    }
                } else {
                    RT *cell = cellbase + (size_t)(ix[0] & (depth - 1u)) * NQ * stride;
                    if (is_st) {
                        row_get<NL>(my + (size_t)SEQ_SRC_A(src) * NQ * stride, stride, x);
                        row_put<NL>(cell, stride, x);
                    } else {
                        row_get<NL>(cell, stride, x);
                        row_put<NL>(my + (size_t)SEQ_W_RD(w) * NQ * stride, stride, x);
                    }
                }
                int step = (int)(in.imm & 0xFFFu);
                if (step & 0x800)
                    step -= 0x1000;
                if (step && !(!is_st && SEQ_W_RD(w) == rb)) {
                    uint32_t cur[NL], mag[NL], zc[NL], o[NL], f = 0;
                    row_get<NL>(rbp, stride, cur);
                    zero<NL>(mag);
                    zero<NL>(zc);
                    mag[0] = (uint32_t)(step < 0 ? -step : step);
                    misc_op<NL>(step > 0 ? 19 : 20, cur, mag, zc, o, f);
                    row_put<NL>(rbp, stride, o);
                }
            }
        }
    }
    if (fl)
        atomicOr(&s_words[0], fl);
    if (status)
        atomicOr(&s_words[1], status);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_words[0]) atomicOr(&words[0], s_words[0]);
        if (s_words[1]) atomicOr(&words[1], s_words[1]);
    }
}

// A segment's start for positions [0, n): registers +0 (r0..r2 from the
// streams at lane order[p]), scratch cleared to `reach`, the lane's state
// as its scratch-in block.  store: the scratch-out block back instead.
template <int NL>
__global__ __launch_bounds__(256) void k_seq_lanes(typename Row<NL>::T *regs,
                                                   typename Row<NL>::T *pool, uint64_t stride,
                                                   uint64_t n, uint32_t nregs, uint32_t reach,
                                                   uint32_t n_io, const uint32_t *order,
                                                   uint32_t *state, const uint32_t *a,
                                                   const uint32_t *b, const uint32_t *c,
                                                   int pa, int pb, int pc, int store)
{
    constexpr int NQ = Row<NL>::NQ;
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const uint32_t id = order[p];
    uint32_t *st = state + (size_t)id * n_io * NL;
    uint32_t x[NL];
    if (store) {
        for (uint32_t s = 0; s < n_io; s++) {
            row_get<NL>(pool + p + (size_t)s * NQ * stride, stride, x);
            el_put<NL>(st, s, x);
        }
        return;
    }
    zero<NL>(x);
    for (uint32_t r = 0; r < nregs; r++)
        row_put<NL>(regs + p + (size_t)r * NQ * stride, stride, x);
    for (uint32_t s = 0; s < reach; s++)
        row_put<NL>(pool + p + (size_t)s * NQ * stride, stride, x);
    const uint32_t *srcs[3] = {a, b, c};
    const int phys[3] = {pa, pb, pc};
    for (int k = 0; k < 3; k++)
        if (srcs[k] && phys[k] >= 0) {
            el_get<NL>(srcs[k], id, x);
            row_put<NL>(regs + p + (size_t)phys[k] * NQ * stride, stride, x);
        }
    for (uint32_t s = 0; s < n_io; s++) {
        el_get<NL>(st, s, x);
        row_put<NL>(pool + p + (size_t)s * NQ * stride, stride, x);
    }
}

template <int NL>
hipError_t launch_op(const SeqInsn &in, void *regs, void *pool, uint64_t stride, uint64_t n,
                     const uint32_t *bank, int strict, uint32_t depth, int log2depth,
                     uint32_t *words, hipStream_t st)
{
    using RT = typename Row<NL>::T;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL((k_seq_op<NL>), dim3(blocks), dim3(256), 0, st, in, (RT *)regs,
                       (RT *)pool, stride, n, bank, strict, depth, log2depth, words);
    return hipGetLastError();
}

template <int NL>
hipError_t launch_lanes(void *regs, void *pool, uint64_t stride, uint64_t n, uint32_t nregs,
                        uint32_t reach, uint32_t n_io, const uint32_t *order, uint32_t *state,
                        const uint32_t *a, const uint32_t *b, const uint32_t *c, int pa, int pb,
                        int pc, int store, hipStream_t st)
{
    using RT = typename Row<NL>::T;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL((k_seq_lanes<NL>), dim3(blocks), dim3(256), 0, st, (RT *)regs,
                       (RT *)pool, stride, n, nregs, reach, n_io, order, state, a, b, c, pa, pb,
                       pc, store);
    return hipGetLastError();
}

// The step records in LDS or read from device memory: one kernel each, so
// that each holds one copy of the step loop (a kernel with both ran up to
// 1.5x slower, or not, by where the loader placed its code).
template <int NL, bool RECS_LDS>
hipError_t launch_wide_as(const SeqArgs &A, const SeqBatch &B, const SeqWide &Wd, int pa, int pb,
                          int pc, int waves, int blocks, size_t lds, hipStream_t st)
{
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_wide<NL, RECS_LDS>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL((k_seq_wide<NL, RECS_LDS>), dim3((uint32_t)blocks), dim3(waves * kWave),
                       lds, st, A, B, Wd, pa, pb, pc);
    return hipGetLastError();
}

template <int NL>
hipError_t launch_wide(const SeqArgs &A, const SeqBatch &B, const SeqWide &Wd, int pa, int pb,
                       int pc, int waves, int blocks, size_t lds, hipStream_t st)
{
    return Wd.recs_in_lds
               ? launch_wide_as<NL, true>(A, B, Wd, pa, pb, pc, waves, blocks, lds, st)
               : launch_wide_as<NL, false>(A, B, Wd, pa, pb, pc, waves, blocks, lds, st);
}

template <int NL>
hipError_t launch_multi(const SeqArgs &A, const SeqMulti &M, int waves, int blocks, size_t lds,
                        hipStream_t st)
{
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_multi<NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL((k_seq_multi<NL>), dim3((uint32_t)blocks), dim3(waves * kWave), lds, st,
                       A, M);
    return hipGetLastError();
}

template <int NL>
int multi_occupancy(int waves, size_t lds)
{
    int nb = 0;
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_multi<NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_seq_multi<NL>, waves * kWave, lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

template <int NL, bool RECS_LDS>
int wide_occupancy_as(int waves, size_t lds)
{
    int nb = 0;
    if (lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_seq_wide<NL, RECS_LDS>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_seq_wide<NL, RECS_LDS>, waves * kWave,
                                                     lds) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

template <int NL>
int wide_occupancy(int waves, size_t lds, bool recs_lds)
{
    return recs_lds ? wide_occupancy_as<NL, true>(waves, lds)
                    : wide_occupancy_as<NL, false>(waves, lds);
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

extern "C" hipError_t cftk_seq_batch(int fmt, const SeqArgs *A, const SeqBatch *B,
                                     uint32_t nregs, int phys_a, int phys_b, int phys_c,
                                     uint32_t trip_pc, int waves, int blocks, hipStream_t st)
{
    const size_t lds = cftk_seq_lds(fmt, A->code_in_lds ? A->n_insns : 0, A->n_consts, nregs,
                                    waves);
    switch (fmt) {
    case 0: return launch_batch<1>(*A, *B, nregs, phys_a, phys_b, phys_c, trip_pc, waves,
                                   blocks, lds, st);
    case 1: return launch_batch<2>(*A, *B, nregs, phys_a, phys_b, phys_c, trip_pc, waves,
                                   blocks, lds, st);
    case 2: return launch_batch<4>(*A, *B, nregs, phys_a, phys_b, phys_c, trip_pc, waves,
                                   blocks, lds, st);
    default: return launch_batch<8>(*A, *B, nregs, phys_a, phys_b, phys_c, trip_pc, waves,
                                    blocks, lds, st);
    }
}

extern "C" int cftk_seq_batch_blocks_per_cu(int fmt, uint32_t n_insns_in_lds,
                                            uint32_t n_consts, uint32_t nregs, int waves)
{
    const size_t lds = cftk_seq_lds(fmt, n_insns_in_lds, n_consts, nregs, waves);
    switch (fmt) {
    case 0: return batch_occupancy<1>(waves, lds);
    case 1: return batch_occupancy<2>(waves, lds);
    case 2: return batch_occupancy<4>(waves, lds);
    default: return batch_occupancy<8>(waves, lds);
    }
}

extern "C" hipError_t cftk_seq_op(int fmt, const SeqInsn *in, void *regs, void *pool,
                                  uint64_t stride, uint64_t n, const uint32_t *bank, int strict,
                                  uint32_t depth, int log2depth, uint32_t *words,
                                  hipStream_t st)
{
    switch (fmt) {
    case 0: return launch_op<1>(*in, regs, pool, stride, n, bank, strict, depth, log2depth,
                                words, st);
    case 1: return launch_op<2>(*in, regs, pool, stride, n, bank, strict, depth, log2depth,
                                words, st);
    case 2: return launch_op<4>(*in, regs, pool, stride, n, bank, strict, depth, log2depth,
                                words, st);
    default: return launch_op<8>(*in, regs, pool, stride, n, bank, strict, depth, log2depth,
                                 words, st);
    }
}

extern "C" hipError_t cftk_seq_lanes(int fmt, void *regs, void *pool, uint64_t stride,
                                     uint64_t n, uint32_t nregs, uint32_t reach, uint32_t n_io,
                                     const uint32_t *order, uint32_t *state, const uint32_t *a,
                                     const uint32_t *b, const uint32_t *c, int pa, int pb,
                                     int pc, int store, hipStream_t st)
{
    switch (fmt) {
    case 0: return launch_lanes<1>(regs, pool, stride, n, nregs, reach, n_io, order, state, a,
                                   b, c, pa, pb, pc, store, st);
    case 1: return launch_lanes<2>(regs, pool, stride, n, nregs, reach, n_io, order, state, a,
                                   b, c, pa, pb, pc, store, st);
    case 2: return launch_lanes<4>(regs, pool, stride, n, nregs, reach, n_io, order, state, a,
                                   b, c, pa, pb, pc, store, st);
    default: return launch_lanes<8>(regs, pool, stride, n, nregs, reach, n_io, order, state, a,
                                    b, c, pa, pb, pc, store, st);
    }
}

extern "C" size_t cftk_seq_wide_lds(int fmt, uint32_t n_consts, uint32_t nsteps_in_lds,
                                    uint32_t width, uint32_t nvregs, int waves)
{
    const size_t row = ((size_t)1 << fmt) * 4;
    const size_t ipw = kWave / width;
    return (size_t)n_consts * row + (size_t)nsteps_in_lds * width * 16 +
           (size_t)waves * nvregs * ipw * row + 8 + (size_t)waves * kWave * 4;
}

extern "C" hipError_t cftk_seq_wide(int fmt, const SeqArgs *A, const SeqBatch *B,
                                    const SeqWide *Wd, int phys_a, int phys_b, int phys_c,
                                    int waves, int blocks, hipStream_t st)
{
    const size_t lds = cftk_seq_wide_lds(fmt, A->n_consts, Wd->recs_in_lds ? Wd->nsteps : 0,
                                         Wd->width, Wd->nvregs, waves);
    switch (fmt) {
    case 0: return launch_wide<1>(*A, *B, *Wd, phys_a, phys_b, phys_c, waves, blocks, lds, st);
    case 1: return launch_wide<2>(*A, *B, *Wd, phys_a, phys_b, phys_c, waves, blocks, lds, st);
    case 2: return launch_wide<4>(*A, *B, *Wd, phys_a, phys_b, phys_c, waves, blocks, lds, st);
    default: return launch_wide<8>(*A, *B, *Wd, phys_a, phys_b, phys_c, waves, blocks, lds, st);
    }
}

extern "C" int cftk_seq_wide_blocks_per_cu(int fmt, uint32_t n_consts, uint32_t nsteps_in_lds,
                                           uint32_t width, uint32_t nvregs, int waves)
{
    const size_t lds = cftk_seq_wide_lds(fmt, n_consts, nsteps_in_lds, width, nvregs, waves);
    const bool recs = nsteps_in_lds != 0;
    switch (fmt) {
    case 0: return wide_occupancy<1>(waves, lds, recs);
    case 1: return wide_occupancy<2>(waves, lds, recs);
    case 2: return wide_occupancy<4>(waves, lds, recs);
    default: return wide_occupancy<8>(waves, lds, recs);
    }
}

extern "C" size_t cftk_seq_row_bytes(int fmt)
{
    switch (fmt) {
    case 0: return sizeof(Row<1>::T);
    case 1: return sizeof(Row<2>::T);
    case 2: return sizeof(Row<4>::T);
    default: return sizeof(Row<8>::T);
    }
}

extern "C" size_t cftk_seq_multi_lds(int fmt, uint32_t region_rows, int waves)
{
    return (size_t)waves * region_rows * cftk_seq_row_bytes(fmt) + 8 +
           (size_t)waves * kWave * 4;
}

extern "C" hipError_t cftk_seq_multi(int fmt, const SeqArgs *A, const SeqMulti *M, int waves,
                                     int blocks, hipStream_t st)
{
    const size_t lds = cftk_seq_multi_lds(fmt, M->region_rows, waves);
    switch (fmt) {
    case 0: return launch_multi<1>(*A, *M, waves, blocks, lds, st);
    case 1: return launch_multi<2>(*A, *M, waves, blocks, lds, st);
    case 2: return launch_multi<4>(*A, *M, waves, blocks, lds, st);
    default: return launch_multi<8>(*A, *M, waves, blocks, lds, st);
    }
}

extern "C" int cftk_seq_multi_blocks_per_cu(int fmt, uint32_t region_rows, int waves)
{
    const size_t lds = cftk_seq_multi_lds(fmt, region_rows, waves);
    switch (fmt) {
    case 0: return multi_occupancy<1>(waves, lds);
    case 1: return multi_occupancy<2>(waves, lds);
    case 2: return multi_occupancy<4>(waves, lds);
    default: return multi_occupancy<8>(waves, lds);
    }
}
