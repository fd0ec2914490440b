// Launchers for the HIP kernels, called by the host side of the C ABI.
// Every pointer is a device pointer; every launch goes on `st`.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "seq_code.h"

extern "C" {

// Flag slots.  The elementwise kernels OR a wavefront's flags into slot
// (workgroup index & slot_mask), the word at flags + slot * CFTK_SLOT_WORDS:
// one 128-byte line a slot, so the waves of a short launch, which all finish
// at once, spread their atomics over up to CFTK_FLAG_SLOTS lines instead of
// queueing on one address.  slot_mask 0: the single word *flags.
#define CFTK_FLAG_SLOTS 64
#define CFTK_SLOT_WORDS 32

// d[i] = op(a[i], b[i], c[i]) over n elements of format fmt (0..3).  An
// operand the opcode does not read may be NULL; bit r of smask makes operand
// r a single element.  The call's flags are ORed into the slots at flags.
hipError_t cftk_elem(int fmt, int op, int rnd, const void *a, const void *b, const void *c,
                     void *d, size_t n, uint32_t smask, uint32_t *flags, uint32_t slot_mask,
                     int grid, hipStream_t st);

// Correctly-rounded a/b (op 0) or sqrt(a) (op 1).
hipError_t cftk_divsqrt(int fmt, int which, int rnd, const void *a, const void *b, void *d,
                        size_t n, uint32_t smask, uint32_t *flags, uint32_t slot_mask, int grid,
                        hipStream_t st);

// For each of nareas slot areas (area b at slots + b * CFTK_FLAG_SLOTS *
// CFTK_SLOT_WORDS), the OR of the words of its slots 0..nslots-1 (nslots <=
// CFTK_FLAG_SLOTS), stored to host_words[b] (the device address of pinned
// host memory) with a system-scope store; every slot is zero again
// afterwards.  One wavefront an area.
hipError_t cftk_flags_gather(uint32_t *slots, int nslots, uint32_t *host_words, int nareas,
                             hipStream_t st);

// One launch of the reductions' tree: for each of nseg segments of m
// partials (segment s at src[s*m ..)), groups of sixteen consecutive partials
// become one node at dst[s*m2 + g], m2 = ceil(m/16).  leaf (the first launch
// only): 0 the partial itself, 1 its rounded product with srcb's, 2 its
// magnitude, 3 its rounded square; kind: 0 ADD nodes, 1 maximum nodes.  Flags are ORed into
// segflags[s] (per_seg), else into the flag slots at segflags (CFTK_FLAG_SLOTS
// of them, slot = workgroup & (CFTK_FLAG_SLOTS - 1)); segspec, when given, gets bit 0 where
// a partial of the segment is an infinity and bit 1 where it is a signalling
// NaN.
hipError_t cftk_reduce(int fmt, const void *src, const void *srcb, void *dst, uint64_t nseg,
                       uint64_t m, uint64_t m2, int leaf, int kind, int rnd, int per_seg,
                       uint32_t *segflags, uint32_t *segspec, int grid, hipStream_t st);

// p[0..n) = v
hipError_t cftk_fill32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);

// The sequencer interpreter over lanes [A->lane0, A->n) in workgroups of
// `waves` wavefronts.  nregs: the program's registers after renaming;
// phys_a/b/c: the renamed registers r0/r1/r2 (-1 where the program never
// names them); trip_pc: the REPEAT whose count A->trips overrides per lane.
hipError_t cftk_seq(int fmt, const SeqArgs *A, uint32_t nregs, int phys_a, int phys_b,
                    int phys_c, uint32_t trip_pc, int waves, hipStream_t st);

// The batched runner: `blocks` persistent workgroups of `waves` wavefronts.
hipError_t cftk_seq_batch(int fmt, const SeqArgs *A, const SeqBatch *B, uint32_t nregs,
                          int phys_a, int phys_b, int phys_c, uint32_t trip_pc, int waves,
                          int blocks, hipStream_t st);

// Workgroups of the batched runner one CU holds at once.
int cftk_seq_batch_blocks_per_cu(int fmt, uint32_t n_insns_in_lds, uint32_t n_consts,
                                 uint32_t nregs, int waves);

// The plain form: one instruction over positions [0, n) of an HBM register
// file and scratch pool (rows `stride` apart).
hipError_t cftk_seq_op(int fmt, const SeqInsn *in, void *regs, void *pool, uint64_t stride,
                       uint64_t n, const uint32_t *bank, int strict, uint32_t depth,
                       int log2depth, uint32_t *words, hipStream_t st);

// The plain form's segment start (store = 0) or end (store = 1) for
// positions [0, n), position p holding lane order[p].
hipError_t cftk_seq_lanes(int fmt, void *regs, void *pool, uint64_t stride, uint64_t n,
                          uint32_t nregs, uint32_t reach, uint32_t n_io, const uint32_t *order,
                          uint32_t *state, const uint32_t *a, const uint32_t *b,
                          const uint32_t *c, int pa, int pb, int pc, int store,
                          hipStream_t st);

// The wide batched runner (an instance over Wd->width lanes).
hipError_t cftk_seq_wide(int fmt, const SeqArgs *A, const SeqBatch *B, const SeqWide *Wd,
                         int phys_a, int phys_b, int phys_c, int waves, int blocks,
                         hipStream_t st);
size_t cftk_seq_wide_lds(int fmt, uint32_t n_consts, uint32_t nsteps_in_lds, uint32_t width,
                         uint32_t nvregs, int waves);
int cftk_seq_wide_blocks_per_cu(int fmt, uint32_t n_consts, uint32_t nsteps_in_lds,
                                uint32_t width, uint32_t nvregs, int waves);

// The multi-program runner; region_rows in rows of the format's LDS row
// type (16 bytes for binary128/256, 8 for binary64, 4 for binary32).
hipError_t cftk_seq_multi(int fmt, const SeqArgs *A, const SeqMulti *M, int waves, int blocks,
                          hipStream_t st);
size_t cftk_seq_multi_lds(int fmt, uint32_t region_rows, int waves);
size_t cftk_seq_row_bytes(int fmt);
int cftk_seq_multi_blocks_per_cu(int fmt, uint32_t region_rows, int waves);

// LDS bytes one interpreter workgroup takes.
size_t cftk_seq_lds(int fmt, uint32_t n_insns_in_lds, uint32_t n_consts, uint32_t nregs,
                    int waves);

}
