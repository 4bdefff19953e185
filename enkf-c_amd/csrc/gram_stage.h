// gram_stage.h -- the chunk pipeline of the item Gram kernels: list entries
// and S rows of a (node, obstype) list copied from global memory straight
// into a wave's private LDS area, several chunks ahead.
#pragma once

#include "enkfg_internal.h"

namespace enkfg {

constexpr int kChunk = 16;   // local obs per chunk

// The list entries and S rows of a chunk are copied from global memory
// straight into LDS (global_load_lds: lane l of an instruction writes its 4 or
// 16 bytes at M0 + 4 l or M0 + 16 l), with no VGPRs held for data in flight.
// Block-major rows: the 16 floats of block X (columns 16X ..) of row r at
// rows[X][r][0..15]; four rows 4s + kq, columns 16X + col, read by one
// instruction fall in 64 distinct banks with no padding.  NB chunk buffers:
// the rows of chunks c+1 .. c+NB-1 are in flight while chunk c is used; list
// entries are copied one chunk further ahead, into NB + 1 slots.
template <int MT, int NB>
struct ChunkLds {
    static_assert(NB >= 2, "at least two chunk buffers");
    alignas(16) float rows[NB][MT][kChunk][16];   // chunk c in rows[c % NB]
    alignas(16) float sv[NB][64];   // s of the chunk's rows as FP64 in dwords 0..31 (32..63: a copy)
    // list entries of chunk c in meta[c % (NB + 1)], as stored (past the
    // list's end too): ids in dwords 0..15, tapers (FP64) in 16..47
    alignas(16) unsigned meta[NB + 1][64];
};

// the copy of the list entries k0 .. k0 + 15 into meta[slot] (lanes 48..63
// repeat the ids into dwords 48..63).  Past the list's end the entries are
// read all the same (the list arrays are padded) and masked where used.
template <int MT, int NB>
__device__ __attribute__((always_inline)) inline void chunk_meta_copy(const CalcArgs& a, ChunkLds<MT, NB>& sh,
                                                                      long long off, int k0, int slot)
{
    // the builtin's LDS operand has no form in the host pass, which never
    // emits these functions
#if defined(__HIP_DEVICE_COMPILE__)
    const int lane = threadIdx.x & 63;
    const unsigned* src = (lane >= 16 && lane < 48)
                              ? reinterpret_cast<const unsigned*>(a.lc + off + k0 + ((lane - 16) >> 1)) + (lane & 1)
                              : reinterpret_cast<const unsigned*>(a.ids + off + k0 + (lane & 15));
    __builtin_amdgcn_global_load_lds(src, &sh.meta[slot][0], 4, 0, 0);
#endif
}

// the MT + 1 copies of the rows of the chunk at list entry k0 (entries in
// meta[slot]) into rows[buf] and sv[buf]: instruction X brings block X of the
// 16 rows (lane l: row l / 4, columns 16X + 4 (l % 4) ..), the last one s of
// the rows.  Rows past the list's end and the columns ms .. MP of the last
// block come from the zero row (row nobs of S, s = 0).
template <int MT, int NB>
__device__ __attribute__((always_inline)) inline void chunk_row_copy(const CalcArgs& a, ChunkLds<MT, NB>& sh, int slot,
                                                                     int buf, int k0, int cnt)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const int lane = threadIdx.x & 63;
    const int r = lane >> 2;
    const int rs = (lane & 31) >> 1;
    // loaded unconditionally, then selected (the compiler would otherwise
    // move the LDS reads under a branch)
    int idr = (int)sh.meta[slot][r];
    int sidr = (int)sh.meta[slot][rs];
    asm volatile("" : "+v"(idr), "+v"(sidr));
    const int id = k0 + r < cnt ? idr : a.nobs;
    const int sid = k0 + rs < cnt ? sidr : a.nobs;
    const float* zero = a.S + (size_t)a.nobs * a.ms;
    const float* row = a.S + (size_t)id * a.ms + 4 * (lane & 3);
#pragma unroll
    for (int X = 0; X < MT - 1; ++X)
        __builtin_amdgcn_global_load_lds(row + 16 * X, &sh.rows[buf][X][0][0], 16, 0, 0);
    const bool inrow = 16 * (MT - 1) + 4 * (lane & 3) < a.ms;
    __builtin_amdgcn_global_load_lds(inrow ? row + 16 * (MT - 1) : zero, &sh.rows[buf][MT - 1][0][0], 16, 0, 0);
    const unsigned* s = reinterpret_cast<const unsigned*>(a.sv + sid) + (lane & 1);
    __builtin_amdgcn_global_load_lds(s, &sh.sv[buf][0], 4, 0, 0);
#endif
}

// Waits, at the top of chunk c, for all copies but the newest chunk's MT + 1
// (NB >= 3: issued in this order, oldest first: rows of chunks c .. c+NB-3,
// list entries of chunk c+NB-1, rows of chunk c+NB-2), or for all of them
// (NB = 2).  The compiler does not count the LDS writes of global_load_lds
// against the LDS reads that follow, so the pipeline waits by hand; the
// memory clobber keeps the compiler's loads and stores on their side of the
// wait.
template <int MT, int NB>
__device__ __attribute__((always_inline)) inline void chunk_wait()
{
    static_assert(MT + 1 < 16, "vmcnt field");
    if constexpr (NB >= 3)
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(MT + 1) : "memory");
    else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

__device__ __attribute__((always_inline)) inline void chunk_fence()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// The prologue, in the order of issue the wait counts on: list entries of
// chunks 0 .. NB-2, copies of chunks 0 .. NB-3, list entries of chunk NB-1,
// copies of chunk NB-2.
template <int MT, int NB>
__device__ __attribute__((always_inline)) inline void chunk_prologue(const CalcArgs& a, ChunkLds<MT, NB>& sh,
                                                                     long long off, int cnt)
{
#pragma unroll
    for (int p = 0; p < NB - 1; ++p)
        chunk_meta_copy<MT, NB>(a, sh, off, p * kChunk, p);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    chunk_fence();
#pragma unroll
    for (int p = 0; p < NB - 2; ++p)
        chunk_row_copy<MT, NB>(a, sh, p, p, p * kChunk, cnt);
    chunk_fence();
    chunk_meta_copy<MT, NB>(a, sh, off, (NB - 1) * kChunk, NB - 1);
    chunk_fence();
    chunk_row_copy<MT, NB>(a, sh, NB - 2, NB - 2, (NB - 2) * kChunk, cnt);
    chunk_fence();
}

// The top of chunk c (buf = c % NB, slot = c % (NB + 1)): the wait, then the
// list entries of chunk c+NB and the copies of chunk c+NB-1 into the slot and
// buffer chunk c-1 was read from.  Past the list's end these read padding and
// the zero row, issued all the same: the count of the wait takes every
// chunk to issue them.
template <int MT, int NB>
__device__ __attribute__((always_inline)) inline void chunk_advance(const CalcArgs& a, ChunkLds<MT, NB>& sh,
                                                                    long long off, int cnt, int c, int buf, int slot)
{
    constexpr int NM = NB + 1;
    chunk_wait<MT, NB>();
    chunk_fence();
    chunk_meta_copy<MT, NB>(a, sh, off, (c + NB) * kChunk, slot == 0 ? NM - 1 : slot - 1);
    chunk_fence();
    chunk_row_copy<MT, NB>(a, sh, slot >= 2 ? slot - 2 : slot + NM - 2, buf == 0 ? NB - 1 : buf - 1,
                           (c + NB - 1) * kChunk, cnt);
    chunk_fence();
}

}  // namespace enkfg
