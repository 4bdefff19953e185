// gram_layout.h -- tile enumeration of the Gram kernels and the layout of the
// Gram scratch that the factor kernel reads.
#pragma once

#include <stddef.h>

#include "enkfg_internal.h"

namespace enkfg {

// Upper-triangular 16x16 tiles (I <= J) of an MT x MT tile grid, enumerated
// row by row; wave w of the 4 waves of a Gram workgroup owns tiles w, w + 4,
// w + 8, ...
__host__ __device__ constexpr int tri_count(int MT) { return MT * (MT + 1) / 2; }

// The item Gram kernels split a segment's tri_count(MT) tiles into NG groups,
// one wavefront each; the two candidates per MT are lo and lo + 1, lo chosen
// so that a group's accumulators and operands fit the 256 registers of a
// lane (21 tiles: 2 or 3 groups; 28: 3 or 4; 36 and 45: 4 or 5; 55: 5 or 6)
__host__ __device__ constexpr int item_groups_lo(int MT)
{
    return MT <= 6 ? 2 : (MT == 7 ? 3 : (MT <= 9 ? 4 : 5));
}

// The items (segment, tile group) of the item Gram kernels are dealt to the 8
// XCDs in runs of kItemRun.  Block i runs on XCD i % 8 (the hardware
// dispatcher deals blocks round-robin), and XCD x takes runs x, x + 8,
// x + 16, ...  A run of 64 items (8 nodes at 2 groups and 4 obstypes) keeps
// neighbouring nodes, whose lists share rows, on one L2, while the runs
// spread each XCD's share over the whole batch.  Contiguous eighths of the
// batch leave the XCD with the heaviest nodes finishing last (FP64 calc on
// the 51x1500x1800 tile: 683 ms, against 667 ms with runs of 16 to 128 items
// and 734 ms with single items).
constexpr int kItemRun = 64;

__device__ inline int item_of_block()
{
    const int x = (int)(blockIdx.x % 8), j = (int)(blockIdx.x / 8);
    return ((j / kItemRun) * 8 + x) * kItemRun + j % kItemRun;
}

// Item `it` of an item Gram launch with NG tile groups per segment: node
// slot bn of the batch, segment seg, group g.  Items run segment by segment
// in the order a.seg_order (obstypes by decreasing total list length), the
// batch's nodes in order within a segment and the NG groups of a (node,
// segment) adjacent.  So the longest items start first and a launch ends on
// short ones: with the nodes' four segments interleaved, the last few of the
// longest items ran alone at the end of each launch (in a copy of the chunk
// loop with 231, 116, 20 and 20 chunks per segment, 47 against 64 TFLOP/s
// for uniform items).  Within a run of kItemRun items the nodes are
// neighbours with lists of one obstype, which share rows in L2.
struct ItemRef {
    int bn, seg, g;
};

__device__ inline ItemRef item_ref(const CalcArgs& a, int it, int NG)
{
    const int per = a.nbatch * NG;
    const int sr = it / per;
    const int rem = it - sr * per;
    ItemRef r;
    r.bn = rem / NG;
    r.g = rem - r.bn * NG;
    r.seg = a.seg_order ? a.seg_order[sr] : sr;
    return r;
}

// blocks of an item Gram launch: a whole number of rounds of 8 runs, over
// which item_of_block is one to one
inline unsigned item_grid(long long nitem)
{
    const long long round = 8LL * kItemRun;
    return (unsigned)(round * ((nitem + round - 1) / round));
}
__host__ __device__ constexpr int tri_I(int MT, int f)
{
    int I = 0;
    while (f >= MT - I) {
        f -= MT - I;
        ++I;
    }
    return I;
}
__host__ __device__ constexpr int tri_J(int MT, int f)
{
    int I = 0;
    while (f >= MT - I) {
        f -= MT - I;
        ++I;
    }
    return I + f;
}

// Position of element (i, j) of an m x m matrix (m <= 16 B) in the Gram
// scratch: the factor kernel's thread (i / B, j / B) of its 16 x 16 thread
// grid owns the B x B block starting at (B (i / B), B (j / B)), stored as one
// contiguous run of B * B doubles per thread.
template <int B>
__host__ __device__ constexpr size_t gram_index(int i, int j)
{
    return ((size_t)((i / B) * 16 + (j / B)) * B + (i % B)) * B + (j % B);
}

}  // namespace enkfg
