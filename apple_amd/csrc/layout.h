// Host-side block partition of a (possibly batched) tet mesh.
//
// The element kernels process runs of consecutive tets as one block, in
// chunks of nth tets (one tet per thread).  The distinct vertices a block's
// tets touch are its "local slots" (first-touch order).  Within a chunk, the
// 4 nth corners are ranked by (slot, tet, corner): a corner's rank is its
// "position", and the corners of one slot occupy a contiguous position range
// (the chunk's slot offsets).  A thread writes each corner's contribution at
// its position in LDS; the slot sums then read contiguous runs.  Each block
// writes one partial sum per slot into a vertex-major buffer: the partials
// of one vertex are contiguous, in ascending block order (the vertex CSR
// below), so the vertex-wise pass reads one contiguous run per vertex.
// Every sum runs in a fixed order, so every result is bitwise reproducible --
// this replaces the float atomics of
// warp/fem/_base.py:261,288-289,321-322,348-349,381.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace apl {

struct Layout {
  int nth = 256;         // threads per block (one tet per thread per chunk)
  int chunks = 1;        // max chunks of nth tets per block
  int nloc_cap = 1024;   // max local slots per block
  int64_t T = 0, V = 0, n_inst = 0;
  int64_t nblocks = 0;
  int nloc_max = 0;        // max local slots over all blocks
  int64_t n_slots = 0;     // total local slots = number of partial sums per pass
  int64_t n_shared_vtx = 0;  // vertices touched by more than one block (statistic)

  std::vector<int32_t> blk_tet_off;   // nblocks+1
  std::vector<int32_t> blk_loc_off;   // nblocks+1 (into loc_gid; slot id = blk_loc_off[b] + s)
  std::vector<int32_t> blk_inst;      // nblocks
  std::vector<int64_t> inst_blk_off;  // n_inst+1
  std::vector<int32_t> loc_gid;       // n_slots: vertex of each slot
  std::vector<int32_t> blk_co_off;    // nblocks+1 into chunk_off
  std::vector<uint16_t> chunk_off;    // per block and chunk: nloc+1 slot offsets into positions
  std::vector<uint16_t> corner;       // 8*T: per tet, the local slots of its 4 corners, then their positions
  std::vector<int32_t> vtx_slot_off;  // V+1: partial range of every vertex
  std::vector<int32_t> loc_dst;       // n_slots: vertex-major partial index of every slot
  // vertex blocks for the vector kernels: aligned to instance ranges, one
  // vertex per thread of a 256-thread workgroup
  int vb = 256;
  std::vector<int32_t> vblk_off;      // nvblk+1
  std::vector<int32_t> vblk_inst;     // nvblk
  std::vector<int64_t> inst_vblk_off; // n_inst+1
  // segments (a model of several potentials): blocks never straddle a
  // segment; seg_of_blk[b] is the segment of block b (0 without segments)
  std::vector<int32_t> seg_of_blk;    // nblocks
};

// Build the partition.  seg_off (n_seg_total + 1 tet offsets, or NULL): the
// tets are split into consecutive segments that blocks must not straddle
// (instance boundaries must be segment boundaries).  Returns an empty string
// on success, else an error.
std::string build_layout(const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                         const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int nth,
                         int chunks, int nloc_cap, Layout& L, int64_t n_seg_total = 0,
                         const int64_t* seg_off = nullptr);

}  // namespace apl
