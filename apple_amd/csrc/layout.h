// Host-side block partition of a (possibly batched) tet mesh.
//
// The element kernels process runs of consecutive tets as one block.  For a
// block, the distinct vertices its tets touch are its "local slots"; a slot
// is *owned* when every tet incident to that vertex lies in the block (its
// vertex sum is complete inside the block and is written straight to the
// output) and *shared* otherwise (the block writes a partial sum that a
// second, vertex-wise pass adds up).  Both the in-block sums and the
// cross-block sums run over CSR lists in ascending order, so every result is
// bitwise reproducible -- this replaces the float atomics of
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
  int nloc_max = 0;      // max local slots over all blocks
  int64_t n_partials = 0;
  int64_t n_shared_vtx = 0;

  std::vector<int32_t> blk_tet_off;   // nblocks+1
  std::vector<int32_t> blk_loc_off;   // nblocks+1 (into loc_gid)
  std::vector<int32_t> blk_nown;      // nblocks
  std::vector<int32_t> blk_sh_base;   // nblocks (first partial index of the block)
  std::vector<int32_t> blk_inst;      // nblocks
  std::vector<int64_t> inst_blk_off;  // n_inst+1
  std::vector<int32_t> loc_gid;       // total slots
  std::vector<uint16_t> loc_inc_off;  // total slots + nblocks (nloc+1 per block)
  std::vector<uint16_t> loc_inc;      // 4*T: per block, (tl*4 + corner) sorted by slot then id
  std::vector<uint16_t> slot;         // 4*T: local slot of each corner
  std::vector<int32_t> sh_vtx;        // shared vertices (ascending)
  std::vector<int32_t> sh_off;        // n_shared_vtx+1 into sh_src
  std::vector<int32_t> sh_src;        // partial indices, ascending
  // vertex blocks for the vector kernels: aligned to instance ranges
  int vb = 256;
  std::vector<int32_t> vblk_off;      // nvblk+1
  std::vector<int32_t> vblk_inst;     // nvblk
  std::vector<int64_t> inst_vblk_off; // n_inst+1
};

// Build the partition.  Returns an empty string on success, else an error.
std::string build_layout(const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                         const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int nth,
                         int chunks, int nloc_cap, Layout& L);

}  // namespace apl
