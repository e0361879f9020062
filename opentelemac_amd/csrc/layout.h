// Host-side construction of the block layout used by the EBE kernels.
//
// Nodes are renumbered for locality (Hilbert curve of the coordinates;
// prisms: whole vertical columns kept together) and cut into blocks of
// consecutive internal node ids. Each workgroup owns one block: it processes
// every element touching one of its owned nodes (elements on a block border
// are therefore processed by every block they touch), accumulates in LDS and
// writes only its owned nodes. Within a block the elements are coloured so
// that no two elements of one colour share an owned node; colours run one
// after the other with a barrier between them, so the LDS scatter-add needs
// no atomics and its summation order is fixed (bitwise reproducible).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bhip {

constexpr int kMaxColors = 48;
// threads of an element-kernel workgroup: one colour phase holds at most
// this many element instances where the colouring allows it
constexpr int kPhaseSlots = 256;

struct Layout {
  int npoin = 0, nelem = 0, ndp = 3, nplan = 1;
  int block_nodes = 0;               // owned nodes per block (last may be fewer)
  int nblocks = 0;
  int nslot = 0;                     // element instances over all blocks
  int nslot_pad = 0;                 // column stride of slot arrays
  int max_colors = 0;                // max colours used by one block
  int max_local = 0;                 // max owned + halo nodes of one block
  int max_owned = 0;
  int64_t halo_total = 0;
  bool col_ids = false;              // prisms: lid[k + 3][s] == lid[k][s] + 1
  std::vector<int32_t> perm;         // internal node -> caller node
  std::vector<int32_t> iperm;        // caller node -> internal node
  std::vector<int32_t> node_start;   // [nblocks+1] internal node ranges
  std::vector<int32_t> halo_start;   // [nblocks+1] ranges in halo
  std::vector<int32_t> halo;         // internal node ids of halo nodes
  std::vector<int32_t> color_off;    // [nblocks*(kMaxColors+1)] slot offsets
  std::vector<int32_t> ncolor;       // [nblocks]
  std::vector<uint16_t> lid;         // [ndp][nslot_pad] local node ids
  std::vector<int32_t> slot_elem;    // [nslot] caller element id
};

// Builds and validates the layout. Returns an empty string on success or a
// description of the problem (invalid connectivity, a block that does not
// fit the 16-bit local ids or the colour limit).
std::string build_layout(int npoin, int nelem, int ndp, const int32_t* ikle,
                         const double* x, const double* y, int nplan,
                         int block_nodes, Layout& L);

// Layout of an ensemble of nsys independent meshes (each laid out on its
// own, then concatenated; no block holds nodes of two meshes). ikle holds
// the meshes' connectivities one after the other (mesh s: ndp x nelem[s],
// structure of arrays, its own 0-based node ids); x, y (may be NULL) the
// coordinates concatenated. sys_blocks receives the block range of each
// mesh, ikle_union the connectivity of the union in global node ids.
std::string build_layout_multi(int nsys, const int* npoin, const int* nelem, int ndp,
                               const int32_t* ikle, const double* x, const double* y,
                               int block_nodes, Layout& L, std::vector<int32_t>& sys_blocks,
                               std::vector<int32_t>& ikle_union);

// Independent check of every invariant the kernels rely on.
std::string validate_layout(const Layout& L, const int32_t* ikle);

}  // namespace bhip
