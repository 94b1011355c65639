// Brick (one reference block / MPI rank per brick) descriptors and launchers.
#pragma once

#include "kernels.hpp"

namespace beams {

struct BrickDesc {
  GridParams g;        // the brick's owned points; g.field points into its ghosted storage
  MapParams m;         // m.map: the brick's phase-1 (local) map
  ShadeParams s;       // per-brick sample distance, mesh epsilon and step bound
  void* map_final;     // phase-3 map (seeded re-march), used by the march
  int64_t nverts;      // map vertices
  int64_t seed_off;    // first vertex of this brick in the seed buffer
  int64_t img_off;     // first pixel of this brick's partial image
  int64_t row_off;     // first entry of this brick's map rows in light-sheet order
  int32_t fp[4];       // screen footprint: x0, y0, width, height (pixels)
  float box32[6];      // bounds inflated for the conservative FP32 cull
  // 4 x 4 x 4-blocked copies read by the march when march_blocked is set (BlockFetch): the
  // owned points of the field (made with the bricks) and map_final (made after the map phases)
  const void* field_blk;
  void* map_blk;
  int64_t fblk_ybs, fblk_zbs, mblk_ybs, mblk_zbs;
  uint32_t fblk_bytes, mblk_bytes;
};

// Work-queue counters per persistent phase (one per XCD).
constexpr int kQueueSlots = 8;

// Map work items: 256 vertices of one brick.  Brick b holds items map_item_off[b] ..
// map_item_off[b+1] - 1; its item i covers its vertices 256 i .. 256 i + 255, x fastest or (the
// map kernel with light_rows) in the light-sheet row order of `rows`.
constexpr int kMapItem = 256;

struct BrickFrame {
  const BrickDesc* bricks;
  int32_t nbricks;
  const int32_t* map_item_off;    // nbricks + 1 prefix sums of the bricks' map items
  const uint32_t* rows;           // per brick: its map rows (j | k << 16) in light-sheet order
  int32_t n_map_items;
  const int32_t* px_item_brick;   // per march work item: brick
  const int32_t* px_item_tile;    // per march work item: 16x16 tile index in the footprint
  int32_t n_px_items;
  const int32_t* depth_order;     // bricks by ascending camera-to-centre distance
  unsigned int* queue;            // work-queue counters, kQueueSlots per phase: map-1, exchange, map-3, march, fallback
  float* seeds;                   // non-local opacity per map vertex (phase 2 -> 3)
  // phase 3: the vertices whose seeded re-march is not a copy or a constant 1
  unsigned int* rm_count;         // per brick: how many k_bricks_finalize listed
  int32_t* rm_item_off;           // nbricks + 1 prefix sums of their map items (k_rm_items)
  uint32_t* rm_list;              // per brick, at its seed offset: their x-fastest map indices
  float4* partial;                // per-brick footprint RGBA
  float* partial_depth;           // per-brick footprint entry distance
  int32_t max_coords;             // the largest brick's n0 + n1 + n2 (LDS: coordinates + reciprocal widths)
  int32_t persistent_groups;      // workgroups of the persistent launches
  // schedule knobs (the same results whatever their values)
  int32_t xcd_ranges;             // persistent queues: each XCD drains its own contiguous range first
  int32_t light_rows;             // map kernel: rows in light-sheet order (rows != null)
  int32_t march_xcd_tiles;        // march: tiles dealt to the XCDs in contiguous blocks
  int32_t map_fast, march_fast;   // branch-free corner fetch in the map / march launches (every
                                  // brick's field and map view below 2^32 bytes, strides below 2^24)
  int32_t map_pipeline;           // map launches: next light step's loads in flight (march_vertex_pipe)
  int32_t march_groups_per_cu;    // march launch: at most this many workgroups per CU (0: no cap)
  int32_t march_blocked;          // march launch: reads the blocked copies (field_blk, map_blk)
};

hipError_t launch_bricks_map(const BrickFrame& f, int32_t field_dtype, int32_t map_dtype,
                             int32_t ntf, bool seeded, hipStream_t st);
hipError_t launch_bricks_exchange(const BrickFrame& f, int32_t map_dtype, int64_t total_verts,
                                  hipStream_t st);
// Phase 3's first part: every final map value that needs no march (a copy of the local map where
// the seed is 0, 1 where it is 1), and the list of the other vertices with its map items; then
// launch_bricks_map(seeded) re-marches the listed vertices only.  f.rm_count starts at zero.
hipError_t launch_bricks_finalize(const BrickFrame& f, int32_t map_dtype, hipStream_t st);
hipError_t launch_bricks_march(const BrickFrame& f, int32_t field_dtype, int32_t map_dtype,
                               int32_t ntf, bool use_map, const CameraParams& c, hipStream_t st);
hipError_t launch_bricks_composite(const BrickFrame& f, const CameraParams& c, const ImageOut& o,
                                   hipStream_t st);
hipError_t launch_extract_brick(const void* src, int32_t elem_bytes, const int64_t src_dims[3],
                                const int64_t lo[3], const int64_t dims[3], void* dst,
                                hipStream_t st);
hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t st);
// The owned points lo .. lo + n - 1 of a field (x fastest, dims src_dims) into a 4 x 4 x 4-blocked
// copy (blocked_index with the strides of n).
hipError_t launch_block_extract(const void* src, int32_t elem_bytes, const int64_t src_dims[3],
                                const int64_t lo[3], const int64_t n[3], void* dst, hipStream_t st);
// Every brick's map_final into its blocked copy map_blk.
hipError_t launch_block_maps(const BrickFrame& f, int32_t map_dtype, hipStream_t st);

}  // namespace beams
