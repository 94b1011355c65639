/*
 * Test oracle — NOT product code.
 *
 * A CPU restatement of the Beams lighted volume renderer's hot path
 * (goodbadwolf/pilot, beams/rendering/) used only by tests/, the smoke check
 * and bench.py's cpu_baseline leg to check and time the HIP implementation.
 * Nothing in the product library links or calls this code.
 *
 * Two twins are built from one body (oracle_body.h):
 *   *_f32 — the reference's FP32 arithmetic in the reference's operation order
 *           (no FMA contraction; FP64 only where the reference promotes to
 *           vtkm::Bounds doubles).  This is the parity oracle.
 *   *_f64 — the same algorithm carried out in FP64 end to end.  This is the
 *           precision oracle for reduced-precision storage and for the
 *           re-discretised fast paths.
 *
 * Every pointer typed `const void*` below points at float data for the _f32
 * twin and at double data for the _f64 twin (field, coordinates, transfer
 * function, maps, rays, images).
 */
#ifndef BEAMS_ORACLE_H
#define BEAMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcScene {
  int64_t dims[3];        /* point dimensions of the block's structured grid   */
  const void* coords[3];  /* rectilinear coordinate arrays, dims[i] values each */
  const void* field;      /* scalars, x fastest (points, or cells if assoc=1)   */
  int32_t assoc;          /* 0 = point field, 1 = cell field                    */
  int32_t ntf;            /* transfer-function entries                          */
  const void* tf;         /* ntf x RGBA                                         */
  double scalar_range[2]; /* vtkm::Range (min, max)                             */
  int32_t light_kind;     /* 0 = point light, 1 = directional (extension)       */
  int32_t use_shadow;     /* SetUseShadowMap                                    */
  double light_pos[3];    /* point: position; directional: travel direction     */
  double light_color[3];  /* Color * Intensity                                  */
  int64_t map_cells[3];   /* SetShadowMapSize (cells per axis)                  */
  double sample_distance; /* <= 0 selects |extent| / 200                        */
  double alpha_stop;      /* early-termination threshold (reference: 1.0)       */
  double light_far;       /* directional light: origin distance (<= 0: 4|extent|) */
} OrcScene;

typedef struct OrcRays {
  int64_t n;
  const void* origin[3];
  const void* dir[3];
  void* min_distance; /* in/out */
  void* max_distance; /* in/out */
  void* distance;     /* out    */
  void* rgba;         /* in/out, 4 interleaved values per ray */
  void* nsamples;     /* optional out: int64 composited samples per ray */
} OrcRays;

#define ORC_DECLARE(SUF)                                                                   \
  int orc_segment_aabb_##SUF(const void* start, const void* end, const double* aabb,       \
                             double t_eps, void* t_min, void* t_max);                      \
  int64_t orc_find_cell_##SUF(const OrcScene* s, const void* point, void* pcoords);        \
  double orc_interp_hex_##SUF(const void* values8, const void* pcoords);                   \
  int orc_tmap_build_##SUF(const OrcScene* s, const void* seed, void* map_out, int nthr);  \
  int orc_tmap_build_range_##SUF(const OrcScene* s, const void* seed, void* map_out,      \
                                 int64_t begin, int64_t end, int nthr);                    \
  int orc_map_points_##SUF(const OrcScene* s, void* const* pts);                           \
  int orc_march_points_##SUF(const OrcScene* s, const void* const* pts, int64_t n,         \
                             void* out, int nthr);                                         \
  int orc_map_estimate_##SUF(const OrcScene* s, const void* map, const void* point,        \
                             void* rgb_out, void* opacity_out);                            \
  int orc_calc_ray_start_##SUF(const OrcScene* s, OrcRays* rays, int nthr);                \
  int orc_sample_rays_##SUF(const OrcScene* s, const void* map, OrcRays* rays, int nthr);  \
  int orc_render_rays_##SUF(const OrcScene* s, void* map_out, OrcRays* rays, int nthr);    \
  int64_t orc_nonlocal_hits_##SUF(int nblocks, const OrcScene* scenes, int self,           \
                                  int64_t* counts, int64_t* offsets, void* hits_out,       \
                                  int64_t hits_cap);                                       \
  int orc_nonlocal_seed_##SUF(int nblocks, const OrcScene* scenes, const void* const* maps,\
                              int self, void* seed_out, int nthr);                         \
  int orc_nonlocal_seed_range_##SUF(int nblocks, const OrcScene* scenes,                   \
                                    const void* const* maps, int self, int64_t begin,      \
                                    int64_t end, void* seed_out, int nthr);                \
  int orc_composite_order_##SUF(int nblocks, const int* order, const void* const* images,   \
                                int64_t npix, void* out);                                 \
  int orc_composite_##SUF(int nblocks, const double* block_bounds, const double* cam_pos,  \
                          const void* const* images, int64_t npix, void* out);

ORC_DECLARE(f32)
ORC_DECLARE(f64)

/* One 32-byte transmittance hit record, TransmittanceMap.h:23-31 (FP32 twin);
 * the FP64 twin uses the same layout with doubles (56 bytes). */
typedef struct OrcHitF32 {
  int32_t ray_id, block_id, from_block_id;
  float point[3];
  float ray_t;
  float opacity;
} OrcHitF32;

typedef struct OrcHitF64 {
  int32_t ray_id, block_id, from_block_id, pad;
  double point[3];
  double ray_t;
  double opacity;
} OrcHitF64;

#ifdef __cplusplus
}
#endif
#endif
