/*
 * beams_mi355x.h — C ABI of the MI355X (gfx950) lighted volume renderer.
 *
 * Drop-in replacement for the hot path of Beams (goodbadwolf/pilot): the
 * transmittance-map build and the shadowed camera-ray march behind
 * beams::rendering::LightedVolumeRenderer and MapperLightedVolume.  Each entry
 * point names the reference interface it replaces (file:line in the pilot
 * repository).
 *
 * Conventions
 *  - Every call returns BEAMS_OK (0) or a BEAMS_ERR_* status; no exception or
 *    signal crosses this boundary.  beams_last_error() gives the message of the
 *    last failing call on a handle.
 *  - One handle owns one HIP stream and its device buffers.  A handle is not
 *    thread-safe; separate handles may be used from separate threads.
 *  - Buffers are caller-owned.  `mem` says whether a pointer is host memory
 *    (BEAMS_MEM_HOST: copied in/out synchronously) or device memory on the
 *    handle's device (BEAMS_MEM_DEVICE: used in place, stream-ordered).
 *  - Arrays are x-fastest, matching vtkm point/cell ordering.
 */
#ifndef BEAMS_MI355X_H
#define BEAMS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct beams_renderer beams_renderer;

/* status codes */
#define BEAMS_OK 0
#define BEAMS_ERR_INVALID 1     /* bad argument: vtkm::cont::ErrorBadValue in the reference */
#define BEAMS_ERR_UNSUPPORTED 2 /* unsupported dtype / association / mode                   */
#define BEAMS_ERR_STATE 3       /* call order (e.g. render before set_field)               */
#define BEAMS_ERR_HIP 4         /* HIP runtime failure                                     */
#define BEAMS_ERR_NO_DEVICE 5   /* no gfx950 device visible                                */

/* element types */
#define BEAMS_F32 0
#define BEAMS_F16 1
#define BEAMS_BF16 2
#define BEAMS_F64 3

/* memory kinds */
#define BEAMS_MEM_HOST 0
#define BEAMS_MEM_DEVICE 1

/* field association (vtkm::cont::Field::Association) */
#define BEAMS_ASSOC_POINTS 0
#define BEAMS_ASSOC_CELLS 1

/* light kinds; the reference has point lights only (PointLight.h:15-30) */
#define BEAMS_LIGHT_POINT 0
#define BEAMS_LIGHT_DIRECTIONAL 1

/* shadow-map construction */
#define BEAMS_SHADOW_REF_MARCH 0 /* per-vertex light march, TransmittanceMap.h:239-340 */
#define BEAMS_SHADOW_SWEEP 1     /* exact sweep: light-aligned slices carrying optical depth
                                    (one pass; needs the light outside the bounds' slab on
                                    one axis; a re-discretisation, not bit-equal) */
#define BEAMS_SHADOW_LEARNED 2   /* coarse 32^3-mip optical depth + network correction per map
                                    vertex, checked against the light march at sparse
                                    vertices; falls back to an exact mode over the budget */

/* phases reported by beams_get_timings (Beams.cxx:30-34 Phase1..5Time) */
#define BEAMS_PHASE_MAP 0       /* phase 1: transmittance map build           */
#define BEAMS_PHASE_EXCHANGE 1  /* phase 2: non-local hit exchange (bricks)   */
#define BEAMS_PHASE_REMAP 2     /* phase 3: seeded map re-march (bricks)      */
#define BEAMS_PHASE_MARCH 3     /* phase 4: CalcRayStart + Sampler            */
#define BEAMS_PHASE_COMPOSITE 4 /* phase 5: brick composite + background      */
#define BEAMS_PHASE_FRAME 5     /* the whole frame (also with graph replay)   */
#define BEAMS_NUM_PHASES 6

/* Pinhole camera for beams_render_image.  Replaces the vtkm::rendering::Camera ->
 * raytracing::Camera::CreateRays step (MapperLightedVolume.cxx:141-152). */
typedef struct beams_camera {
  double eye[3];
  double look_at[3];
  double up[3];
  double fov_y_deg;  /* vertical field of view                            */
  double near_dist;  /* initial Ray::MinDistance (0 for a fresh canvas)  */
  double far_dist;   /* initial Ray::MaxDistance (<= 0: unbounded)       */
} beams_camera;

/* ---- lifetime: LightedVolumeRenderer::LightedVolumeRenderer (LightedVolumeRenderer.cxx:752-759)
 * Defaults follow the reference: sample distance unset (|extent|/200 at render),
 * shadow map on, 16^3 map cells, early termination at alpha >= 1. */
int beams_create(beams_renderer** out);
int beams_destroy(beams_renderer* h);
const char* beams_last_error(const beams_renderer* h);
/* Wait for all work queued on the handle's stream. */
int beams_synchronize(beams_renderer* h);

/* ---- LightedVolumeRenderer::SetData (LightedVolumeRenderer.cxx:766-779) -------------------
 * values: scalars, dims[0]*dims[1]*dims[2] of them for points, (dims-1) products for cells.
 * dims: point dimensions of the structured grid (CellSetStructured<3>).
 * coords: three rectilinear coordinate arrays (ArrayHandleCartesianProduct), host memory,
 *   dims[i] increasing values each; or NULL for uniform coordinates given by origin and
 *   spacing (held as x_i = float(origin) + float(spacing) * i, what ArrayHandleCounting holds).
 * dtype: BEAMS_F32, BEAMS_F16 or BEAMS_BF16 storage; arithmetic is FP32.
 * Limits (BEAMS_ERR_INVALID / BEAMS_ERR_UNSUPPORTED otherwise): each dims[i] in [2, 2^20],
 *   dims[0]*dims[1] below 2^31, dims[0]+dims[1]+dims[2] at most 16384 (the coordinate tables
 *   are staged in LDS), coordinates finite with magnitude at most 2^60 and cell widths at
 *   least 2^-60 (the range of the kernels' corrected-reciprocal quotients).
 * Joint limit, checked by the render calls (BEAMS_ERR_UNSUPPORTED): 16 B per colormap entry
 *   plus 8 B per coordinate point (12 B for unevenly spaced coordinates, which carry bin
 *   tables) within 159 KiB, the LDS a workgroup stages; the learned closure takes at most
 *   1024 colormap entries. */
int beams_set_field(beams_renderer* h, const void* values, int32_t dtype, int32_t assoc,
                    const int64_t dims[3], const float* const coords[3], const double origin[3],
                    const double spacing[3], int32_t mem);

/* ---- BoundsMap (BoundsMap.cxx:77-127, one block per rank) -> bricks on this device ---------
 * Cuts the field of beams_set_field into nbricks blocks, each rendered as the reference
 * renders one MPI rank's block: point_ranges holds 6 inclusive point-index bounds per brick
 * (i0,i1,j0,j1,k0,k1; neighbours share their boundary plane), each brick is stored with
 * `ghost` extra point layers (clipped at the field's edge; outside the brick's bounds), and
 * map_cells gives 3 map cells per brick (NULL: the brick's own cell counts; fewer than 2^31
 * map vertices per brick, else BEAMS_ERR_UNSUPPORTED).  A frame then
 * runs, for all bricks at once: local maps (phase 1), the non-local hit exchange and
 * composite (TransmittanceMap.h:342-513, LightedVolumeRenderer.cxx:968-1235, phase 2), the
 * seeded re-march (phase 3), the per-brick march (phase 4) and the depth-ordered over
 * composite (MapperLightedVolume.cxx:215-303, phase 5).  nbricks = 0 returns to one block;
 * beams_set_field drops the bricks. */
int beams_set_bricks(beams_renderer* h, int32_t nbricks, const int64_t* point_ranges,
                     int32_t ghost, const int64_t* map_cells);

/* Copy of brick b's map: which = 0 the phase-1 local map, 1 the final (seeded) map;
 * (cells+1)^3 FP32 values. */
int beams_get_brick_map(beams_renderer* h, int32_t brick, int32_t which, float* out);

/* Workgroups of the persistent brick launches (a schedule knob; default 2048). */
int beams_set_persistent_groups(beams_renderer* h, int32_t groups);

/* Schedule of the brick launches: how work is dealt to the GPU, never what it computes (the
 * maps and images are the same bit for bit under every schedule).  Defaults: 2048 persistent
 * workgroups, every switch on.  Extension: the reference schedules through VTK-m's device
 * adapter. */
typedef struct beams_schedule {
  int32_t persistent_groups; /* workgroups of the persistent map and exchange launches, >= 1    */
  int32_t xcd_ranges;        /* 1: each XCD first drains its own contiguous range of map items */
  int32_t light_rows;        /* 1: the map launches walk each brick's rows in light-sheet order */
  int32_t march_xcd_tiles;   /* 1: the march deals screen tiles to the XCDs in contiguous blocks */
  int32_t map_fast_fetch;    /* 1: the map launches take the branch-free corner fetch (24-bit
                                offsets) when every brick's field and map allow it              */
  int32_t march_fast_fetch;  /* 1: the same for the march launch                                */
  int32_t map_pipeline;      /* 1: the map launches issue each light step's corner loads one
                                step ahead                                                      */
  int32_t sweep_block;       /* 1: the exact sweep advances several slices per launch from an
                                LDS halo (one sigma pass first); 0: one launch per slice.  The
                                maps are bit-identical; a light too oblique for the halo always
                                takes one launch per slice                                      */
  int32_t march_groups_per_cu; /* 1..16: at most that many march workgroups share a CU (a larger
                                  share of the L2 per screen tile); 0: no cap                   */
  int32_t march_blocked;     /* 1: the brick march reads 4 x 4 x 4-blocked copies of the bricks'
                                fields (made with the bricks) and final maps (made after the map
                                phases), so a wave's samples touch fewer cache lines            */
} beams_schedule;
int beams_set_schedule(beams_renderer* h, const beams_schedule* s);
int beams_get_schedule(const beams_renderer* h, beams_schedule* s);

/* Order in which the bricks' partial images are composited (bricks only).
 * BEAMS_ORDER_CENTER_DISTANCE (default) is the reference's: ascending |camera - block centre|
 * (MapperLightedVolume.cxx:222-236), which can put a farther brick first when bricks differ
 * in size.  BEAMS_ORDER_KD_FRONT_TO_BACK (an extension) is a visibility order of a kd
 * decomposition: split the bricks by an axis-aligned plane that no brick straddles, the
 * camera's side first, recursively; a frame fails with BEAMS_ERR_UNSUPPORTED when some set of
 * bricks has no such plane (beams_gen_kd_bricks always has one). */
#define BEAMS_ORDER_CENTER_DISTANCE 0
#define BEAMS_ORDER_KD_FRONT_TO_BACK 1
int beams_set_composite_order(beams_renderer* h, int32_t order);

/* Frames of beams_render_image replay a hipGraph captured from their launches; the graph
 * is captured again whenever a kernel argument changes (camera, outputs, scene state). */
int beams_set_graph(beams_renderer* h, int32_t enabled);

/* Times `frames` frames of beams_render_image for candidate schedules of the brick launches
 * (coordinate descent from the current schedule: the persistent workgroup count, then each
 * switch of beams_schedule but sweep_block, a flip kept only when it gains more than 0.5%),
 * keeps the fastest (beams_get_schedule reads it) and reports its workgroup count and time. */
int beams_autotune(beams_renderer* h, const beams_camera* cam, int32_t width, int32_t height,
                   int32_t frames, int32_t* best_groups, float* best_ms);

/* scalarRange argument of SetData (vtkm::Range). */
int beams_set_scalar_range(beams_renderer* h, double min_value, double max_value);

/* ---- LightedVolumeRenderer::SetColorMap (LightedVolumeRenderer.cxx:761-764) --------------
 * n RGBA float entries (ArrayHandle<Vec4f_32>), host memory. */
int beams_set_colormap(beams_renderer* h, const float* rgba, int64_t n);

/* ---- LightedVolumeRenderer::ClearLights + AddLight (LightedVolumeRenderer.cxx:835-846) ---
 * Point light: xyz = position.  Directional (extension): xyz = direction of travel.
 * The light colour used is color * intensity, as Lights::AddLight stores it. */
int beams_set_light(beams_renderer* h, int32_t kind, const float xyz[3], const float color[3],
                    float intensity);

/* ---- SetUseShadowMap / SetShadowMapSize (LightedVolumeRenderer.h:59-62) ------------------
 * map_cells: cells per axis of the transmittance map (the map has cells+1 vertices).
 * mode: BEAMS_SHADOW_REF_MARCH (the reference), BEAMS_SHADOW_SWEEP or BEAMS_SHADOW_LEARNED
 * (extensions; bricks take REF_MARCH or LEARNED).  map_dtype: BEAMS_F32 or BEAMS_F16 storage. */
int beams_set_shadow(beams_renderer* h, int32_t enabled, const int64_t map_cells[3],
                     int32_t mode, int32_t map_dtype);

/* ---- LightedVolumeRenderer::SetSampleDistance (LightedVolumeRenderer.cxx:1374-1379) ------
 * Non-positive distances are rejected (BEAMS_ERR_INVALID), as the reference throws. */
int beams_set_sample_distance(beams_renderer* h, float distance);

/* Early-termination threshold on accumulated alpha; the reference uses 1.0
 * (LightedVolumeRenderer.cxx:491).  Extension. */
int beams_set_alpha_stop(beams_renderer* h, float alpha);

/* ---- LightedVolumeRenderer::Render(Ray<Float32>&) (LightedVolumeRenderer.cxx:807-825,910-1372)
 * The parity entry point: builds the transmittance map, then runs CalcRayStart and the
 * sampler over n rays given as the Ray<Float32> SoA.  min_d/max_d/dist are in/out as in
 * the reference (MinDistance, MaxDistance, Distance); rgba is Buffers[0], four
 * interleaved floats per ray, read and accumulated in place. */
int beams_render_rays(beams_renderer* h, int64_t n, const float* const origin[3],
                      const float* const dir[3], float* min_d, float* max_d, float* dist,
                      float* rgba, int32_t mem);

/* beams_render_rays without allocating: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The same map build and march as beams_render_rays with BEAMS_MEM_DEVICE
 * rays (device memory only, used in place, stream-ordered on the handle's stream, nothing
 * waited for), but the transmittance map goes to a scratch array of 64 MiB that the library
 * holds per device from load until the process ends (at most 2^24 FP32 or 2^25 FP16 map
 * vertices; larger maps are refused with BEAMS_ERR_INVALID).  Every handle on a device shares
 * that device's array and it keeps the last call's map: calls that overlap on one device race.
 * Single block, reference-march or no shadow map only (BEAMS_ERR_UNSUPPORTED otherwise); no
 * phase timings or sample counts are recorded, and beams_get_map does not see this map. */
int beams_render_rays_noalloc(beams_renderer* h, int64_t n, const float* const origin[3],
                              const float* const dir[3], float* min_d, float* max_d, float* dist,
                              float* rgba);

/* ---- MapperLightedVolume::RenderCells (MapperLightedVolume.cxx:114-193) ------------------
 * Full frame for a width x height image: camera rays, map build, march, optional
 * background blend (background may be NULL: none; else RGBA blended under the image as
 * CanvasRayTracer::BlendBackground does).  rgba: width*height*4 floats; depth:
 * width*height floats (the ray entry distance, +inf where the ray misses), may be NULL. */
int beams_render_image(beams_renderer* h, const beams_camera* cam, int32_t width,
                       int32_t height, const float* background, float* rgba, float* depth,
                       int32_t mem);

/* Copy of the last transmittance (opacity) map, (cells+1)^3 FP32 values; the debugging
 * dump SaveTransmittance (TransmittanceMap.h:527-537) writes the same field.  After a change
 * of field, bricks, light, colour table or shadow settings the map is stale and this returns
 * BEAMS_ERR_STATE until the next frame with shadows builds one. */
int beams_get_map(beams_renderer* h, float* out, int32_t mem);

/* Per-phase device times of the last render in milliseconds (BEAMS_NUM_PHASES values),
 * replacing the Phase1..5Time globals (Beams.cxx:30-34). */
int beams_get_timings(beams_renderer* h, float* phase_ms, int32_t n);

/* Mean per-phase device times over the frames rendered since the last reset (at most
 * the last 256 frames), and their number; the Profiler's per-frame collection
 * (Profiler.cxx:24-74) in one call.  beams_reset_timings starts a new window. */
int beams_get_timing_stats(beams_renderer* h, float* avg_ms, int32_t n, int32_t* nframes);
int beams_reset_timings(beams_renderer* h);

/* Optional device counters of the next renders: composited camera-ray samples (the
 * iterations of the Sampler loop that composite, LightedVolumeRenderer.cxx:404-493)
 * and light-march steps inside the volume (TransmittanceMap.h:299-331) of the last
 * frame.  Extension; counting adds one atomic per ray. */
int beams_set_sample_counting(beams_renderer* h, int32_t enabled);
int beams_get_sample_counts(beams_renderer* h, int64_t* camera_samples, int64_t* light_steps);

/* ---- DataSetFactory / SceneFactory inputs (DataSetFactory.h:15, SceneFactory.h:13) -------
 * Seeded synthetic scenes for the benchmark configurations. */
typedef struct beams_field_recipe {
  uint64_t seed;
  int32_t nblobs;        /* isotropic Gaussian blobs                                   */
  int32_t noise_octaves; /* lattice value-noise octaves (0: none)                      */
  double noise_amp;      /* amplitude of octave 0                                      */
  double noise_freq;     /* base lattice frequency                                     */
  double time;           /* time step: blobs drift, noise phase advances               */
  double drift_speed;    /* blob speed in box lengths per time step                    */
  double noise_phase_rate; /* lattice offset per time step                             */
} beams_field_recipe;

/* Field on the unit cube sampled at dims points per axis, normalised to [0,1].
 * out: dims product values of dtype (BEAMS_F32/F16/BF16/F64) in `mem` memory.
 * Host and device generation give identical values. */
int beams_gen_field(const beams_field_recipe* r, const int64_t dims[3], int32_t dtype,
                    void* out, int32_t mem);

/* Piecewise-linear RGBA table of n entries between npoints seeded control points
 * (x = 0 and x = 1 included, RGBA uniform in [0,1]).  rgba32 and/or rgba64 may be NULL. */
int beams_gen_colormap(uint64_t seed, int32_t npoints, int64_t n, float* rgba32, double* rgba64);

/* Seeded kd decomposition of a grid of dims points into nbricks bricks whose cell extents
 * per axis lie in [min_cells, max_cells]; point_ranges receives 6 values per brick in the
 * layout of beams_set_bricks. */
int beams_gen_kd_bricks(uint64_t seed, const int64_t dims[3], int32_t nbricks, int64_t min_cells,
                        int64_t max_cells, int64_t* point_ranges);

/* Orbit camera looking at the centre of bounds (xmin,xmax,ymin,ymax,zmin,zmax) from a
 * distance of one box diagonal, at azimuth/elevation in degrees, up = +y. */
int beams_gen_orbit_camera(const double bounds[6], double azimuth_deg, double elevation_deg,
                           double fov_y_deg, beams_camera* out);

/* The FP32 rays beams_render_image generates for cam at width x height, as SoA
 * (host memory, width*height each, pixel index = row * width + column, row 0 on top);
 * dir64/org64 (may be NULL) receive the same rays before rounding to FP32. */
int beams_gen_camera_rays(const beams_camera* cam, int32_t width, int32_t height,
                          float* const origin[3], float* const dir[3], double* const origin64[3],
                          double* const dir64[3]);

/* ==== Learned shadow closure (extension: no counterpart in the reference) =================
 * A network of 32 inputs (31 features + a zero pad) -> 64 -> 64 -> 1 with ReLU, BF16 weights
 * and activations, FP32 accumulation, run on the matrix cores.  It predicts, per map vertex,
 * y = min(tau_coarse, 16) - min(tau_exact, 16) with tau = -ln(1 - opacity), from the coarse
 * optical depth, the 3x3x3 alpha/max-alpha patch around the nearest field vertex and the
 * light direction.  Parameters are FP32 master copies packed as W1[32][64], b1[64],
 * W2[64][64], b2[64], W3[64], b3 (6,337 floats); W1 row 31 multiplies the zero pad. */
typedef struct beams_mlp beams_mlp;
#define BEAMS_MLP_FEATURES 32

int beams_mlp_num_params(void);
/* He-uniform weights from SplitMix64(seed), zero biases, zero optimiser state. */
int beams_mlp_create(beams_mlp** out, uint64_t seed);
int beams_mlp_destroy(beams_mlp* h);
const char* beams_mlp_last_error(const beams_mlp* h);
int beams_mlp_get_params(beams_mlp* h, float* out);
/* Replaces the parameters (finite values only) and resets the optimiser state. */
int beams_mlp_set_params(beams_mlp* h, const float* in);
/* y[n] = network(x[n][32]). */
int beams_mlp_forward(beams_mlp* h, const float* x, int64_t n, float* y, int32_t mem);
/* Gradients (6,337 floats, host memory) of the mean squared error over n rows with respect to
 * the BF16-rounded parameters the forward pass uses; *loss (may be NULL) = that error. */
int beams_mlp_gradients(beams_mlp* h, const float* x, const float* t, int64_t n, float* grads,
                        float* loss, int32_t mem);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over consecutive minibatches of `batch` rows for `epochs`
 * passes; *loss (may be NULL) = mean squared error over the last pass. */
int beams_mlp_train(beams_mlp* h, const float* x, const float* t, int64_t n, int32_t epochs,
                    int32_t batch, float lr, float* loss, int32_t mem);

/* Attach the network used by BEAMS_SHADOW_LEARNED (net NULL: the coarse optical depth alone).
 * fallback_mode (BEAMS_SHADOW_REF_MARCH or BEAMS_SHADOW_SWEEP) builds the map instead when the
 * largest opacity error at the check vertices (every check_stride-th map vertex per axis and
 * the last) exceeds error_budget; a negative budget sends every map to the fallback (the
 * learned frame's exact counterpart), a NaN one is refused.  With the reference-march fallback
 * the map is rebuilt tile by tile (below); the sweep fallback rebuilds the whole map.  The decision is taken on the
 * device (the fallback's launches return at once when the check is within the budget), so a
 * learned frame is captured in one graph like the exact ones.  The learned map approximates
 * the single-block map (the whole volume's light march, step |size| / 128).  With bricks that
 * is the single-block map at each brick's vertices (no non-local exchange and no per-brick
 * light step, unlike the exact brick frame).  The check vertices cut each map (the single
 * block's, or each brick's) into tiles of up to check_stride^3 cells, and a tile is re-marched
 * when the error at its own or a neighbouring tile's corners exceeds the budget.  The renderer
 * keeps the pointer: destroy the renderer or attach another network before destroying net. */
int beams_set_learned(beams_renderer* h, const beams_mlp* net, int32_t fallback_mode,
                      float error_budget, int32_t check_stride);
/* Training data from the current field, colormap and light: builds the exact map with
 * `mode` (BEAMS_SHADOW_REF_MARCH or BEAMS_SHADOW_SWEEP) and writes the features
 * (n x 32) and targets (n) of n map vertices drawn uniformly by SplitMix64(seed). */
int beams_learned_samples(beams_renderer* h, int32_t mode, int64_t n, uint64_t seed,
                          float* features, float* targets, int32_t mem);
/* The last learned map build (read back from the device: waits for the stream): err[0] =
 * largest |learned - exact| opacity at the check vertices, err[1] = the same for the coarse
 * optical depth alone (single block only, else -1); *fell_back = what the fallback
 * rebuilt: the number of check tiles re-marched by the reference march (bricks, and the single
 * block with that fallback), or 0 or 1 map (the single block with the sweep fallback).  -1, -1 and 0 before any. */
int beams_get_learned_stats(beams_renderer* h, float err[2], int32_t* fell_back);

#ifdef __cplusplus
}
#endif
#endif
