/*
 * Test oracle body (see oracle.h) — included twice by oracle.c with
 *   R = float,  SUF = f32,  FMA_ = fmaf, SQRT_ = sqrtf, FMIN_ = fminf, FMAX_ = fmaxf
 *   R = double, SUF = f64,  FMA_ = fma,  SQRT_ = sqrt,  FMIN_ = fmin,  FMAX_ = fmax
 * Compiled with -ffp-contract=off so every a*b+c rounds twice, as the
 * reference's CPU build (x86-64 baseline, no FMA) does.
 *
 * Reference files restated here (paths relative to the pilot repository):
 *   beams/Intersections.cxx:8-49            SegmentAABB
 *   beams/Math.h:12-20                      BoundsMagnitude
 *   beams/rendering/RectilinearMeshOracle.h:87-195   IsInside/FindCell/Interpolate
 *   beams/rendering/LightRayOperations.h:17-65       light rays
 *   beams/rendering/TransmittanceMap.h:33-237        TransmittanceLocator/Estimator
 *   beams/rendering/TransmittanceMap.h:239-340       TransmittanceMapGenerator
 *   beams/rendering/TransmittanceMap.h:342-453,487-513,550-649  non-local hits, fetch
 *   beams/rendering/LightedVolumeRenderer.cxx:45-186   RectilinearLocator
 *   beams/rendering/LightedVolumeRenderer.cxx:294-499  Sampler
 *   beams/rendering/LightedVolumeRenderer.cxx:501-642  SamplerCellAssoc
 *   beams/rendering/LightedVolumeRenderer.cxx:644-718  CalcRayStart
 *   beams/rendering/LightedVolumeRenderer.cxx:807-825,910-1372  Render/RenderOnDevice
 *   beams/rendering/MapperLightedVolume.cxx:215-303    GlobalComposite
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

typedef struct {
  int64_t n[3];      /* point dims */
  int64_t c[3];      /* cell dims */
  const R* x[3];     /* coordinate arrays */
  R minp[3], maxp[3];
  double bounds[6];  /* vtkm::Bounds of the coordinates (FP64) */
  R ext_mag;         /* BoundsMagnitude<Float32> */
  /* transmittance map geometry (CreateDataSetForOpacityMap) */
  int64_t mp[3];     /* map point dims = cells + 1 */
  R morg[3], msp[3], minv[3], mmax[3];
  R mstep;           /* |size| / 128 */
  R max_density;     /* GetMaxAlpha */
  R min_scalar, inv_delta;
  R light[3], light_color[3];
  R dir_far;         /* directional-light origin distance */
  int64_t ntf;
  const R* tf;
} FN(Geo);

/* vtkm::Min/Max on floating point are fmin/fmax on the CPU build. */
static inline R FN(vmin)(R a, R b) { return FMIN_(a, b); }
static inline R FN(vmax)(R a, R b) { return FMAX_(a, b); }
/* lcl's internal::lerp used by vtkm::exec::CellInterpolate (VTK-m 1.9 thirdparty lcl):
 * fma(t, b, fma(-t, a, a)). */
static inline R FN(lerp)(R a, R b, R t) { return FMA_(t, b, FMA_(-t, a, a)); }

static inline R FN(dot3)(const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void FN(geo_init)(FN(Geo) * g, const OrcScene* s)
{
  for (int d = 0; d < 3; ++d) {
    g->n[d] = s->dims[d];
    g->c[d] = s->dims[d] - 1;
    g->x[d] = (const R*)s->coords[d];
    double lo = (double)g->x[d][0], hi = lo;
    for (int64_t i = 1; i < g->n[d]; ++i) {
      double v = (double)g->x[d][i];
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
    g->bounds[2 * d] = lo;
    g->bounds[2 * d + 1] = hi;
    g->minp[d] = g->x[d][0];
    g->maxp[d] = g->x[d][g->n[d] - 1];
  }
  /* Math.h:12-20 */
  R e[3] = { (R)(g->bounds[1] - g->bounds[0]), (R)(g->bounds[3] - g->bounds[2]),
             (R)(g->bounds[5] - g->bounds[4]) };
  g->ext_mag = SQRT_(FN(dot3)(e, e));
  /* TransmittanceMap.h:550-563, 582-585; LightedVolumeRenderer.cxx:926-931 */
  for (int d = 0; d < 3; ++d) {
    g->mp[d] = s->map_cells[d] + 1;
    g->morg[d] = (R)g->bounds[2 * d];
    R size = (R)(g->bounds[2 * d + 1] - g->bounds[2 * d]);
    g->msp[d] = size / (R)s->map_cells[d];
    g->minv[d] = (R)1 / g->msp[d];
    g->mmax[d] = g->morg[d] + g->msp[d] * (R)(g->mp[d] - 1);
    e[d] = size;
  }
  g->mstep = SQRT_(FN(dot3)(e, e)) / (R)128;
  /* TransmittanceMap.h:515-525 */
  const R* tf = (const R*)s->tf;
  R ma = 0;
  for (int64_t i = 0; i < s->ntf; ++i) ma = FN(vmax)(ma, tf[4 * i + 3]);
  g->max_density = ma;
  g->ntf = s->ntf;
  g->tf = tf;
  /* TransmittanceMap.h:254-261 and LightedVolumeRenderer.cxx:331-334 (same rule) */
  R smin = (R)s->scalar_range[0], smax = (R)s->scalar_range[1];
  g->min_scalar = smin;
  g->inv_delta = ((smax - smin) != (R)0) ? (R)1 / (smax - smin) : smin;
  for (int d = 0; d < 3; ++d) {
    g->light[d] = (R)s->light_pos[d];
    g->light_color[d] = (R)s->light_color[d];
  }
  if (s->light_kind == 1) { /* directional extension: unit travel direction */
    R rs = (R)1 / SQRT_(FN(dot3)(g->light, g->light));
    for (int d = 0; d < 3; ++d) g->light[d] = g->light[d] * rs;
  }
  g->dir_far = s->light_far > 0 ? (R)s->light_far : (R)4 * g->ext_mag;
}

/* ---- Intersections.cxx:8-49 ------------------------------------------------------ */
static int FN(segment_aabb)(const R* start, const R* end, const double* aabb, R tEps, R* tMinOut,
                            R* tMaxOut)
{
  R d[3] = { end[0] - start[0], end[1] - start[1], end[2] - start[2] };
  R len = SQRT_(FN(dot3)(d, d));
  d[0] = d[0] / len;
  d[1] = d[1] / len;
  d[2] = d[2] / len;
  R tMin = 0, tMax = (R)INFINITY;
  for (int a = 0; a < 3; ++a) {
    /* vtkm::Bounds are Float64: the subtraction and division promote. */
    R t0 = (R)((aabb[2 * a] - (double)start[a]) / (double)d[a]);
    R t1 = (R)((aabb[2 * a + 1] - (double)start[a]) / (double)d[a]);
    tMin = FN(vmax)(tMin, FN(vmin)(t0, t1));
    tMax = FN(vmin)(tMax, FN(vmax)(t0, t1));
  }
  if (tMin < (R)0) tMin = tEps;
  if (tMax > len) tMax = len;
  *tMinOut = tMin;
  *tMaxOut = tMax;
  return tMin <= tMax;
}

int FN(orc_segment_aabb)(const void* start, const void* end, const double* aabb, double t_eps,
                         void* t_min, void* t_max)
{
  return FN(segment_aabb)((const R*)start, (const R*)end, aabb, (R)t_eps, (R*)t_min, (R*)t_max);
}

/* ---- RectilinearMeshOracle.h:87-195 ------------------------------------------------ */
static inline int FN(inside)(const FN(Geo) * g, const R* p)
{
  int in = 1;
  for (int d = 0; d < 3; ++d)
    if (p[d] < g->minp[d] || p[d] > g->maxp[d]) in = 0;
  return in;
}

static int64_t FN(find_cell)(const FN(Geo) * g, const R* p, R* pc)
{
  if (!FN(inside)(g, p)) return -1;
  int64_t cell[3];
  for (int d = 0; d < 3; ++d) {
    int64_t lo = 0, hi = g->c[d] - 1;
    const R v = p[d];
    while (lo <= hi) {
      int64_t mid = (lo + hi) / 2;
      R pm = g->x[d][mid + 1];
      if (v <= pm)
        hi = mid - 1;
      else
        lo = mid + 1;
    }
    cell[d] = lo;
  }
  for (int d = 0; d < 3; ++d) {
    R lo = g->x[d][cell[d]];
    int64_t j = cell[d] + 1;
    j = j >= g->n[d] ? cell[d] : j;
    R hi = g->x[d][j];
    R sp = hi - lo; /* VecAxisAlignedPointCoordinates(min, max - min) */
    pc[d] = (p[d] - lo) / sp;
  }
  return (cell[2] * g->c[1] + cell[1]) * g->c[0] + cell[0];
}

static inline void FN(cell_points)(const int64_t* pdims, int64_t cellId, const int64_t* cdims,
                                   int64_t* idx)
{
  int64_t i = cellId % cdims[0];
  int64_t j = (cellId / cdims[0]) % cdims[1];
  int64_t k = cellId / (cdims[0] * cdims[1]);
  idx[0] = (k * pdims[1] + j) * pdims[0] + i;
  idx[1] = idx[0] + 1;
  idx[2] = idx[1] + pdims[0];
  idx[3] = idx[2] - 1;
  idx[4] = idx[0] + pdims[0] * pdims[1];
  idx[5] = idx[4] + 1;
  idx[6] = idx[5] + pdims[0];
  idx[7] = idx[6] - 1;
}

/* lcl Hexahedron interpolate: x on (0,1),(3,2),(4,5),(7,6); then y; then z. */
static inline R FN(interp_hex)(const R* v, const R* pc)
{
  R a = FN(lerp)(v[0], v[1], pc[0]);
  R b = FN(lerp)(v[3], v[2], pc[0]);
  R c = FN(lerp)(v[4], v[5], pc[0]);
  R d = FN(lerp)(v[7], v[6], pc[0]);
  R lo = FN(lerp)(a, b, pc[1]);
  R hi = FN(lerp)(c, d, pc[1]);
  return FN(lerp)(lo, hi, pc[2]);
}

int64_t FN(orc_find_cell)(const OrcScene* s, const void* point, void* pcoords)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  return FN(find_cell)(&g, (const R*)point, (R*)pcoords);
}

double FN(orc_interp_hex)(const void* values8, const void* pcoords)
{
  return (double)FN(interp_hex)((const R*)values8, (const R*)pcoords);
}

/* Map vertex coordinate: ArrayPortalUniformPointCoordinates::Get (origin + spacing * i). */
static inline void FN(map_vertex)(const FN(Geo) * g, int64_t id, R* v)
{
  int64_t ijk[3] = { id % g->mp[0], (id / g->mp[0]) % g->mp[1], id / (g->mp[0] * g->mp[1]) };
  for (int d = 0; d < 3; ++d) v[d] = g->morg[d] + g->msp[d] * (R)ijk[d];
}

/* Light-ray origin for a map vertex: the point light itself
 * (LightRayOperations.h:38-43); for the directional extension, a point
 * 4 * |extent| up-light of the vertex. */
static inline void FN(light_origin)(const FN(Geo) * g, int kind, const R* v, R* o)
{
  if (kind == 1) {
    for (int d = 0; d < 3; ++d) o[d] = v[d] - g->light[d] * g->dir_far;
  } else {
    for (int d = 0; d < 3; ++d) o[d] = g->light[d];
  }
}

/* ---- TransmittanceMap.h:279-331 (the light march to one point v) ------------------------ */
static R FN(march_point)(const FN(Geo) * g, const R* field, int64_t nfield, int kind, const R* v,
                         R opacity)
{
  R o[3], dir[3];
  FN(light_origin)(g, kind, v, o);
  /* vtkm::Normal(dest - origin) = x * (1 / sqrt(|x|^2)) */
  R dd[3] = { v[0] - o[0], v[1] - o[1], v[2] - o[2] };
  R rs = (R)1 / SQRT_(FN(dot3)(dd, dd));
  for (int d = 0; d < 3; ++d) dir[d] = dd[d] * rs;

  R tMin, tMax;
  (void)FN(segment_aabb)(o, v, g->bounds, (R)1e-4f, &tMin, &tMax);
  R start[3], p[3];
  for (int d = 0; d < 3; ++d) {
    start[d] = o[d] + tMin * dir[d];
    p[d] = start[d];
  }
  const R len2 = (tMax - tMin) * (tMax - tMin);
  const int64_t tfmax = g->ntf - 1;
  for (int64_t it = 0; it < 1024; ++it) { /* no in-bounds segment takes > 130 steps */
    R df[3] = { p[0] - start[0], p[1] - start[1], p[2] - start[2] };
    if (!(FN(dot3)(df, df) <= len2)) break;
    for (int d = 0; d < 3; ++d) p[d] = p[d] + g->mstep * dir[d];
    R pc[3];
    int64_t cellId = FN(find_cell)(g, p, pc);
    if (cellId == -1) break;
    int64_t idx[8];
    FN(cell_points)(g->n, cellId, g->c, idx);
    R vals[8];
    for (int i = 0; i < 8; ++i) {
      int64_t j = idx[i] >= nfield ? nfield - 1 : idx[i];
      vals[i] = field[j];
    }
    R sc = FN(interp_hex)(vals, pc);
    sc = (sc - g->min_scalar) * g->inv_delta;
    int64_t ci = (int64_t)(sc * (R)tfmax);
    ci = ci < tfmax ? ci : tfmax;
    ci = ci > 0 ? ci : 0;
    R a = g->tf[4 * ci + 3] / g->max_density;
    opacity = opacity + ((R)1 - opacity) * a;
  }
  return opacity;
}

static R FN(march_vertex)(const FN(Geo) * g, const R* field, int64_t nfield, int kind,
                          int64_t id, R opacity)
{
  R v[3];
  FN(map_vertex)(g, id, v);
  return FN(march_point)(g, field, nfield, kind, v, opacity);
}

static int64_t FN(nfield)(const OrcScene* s)
{
  if (s->assoc == 1) return (s->dims[0] - 1) * (s->dims[1] - 1) * (s->dims[2] - 1);
  return s->dims[0] * s->dims[1] * s->dims[2];
}

/* The positions of block s's map vertices (x fastest) into the x, y, z arrays pts[0..2]. */
int FN(orc_map_points)(const OrcScene* s, void* const* pts)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  const int64_t nv = g.mp[0] * g.mp[1] * g.mp[2];
  for (int64_t id = 0; id < nv; ++id) {
    R v[3];
    FN(map_vertex)(&g, id, v);
    for (int d = 0; d < 3; ++d) ((R*)pts[d])[id] = v[d];
  }
  return 0;
}

/* The light march of block s (its bounds, its light step |size| / 128) to n given points
 * (x, y, z arrays): the map value a vertex at each point would get. */
int FN(orc_march_points)(const OrcScene* s, const void* const* pts, int64_t n, void* out, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  const R* field = (const R*)s->field;
  const int64_t nf = FN(nfield)(s);
  const R* px = (const R*)pts[0];
  const R* py = (const R*)pts[1];
  const R* pz = (const R*)pts[2];
  R* o = (R*)out;
  (void)nthr;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t i = 0; i < n; ++i) {
    const R v[3] = { px[i], py[i], pz[i] };
    o[i] = FN(march_point)(&g, field, nf, s->light_kind, v, (R)0);
  }
  return 0;
}

int FN(orc_tmap_build)(const OrcScene* s, const void* seed, void* map_out, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  const int64_t nv = g.mp[0] * g.mp[1] * g.mp[2];
  const R* sd = (const R*)seed;
  R* out = (R*)map_out;
  const R* field = (const R*)s->field;
  const int64_t nf = FN(nfield)(s);
  (void)nthr;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t id = 0; id < nv; ++id)
    out[id] = FN(march_vertex)(&g, field, nf, s->light_kind, id, sd ? sd[id] : (R)0);
  return 0;
}

/* The map vertices [begin, end) only (a slab of the map, for timing samples); seed, when
 * given, is a full-size array of phase-3 seed opacities. */
int FN(orc_tmap_build_range)(const OrcScene* s, const void* seed, void* map_out, int64_t begin,
                             int64_t end, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  R* out = (R*)map_out;
  const R* sd = (const R*)seed;
  const R* field = (const R*)s->field;
  const int64_t nf = FN(nfield)(s);
  (void)nthr;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t id = begin; id < end; ++id)
    out[id] = FN(march_vertex)(&g, field, nf, s->light_kind, id, sd ? sd[id] : (R)0);
  return 0;
}

/* ---- TransmittanceMap.h:93-112,166-229 ------------------------------------------------ */
static R FN(map_opacity)(const FN(Geo) * g, const R* map, const R* p)
{
  R t[3];
  int64_t cell[3];
  for (int d = 0; d < 3; ++d) {
    t[d] = (p[d] - g->morg[d]) * g->minv[d];
    if (t[d] < (R)0) t[d] = (R)0;
    if (t[d] >= (R)(g->mp[d] - 1)) t[d] = (R)(g->mp[d] - 2);
    cell[d] = (int64_t)t[d];
  }
  int64_t idx[8];
  idx[0] = (cell[2] * g->mp[1] + cell[1]) * g->mp[0] + cell[0];
  idx[1] = idx[0] + 1;
  idx[2] = idx[1] + g->mp[0];
  idx[3] = idx[2] - 1;
  idx[4] = idx[0] + g->mp[0] * g->mp[1];
  idx[5] = idx[4] + 1;
  idx[6] = idx[5] + g->mp[0];
  idx[7] = idx[6] - 1;
  R pc[3];
  for (int d = 0; d < 3; ++d) {
    R lo = g->morg[d] + g->msp[d] * (R)cell[d];
    R hi = g->morg[d] + g->msp[d] * (R)(cell[d] + 1);
    pc[d] = (p[d] - lo) / (hi - lo);
  }
  R vals[8];
  for (int i = 0; i < 8; ++i) vals[i] = map[idx[i]];
  return FN(interp_hex)(vals, pc);
}

int FN(orc_map_estimate)(const OrcScene* s, const void* map, const void* point, void* rgb_out,
                         void* opacity_out)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  R o = FN(map_opacity)(&g, (const R*)map, (const R*)point);
  R tr = (R)1 - o;
  if (rgb_out)
    for (int d = 0; d < 3; ++d) ((R*)rgb_out)[d] = tr * g.light_color[d];
  if (opacity_out) *(R*)opacity_out = o;
  return 0;
}

/* ---- LightedVolumeRenderer.cxx:644-718 ------------------------------------------------ */
static inline R FN(rcp_safe)(R f) { return (R)1 / ((FABS_(f) < (R)1e-8f) ? (R)1e-8f : f); }

static void FN(calc_ray_start)(const FN(Geo) * g, const R* org, const R* dir, R* minD, R* dist,
                               R* maxD)
{
  R xmin_b = (R)g->bounds[0], xmax_b = (R)g->bounds[1];
  R ymin_b = (R)g->bounds[2], ymax_b = (R)g->bounds[3];
  R zmin_b = (R)g->bounds[4], zmax_b = (R)g->bounds[5];
  R ix = FN(rcp_safe)(dir[0]), iy = FN(rcp_safe)(dir[1]), iz = FN(rcp_safe)(dir[2]);
  R ox = org[0] * ix, oy = org[1] * iy, oz = org[2] * iz;
  R xmin = xmin_b * ix - ox, ymin = ymin_b * iy - oy, zmin = zmin_b * iz - oz;
  R xmax = xmax_b * ix - ox, ymax = ymax_b * iy - oy, zmax = zmax_b * iz - oz;
  *minD = FN(vmax)(FN(vmax)(FN(vmax)(FN(vmin)(ymin, ymax), FN(vmin)(xmin, xmax)), FN(vmin)(zmin, zmax)),
                   *minD);
  R ex = FN(vmin)(FN(vmin)(FN(vmax)(ymin, ymax), FN(vmax)(xmin, xmax)), FN(vmax)(zmin, zmax));
  *maxD = FN(vmin)(*maxD, ex);
  if (*maxD < *minD)
    *minD = (R)-1;
  else
    *dist = *minD;
}

int FN(orc_calc_ray_start)(const OrcScene* s, OrcRays* rays, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  R* mn = (R*)rays->min_distance;
  R* mx = (R*)rays->max_distance;
  R* ds = (R*)rays->distance;
  (void)nthr;
#pragma omp parallel for schedule(static) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t r = 0; r < rays->n; ++r) {
    R o[3], d[3];
    for (int c = 0; c < 3; ++c) {
      o[c] = ((const R*)rays->origin[c])[r];
      d[c] = ((const R*)rays->dir[c])[r];
    }
    FN(calc_ray_start)(&g, o, d, &mn[r], &ds[r], &mx[r]);
  }
  return 0;
}

/* ---- LightedVolumeRenderer.cxx:119-165 (RectilinearLocator::LocateCell) --------------- */
static inline void FN(locate_cell)(const FN(Geo) * g, int64_t* cell, const R* p, R* invsp)
{
  for (int d = 0; d < 3; ++d) {
    if (p[d] == g->maxp[d]) {
      cell[d] = g->n[d] - 2;
      continue; /* invSpacing[d] keeps its previous value */
    }
    const R* x = g->x[d];
    R lo = x[cell[d]];
    const int64_t dirn = (p[d] - lo >= (R)0) ? 1 : -1;
    R hi = x[cell[d] + 1];
    for (;;) {
      if (p[d] >= lo && p[d] < hi) break;
      cell[d] += dirn;
      int64_t nxt = dirn == 1 ? cell[d] + 1 : cell[d];
      R nv = x[nxt];
      if (dirn == 1) {
        lo = hi;
        hi = nv;
      } else {
        hi = lo;
        lo = nv;
      }
    }
    invsp[d] = (R)1 / (hi - lo);
  }
}

/* ---- LightedVolumeRenderer.cxx:345-498 (Sampler, point field) -------------------------- */
static int64_t FN(sample_point)(const FN(Geo) * g, const R* field, const R* map, int use_map,
                             R sampleDist, R meshEps, R alphaStop, const R* org, const R* dir,
                             R minD, R maxD, R* color)
{
  if (minD == (R)-1) return 0;
  int64_t composited = 0;
  R dist = minD + meshEps;
  R p[3];
  /* step bound (never reached by a finite ray): 2 diagonals / step + 64 */
  const int64_t maxSteps = (int64_t)((R)2 * g->ext_mag / sampleDist) + 64;
  int64_t steps = 0;
  for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  while (!FN(inside)(g, p) && dist < maxD && ++steps < maxSteps) {
    dist += sampleDist;
    for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  }
  R bl[3] = { 0, 0, 0 };
  int newCell = 1;
  R tx = 0, ty = 0, tz = 0;
  R s0 = 0, s10 = 0, s23 = 0, s3 = 0, s4 = 0, s54 = 0, s67 = 0, s7 = 0;
  int64_t cell[3] = { 0, 0, 0 };
  R invsp[3] = { 0, 0, 0 };
  const int64_t cms = g->ntf - 1;
  const int64_t nx = g->n[0], nxy = g->n[0] * g->n[1];
  while (FN(inside)(g, p) && dist < maxD && ++steps < maxSteps) {
    R mint = FN(vmin)(tx, FN(vmin)(ty, tz));
    R maxt = FN(vmax)(tx, FN(vmax)(ty, tz));
    if (maxt > (R)1 || mint < (R)0) newCell = 1;
    if (newCell) {
      FN(locate_cell)(g, cell, p, invsp);
      int64_t i0 = (cell[2] * g->n[1] + cell[1]) * nx + cell[0];
      int64_t i1 = i0 + 1, i2 = i1 + nx, i3 = i2 - 1;
      int64_t i4 = i0 + nxy, i5 = i4 + 1, i6 = i5 + nx, i7 = i6 - 1;
      bl[0] = g->x[0][cell[0]];
      bl[1] = g->x[1][cell[1]];
      bl[2] = g->x[2][cell[2]];
      s0 = field[i0];
      R s1 = field[i1], s2 = field[i2];
      s3 = field[i3];
      s4 = field[i4];
      R s5 = field[i5], s6 = field[i6];
      s7 = field[i7];
      s67 = s6 - s7;
      s54 = s5 - s4;
      s10 = s1 - s0;
      s23 = s2 - s3;
      tx = (p[0] - bl[0]) * invsp[0];
      ty = (p[1] - bl[1]) * invsp[1];
      tz = (p[2] - bl[2]) * invsp[2];
      newCell = 0;
    }
    R l76 = s7 + tx * s67;
    R l45 = s4 + tx * s54;
    R top = l45 + ty * (l76 - l45);
    R l01 = s0 + tx * s10;
    R l32 = s3 + tx * s23;
    R bot = l01 + ty * (l32 - l01);
    R fs = bot + tz * (top - bot);
    fs = (fs - g->min_scalar) * g->inv_delta;
    int64_t ci = (int64_t)(fs * (R)cms);
    /* The reference `continue`s here without advancing (an endless loop for
     * scalars outside the range); this restatement advances instead. */
    int skip = (ci < 0 || ci > cms);
    if (!skip) {
      R sc[4] = { g->tf[4 * ci + 0], g->tf[4 * ci + 1], g->tf[4 * ci + 2], g->tf[4 * ci + 3] };
      sc[3] *= ((R)1 - color[3]);
      if (use_map) {
        R tr = (R)1 - FN(map_opacity)(g, map, p);
        sc[0] = sc[0] * (tr * g->light_color[0]);
        sc[1] = sc[1] * (tr * g->light_color[1]);
        sc[2] = sc[2] * (tr * g->light_color[2]);
      }
      for (int c = 0; c < 3; ++c) sc[c] = sc[c] < (R)0 ? (R)0 : (sc[c] > (R)1 ? (R)1 : sc[c]);
      color[0] = color[0] + sc[0] * sc[3];
      color[1] = color[1] + sc[1] * sc[3];
      color[2] = color[2] + sc[2] * sc[3];
      color[3] = sc[3] + color[3];
      ++composited;
    }
    dist += sampleDist;
    for (int d = 0; d < 3; ++d) p[d] = p[d] + sampleDist * dir[d];
    tx = (p[0] - bl[0]) * invsp[0];
    ty = (p[1] - bl[1]) * invsp[1];
    tz = (p[2] - bl[2]) * invsp[2];
    if (color[3] >= alphaStop) break;
  }
  return composited;
}

/* ---- LightedVolumeRenderer.cxx:540-641 (SamplerCellAssoc) ------------------------------ */
static int64_t FN(sample_cell)(const FN(Geo) * g, const R* field, R sampleDist, R meshEps,
                            R alphaStop, const R* org, const R* dir, R minD, R maxD, R* color)
{
  if (minD == (R)-1) return 0;
  int64_t composited = 0;
  R dist = minD + meshEps;
  R p[3];
  /* step bound (never reached by a finite ray): 2 diagonals / step + 64 */
  const int64_t maxSteps = (int64_t)((R)2 * g->ext_mag / sampleDist) + 64;
  int64_t steps = 0;
  for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  while (!FN(inside)(g, p) && dist < maxD && ++steps < maxSteps) {
    dist += sampleDist;
    for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  }
  int newCell = 1;
  R tx = 2, ty = 2, tz = 2;
  R scol[4] = { 0, 0, 0, 0 };
  R bl[3] = { 0, 0, 0 };
  R invsp[3] = { 0, 0, 0 };
  int64_t cell[3] = { 0, 0, 0 };
  const int64_t cms = g->ntf - 1;
  while (FN(inside)(g, p) && dist < maxD && ++steps < maxSteps) {
    R mint = FN(vmin)(tx, FN(vmin)(ty, tz));
    R maxt = FN(vmax)(tx, FN(vmax)(ty, tz));
    if (maxt > (R)1 || mint < (R)0) newCell = 1;
    if (newCell) {
      FN(locate_cell)(g, cell, p, invsp);
      int64_t cid = (cell[2] * g->c[1] + cell[1]) * g->c[0] + cell[0];
      R s0 = field[cid];
      R ns = (s0 - g->min_scalar) * g->inv_delta;
      int64_t ci = (int64_t)(ns * (R)cms);
      if (ci < 0) ci = 0;
      if (ci > cms) ci = cms;
      for (int c = 0; c < 4; ++c) scol[c] = g->tf[4 * ci + c];
      bl[0] = g->x[0][cell[0]];
      bl[1] = g->x[1][cell[1]];
      bl[2] = g->x[2][cell[2]];
      tx = (p[0] - bl[0]) * invsp[0];
      ty = (p[1] - bl[1]) * invsp[1];
      tz = (p[2] - bl[2]) * invsp[2];
      newCell = 0;
    }
    R a = scol[3] * ((R)1 - color[3]);
    color[0] = color[0] + scol[0] * a;
    color[1] = color[1] + scol[1] * a;
    color[2] = color[2] + scol[2] * a;
    color[3] = a + color[3];
    ++composited;
    dist += sampleDist;
    for (int d = 0; d < 3; ++d) p[d] = p[d] + sampleDist * dir[d];
    if (color[3] >= alphaStop) break;
    tx = (p[0] - bl[0]) * invsp[0];
    ty = (p[1] - bl[1]) * invsp[1];
    tz = (p[2] - bl[2]) * invsp[2];
  }
  for (int c = 0; c < 4; ++c) color[c] = FN(vmin)(color[c], (R)1);
  return composited;
}

static R FN(sample_distance)(const FN(Geo) * g, const OrcScene* s)
{
  /* LightedVolumeRenderer.cxx:815-819 */
  if (s->sample_distance <= 0) return g->ext_mag / (R)200;
  return (R)s->sample_distance;
}

int FN(orc_sample_rays)(const OrcScene* s, const void* map, OrcRays* rays, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, s);
  const R sd = FN(sample_distance)(&g, s);
  const R eps = g.ext_mag * (R)0.0001f; /* LightedVolumeRenderer.cxx:1261 */
  const R stop = (R)s->alpha_stop;
  const R* field = (const R*)s->field;
  R* rgba = (R*)rays->rgba;
  const R* mn = (const R*)rays->min_distance;
  const R* mx = (const R*)rays->max_distance;
  const int use_map = s->use_shadow && map != 0;
  (void)nthr;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t r = 0; r < rays->n; ++r) {
    R o[3], d[3];
    for (int c = 0; c < 3; ++c) {
      o[c] = ((const R*)rays->origin[c])[r];
      d[c] = ((const R*)rays->dir[c])[r];
    }
    R* col = rgba + 4 * r;
    int64_t ns;
    if (s->assoc == 1)
      ns = FN(sample_cell)(&g, field, sd, eps, stop, o, d, mn[r], mx[r], col);
    else
      ns = FN(sample_point)(&g, field, (const R*)map, use_map, sd, eps, stop, o, d, mn[r], mx[r], col);
    if (rays->nsamples) ((int64_t*)rays->nsamples)[r] = ns;
  }
  return 0;
}

/* Single-block RenderOnDevice: map (phases 1 and 3 coincide for one block), then
 * CalcRayStart and the sampler. */
int FN(orc_render_rays)(const OrcScene* s, void* map_out, OrcRays* rays, int nthr)
{
  if (s->use_shadow && s->assoc == 0) {
    if (!map_out) return -1;
    FN(orc_tmap_build)(s, 0, map_out, nthr);
  }
  FN(orc_calc_ray_start)(s, rays, nthr);
  return FN(orc_sample_rays)(s, s->use_shadow ? map_out : 0, rays, nthr);
}

/* ---- TransmittanceMap.h:342-453, 623-649: non-local hits of one block ----------------- */
typedef struct {
  int32_t ray_id, block_id, from_block_id;
#if ORC_IS_F64
  int32_t pad;
#endif
  R point[3];
  R ray_t;
  R opacity;
} FN(Hit);

/* One pass over the map vertices [begin, end) of block `self`: the number of other blocks each
 * light segment crosses (CountNonLocalBlockHits, into counts), or, with hits given, the hit
 * records at offsets (CalculateNonLocalBlockHits); counts and offsets are indexed from begin.
 * Vertices are independent: nthr threads share them. */
static void FN(hits_pass)(const FN(Geo) * g, const double* bb, int nblocks, int self, int kind,
                          int64_t* counts, const int64_t* offsets, FN(Hit) * hits, int64_t begin,
                          int64_t end, int nthr)
{
  (void)nthr;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t id = begin; id < end; ++id) {
    R v[3], o[3];
    FN(map_vertex)(g, id, v);
    FN(light_origin)(g, kind, v, o);
    if (!hits) {
      int64_t n = 0;
      for (int b = 0; b < nblocks; ++b) {
        if (b == self) continue;
        R t0, t1;
        if (FN(segment_aabb)(o, v, bb + 6 * b, (R)1e-4f, &t0, &t1)) ++n; /* glancing hits kept */
      }
      counts[id - begin] = n;
      continue;
    }
    R dd[3] = { v[0] - o[0], v[1] - o[1], v[2] - o[2] };
    R rs = (R)1 / SQRT_(FN(dot3)(dd, dd)); /* vtkm::Normalize */
    R dir[3] = { dd[0] * rs, dd[1] * rs, dd[2] * rs };
    int64_t off = offsets ? offsets[id - begin] : 0;
    for (int b = 0; b < nblocks; ++b) {
      if (b == self) continue;
      R t0, t1;
      /* the reference passes SelfBlockId as tEps here; tMin is never < 0, so it is inert */
      if (!FN(segment_aabb)(o, v, bb + 6 * b, (R)self, &t0, &t1)) continue;
      FN(Hit)* h = &hits[off++];
      h->ray_id = (int32_t)id;
      h->block_id = b;
      h->from_block_id = self;
      h->ray_t = t1;
      for (int d = 0; d < 3; ++d) h->point[d] = o[d] + dir[d] * t1;
      h->opacity = 0;
    }
  }
}

/* CountNonLocalBlockHits + ScanExclusive + CalculateNonLocalBlockHits for block `self`
 * (TransmittanceMap.h:342-453, 623-649); returns the number of hits. */
static int64_t FN(nonlocal_hits)(int nblocks, const OrcScene* scenes, int self, int64_t* counts,
                                 int64_t* offsets, void* hits_out, int64_t hits_cap, int nthr)
{
  FN(Geo) g, gb;
  FN(geo_init)(&g, &scenes[self]);
  const int64_t nv = g.mp[0] * g.mp[1] * g.mp[2];
  double* bb = (double*)malloc(sizeof(double) * 6 * (size_t)nblocks);
  for (int b = 0; b < nblocks; ++b) {
    FN(geo_init)(&gb, &scenes[b]);
    memcpy(bb + 6 * b, gb.bounds, sizeof(double) * 6);
  }
  const int kind = scenes[self].light_kind;
  int64_t* cnt = counts ? counts : (int64_t*)malloc(sizeof(int64_t) * (size_t)nv);
  FN(hits_pass)(&g, bb, nblocks, self, kind, cnt, 0, 0, 0, nv, nthr);
  int64_t total = 0;
  for (int64_t id = 0; id < nv; ++id) {
    if (offsets) offsets[id] = total;
    total += cnt[id];
  }
  if (hits_out && hits_cap >= total)
    FN(hits_pass)(&g, bb, nblocks, self, kind, cnt, offsets, (FN(Hit)*)hits_out, 0, nv, nthr);
  if (!counts) free(cnt);
  free(bb);
  return total;
}

int64_t FN(orc_nonlocal_hits)(int nblocks, const OrcScene* scenes, int self, int64_t* counts,
                              int64_t* offsets, void* hits_out, int64_t hits_cap)
{
  return FN(nonlocal_hits)(nblocks, scenes, self, counts, offsets, hits_out, hits_cap, 1);
}

/* Phase 2 fetch (TransmittanceFetcher2, TransmittanceMap.h:505-508) and phase 3
 * composite (LightedVolumeRenderer.cxx:1214-1235) for block `self`: the seed
 * opacities that the re-march starts from.  `maps[b]` is block b's phase-1 map. */
static int FN(hit_less)(const FN(Hit) * a, const FN(Hit) * b)
{
  if (a->ray_id != b->ray_id) return a->ray_id < b->ray_id;
  if (a->ray_t != b->ray_t) return a->ray_t < b->ray_t;
  return a->block_id < b->block_id; /* tie order fixed by block id */
}

/* The seed opacities of the map vertices [begin, end) of block `self` (seed_out is
 * full-size; other entries are left alone). */
int FN(orc_nonlocal_seed_range)(int nblocks, const OrcScene* scenes, const void* const* maps,
                                int self, int64_t begin, int64_t end, void* seed_out, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, &scenes[self]);
  const int64_t nv = g.mp[0] * g.mp[1] * g.mp[2];
  if (begin < 0) begin = 0;
  if (end > nv) end = nv;
  if (end <= begin) return 0;
  const int64_t nr = end - begin;
  /* counts and offsets of the range, indexed from `begin` */
  int64_t* counts = (int64_t*)malloc(sizeof(int64_t) * (size_t)nr);
  int64_t* offs = (int64_t*)malloc(sizeof(int64_t) * (size_t)nr);
  FN(Geo)* gs = (FN(Geo)*)malloc(sizeof(FN(Geo)) * (size_t)nblocks);
  double* bb = (double*)malloc(sizeof(double) * 6 * (size_t)nblocks);
  for (int b = 0; b < nblocks; ++b) {
    FN(geo_init)(&gs[b], &scenes[b]);
    memcpy(bb + 6 * b, gs[b].bounds, sizeof(double) * 6);
  }
  const int kind = scenes[self].light_kind;
  FN(hits_pass)(&g, bb, nblocks, self, kind, counts, 0, 0, begin, end, nthr);
  int64_t total = 0;
  for (int64_t i = 0; i < nr; ++i) {
    offs[i] = total;
    total += counts[i];
  }
  FN(Hit)* hits = (FN(Hit)*)malloc(sizeof(FN(Hit)) * (size_t)(total > 0 ? total : 1));
  FN(hits_pass)(&g, bb, nblocks, self, kind, counts, offs, hits, begin, end, nthr);
  free(bb);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t h = 0; h < total; ++h) {
    int b = hits[h].block_id;
    hits[h].opacity = FN(map_opacity)(&gs[b], (const R*)maps[b], hits[h].point);
  }
  R* seed = (R*)seed_out;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthr > 0 ? nthr : 1)
  for (int64_t id = begin; id < end; ++id) {
    /* insertion sort of this vertex's hits by RayT (a vertex crosses few blocks) */
    FN(Hit)* hs = hits + offs[id - begin];
    int64_t n = counts[id - begin];
    for (int64_t i = 1; i < n; ++i) {
      FN(Hit) t = hs[i];
      int64_t j = i - 1;
      while (j >= 0 && FN(hit_less)(&t, &hs[j])) {
        hs[j + 1] = hs[j];
        --j;
      }
      hs[j + 1] = t;
    }
    R o = 0;
    for (int64_t i = 0; i < n; ++i) o = o + ((R)1 - o) * hs[i].opacity;
    seed[id] = o;
  }
  free(gs);
  free(hits);
  free(offs);
  free(counts);
  return 0;
}

int FN(orc_nonlocal_seed)(int nblocks, const OrcScene* scenes, const void* const* maps, int self,
                          void* seed_out, int nthr)
{
  FN(Geo) g;
  FN(geo_init)(&g, &scenes[self]);
  return FN(orc_nonlocal_seed_range)(nblocks, scenes, maps, self, 0, g.mp[0] * g.mp[1] * g.mp[2],
                                     seed_out, nthr);
}

/* ---- MapperLightedVolume.cxx:215-303 --------------------------------------------------- */
/* GlobalComposite's over operator over an explicit block order (first = nearest). */
int FN(orc_composite_order)(int nblocks, const int* order, const void* const* images, int64_t npix, void* out)
{
  R* o = (R*)out;
  memcpy(o, images[order[0]], sizeof(R) * 4 * (size_t)npix);
  for (int i = 1; i < nblocks; ++i) {
    const R* b = (const R*)images[order[i]];
    for (int64_t j = 0; j < npix; ++j) {
      R* a = o + 4 * j;
      const R* q = b + 4 * j;
      R ia = (R)1 - a[3];
      a[0] = a[0] + q[0] * ia;
      a[1] = a[1] + q[1] * ia;
      a[2] = a[2] + q[2] * ia;
      a[3] = a[3] + q[3] * ia;
    }
  }
  return 0;
}

int FN(orc_composite)(int nblocks, const double* block_bounds, const double* cam_pos,
                      const void* const* images, int64_t npix, void* out)
{
  float* depth = (float*)malloc(sizeof(float) * (size_t)nblocks);
  int* order = (int*)calloc((size_t)(nblocks > 0 ? nblocks : 1), sizeof(int));
  for (int b = 0; b < nblocks; ++b) {
    const double* bb = block_bounds + 6 * b;
    /* vtkm::Bounds::Center in FP64, then Vec3f_32 */
    float c[3] = { (float)((bb[0] + bb[1]) * 0.5), (float)((bb[2] + bb[3]) * 0.5),
                   (float)((bb[4] + bb[5]) * 0.5) };
    float d[3] = { (float)cam_pos[0] - c[0], (float)cam_pos[1] - c[1], (float)cam_pos[2] - c[2] };
    depth[b] = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    order[b] = b;
  }
  /* ascending depth; equal depths keep block order */
  for (int i = 1; i < nblocks; ++i) {
    int t = order[i];
    int j = i - 1;
    while (j >= 0 && depth[order[j]] > depth[t]) {
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = t;
  }
  FN(orc_composite_order)(nblocks, order, images, npix, out);
  free(order);
  free(depth);
  return 0;
}

#undef FN
#undef CAT
#undef CAT2
