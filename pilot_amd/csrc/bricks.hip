// Many bricks on one GPU, standing in for the reference's one block per MPI rank
// (RenderOnDevice phases 1-4, LightedVolumeRenderer.cxx:910-1372, and
// MapperLightedVolume::GlobalComposite, MapperLightedVolume.cxx:215-303):
//   phase 1  k_bricks_map (local maps)              persistent, all bricks
//   phase 2  k_bricks_exchange                       per map vertex: every other brick the
//            light segment crosses (SegmentAABB), that brick's local opacity at the segment's
//            exit, over-composited in exit-distance order -> seed opacity
//   phase 3  k_bricks_finalize (copies, constants,   per map vertex
//            and the list of the rest), then
//            k_bricks_map (seeded re-march)          persistent, the listed vertices
//   phase 4  k_bricks_march (footprint tiles)        one workgroup per (brick, tile) item
//   phase 5  k_bricks_composite (depth-ordered over) per pixel
// The reference routes 32-byte hit records through rank 0 with MPI; here the hits of a
// vertex live in LDS for the duration of its loop and are never written to memory.
#include "bricks.hpp"

#include <algorithm>
#include "workqueue.hpp"
#include "camera.hpp"
#include "march_dev.hpp"

namespace beams {

namespace {

// ---- phases 1 and 3: per-brick TransmittanceMapGenerator -------------------------------------
// Phase 1 marches every map vertex of every brick; phase 3 (seeded) re-marches the vertices
// k_bricks_finalize listed, from their seeds, densely packed into work items.
template <int FDT, int MDT, bool FAST, bool PIPE>
__global__ void __launch_bounds__(256) k_bricks_map(BrickFrame f, int seeded)
{
  extern __shared__ float lds[];
  __shared__ ItemQueue queue;
  const int ntf = f.bricks[0].s.ntf;
  float* a_tab = lds;
  float* xs = lds + ntf;
  unsigned int* counters = f.queue + (seeded ? 2 : 0) * kQueueSlots;
  // the listed vertices' items (k_rm_items wrote their count), or every vertex's
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  const int nitems = seeded ? __builtin_amdgcn_readfirstlane(((ConstInt*)(const void*)f.rm_item_off)[f.nbricks])
                            : f.n_map_items;
  queue_init(&queue);
  int staged = -1;
  // the current brick and its item range: a workgroup's items mostly follow one another in a
  // brick, so the brick is searched for only when an item leaves the range
  ConstInt* offs = (ConstInt*)(const void*)(seeded ? f.rm_item_off : f.map_item_off);
  int b = 0, b_lo = 0, b_hi = 0;
  for (;;) {
    const int item = next_item(counters, &queue, nitems, 1, f.xcd_ranges);
    if (item >= nitems) break;
    if (item < b_lo || item >= b_hi) {
      b = seeded ? item_brick_in(f.rm_item_off, f.nbricks, item) : item_brick(f, item);
      b_lo = offs[b];
      b_hi = offs[b + 1];
    }
    // the brick's parameters in scalar registers (the same brick for the whole workgroup)
    const BrickDesc* bk = f.bricks + b;
    const GridParams g = load_uniform(&bk->g);
    const MapParams m = load_uniform(&bk->m);
    const ShadeParams sp = load_uniform(&bk->s);
    if (staged < 0)
      for (int i = threadIdx.x; i < ntf; i += blockDim.x) a_tab[i] = sp.tf[4 * i + 3] / sp.max_density;
    if (b != staged) stage_coords(g, xs);
    staged = b;
    __syncthreads();
    const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
    // (a brick's map holds fewer than 2^31 vertices: 32-bit divisions)
    const uint32_t mp0 = (uint32_t)m.mp[0], mp1 = (uint32_t)m.mp[1];
    uint32_t i, j, k;
    int64_t v;
    bool live;
    if (seeded) {
      // the item's share of the brick's list
      const int first = (item - b_lo) * kMapItem;
      const int e = first + (int)threadIdx.x;
      live = e < (int)__builtin_amdgcn_readfirstlane(f.rm_count[b]);
      const uint32_t vv = live ? f.rm_list[bk->seed_off + e] : 0u;
      k = vv / (mp0 * mp1);
      const uint32_t rr = vv - k * (mp0 * mp1);
      j = rr / mp0;
      i = rr - j * mp0;
      v = (int64_t)vv;
    } else {
      // the item's ordinal of this lane's vertex: x fastest, or along the brick's rows in
      // light-sheet order (rows whose light paths share a sheet follow one another, so the
      // workgroups in flight read a few sheets of the field instead of a slab of the brick)
      const int64_t q = (int64_t)(item - b_lo) * kMapItem + threadIdx.x;
      live = q < bk->nverts;
      const uint32_t r = (uint32_t)q / mp0;
      i = (uint32_t)q - r * mp0;
      if (f.light_rows) {
        const uint32_t jk = live ? f.rows[bk->row_off + r] : 0u;
        j = jk & 0xffffu;
        k = jk >> 16;
      } else {
        k = r / mp1;
        j = r - k * mp1;
      }
      v = (int64_t)i + (int64_t)mp0 * ((int64_t)j + (int64_t)mp1 * (int64_t)k);
    }
    if (live) {
      // ArrayPortalUniformPointCoordinates::Get, as map_vertex computes it
      const float pos[3] = { m.morg[0] + m.msp[0] * (float)i, m.morg[1] + m.msp[1] * (float)j,
                             m.morg[2] + m.msp[2] * (float)k };
      int64_t taken = 0;
      const float seed = seeded ? f.seeds[bk->seed_off + v] : 0.0f;
      const float op = PIPE ? march_vertex_pipe<FDT, FAST>(g, m, sp, a_tab, x, pos, seed, &taken)
                            : march_vertex_at<FDT, false, FAST>(g, m, sp, a_tab, x, pos, seed, &taken);
      store_t<MDT>(seeded ? bk->map_final : m.map, v, op);
      if (sp.counters) count_add(&sp.counters[1], (uint32_t)taken);
    }
  }
}

// Phase 3's bulk pass (grid (G, nbricks), one contiguous run of vertices per workgroup): a
// vertex seeded at 0 would repeat its phase-1 march, so its local value is copied; one seeded at
// 1 keeps opacity 1 (each step adds (1 - 1) a = 0 for finite a); the others (all of them while
// steps are counted) go to the brick's list, in vertex order within each wave.  A workgroup
// gathers them in LDS and reserves list space once per kFinalBuf - 256 of them (one counter per
// brick takes every reservation, so per-wave reservations would queue on it), writing each
// flush as one contiguous run.
constexpr int kFinalBuf = 2048;

template <int MDT>
__global__ void __launch_bounds__(256) k_bricks_finalize(BrickFrame f)
{
  __shared__ uint32_t buf[kFinalBuf];
  __shared__ unsigned int nbuf, base;
  const int b = blockIdx.y;
  const BrickDesc* bk = f.bricks + b;
  const int64_t nv = bk->nverts;
  const bool counting = bk->s.counters != nullptr, finite = bk->s.alpha_finite != 0;
  const int64_t per = ((nv + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
  const int64_t end = std::min<int64_t>(nv, per * (blockIdx.x + 1));
  const int lane = (int)__lane_id();
  if (threadIdx.x == 0) nbuf = 0;
  __syncthreads();
  auto flush = [&]() {
    const unsigned int n = nbuf;  // every thread's appends are in (a barrier precedes)
    if (threadIdx.x == 0) base = n ? atomicAdd(&f.rm_count[b], n) : 0u;
    __syncthreads();
    for (unsigned int i = threadIdx.x; i < n; i += 256) f.rm_list[bk->seed_off + base + i] = buf[i];
    __syncthreads();
    if (threadIdx.x == 0) nbuf = 0;
    __syncthreads();
  };
  // each iteration's seeds and local map words are read one iteration ahead, so their latency
  // overlaps the previous iteration's stores and barrier (the local word of every vertex is
  // read, most of them being copied)
  typedef typename std::conditional<MDT == kF32, uint32_t, uint16_t>::type Word;
  const float* seeds = f.seeds + bk->seed_off;
  const Word* local = (const Word*)bk->m.map;
  int64_t v0 = per * blockIdx.x;
  float next = 0.0f;
  Word next_w = 0;
  if (v0 + threadIdx.x < end) {
    next = seeds[v0 + threadIdx.x];
    next_w = local[v0 + threadIdx.x];
  }
  for (; v0 < end; v0 += 256) {
    const int64_t v = v0 + threadIdx.x;
    const float seed = next;
    const Word w = next_w;
    if (v + 256 < end) {
      next = seeds[v + 256];
      next_w = local[v + 256];
    }
    bool listed = false;
    if (v < end) {
      if (!counting && seed == 0.0f)
        ((Word*)bk->map_final)[v] = w;  // the local value, bit for bit
      else if (!counting && seed == 1.0f && finite)
        store_t<MDT>(bk->map_final, v, 1.0f);
      else
        listed = true;
    }
    const unsigned long long bal = __ballot(listed);
    if (bal) {
      const int leader = __builtin_ctzll(bal);
      unsigned int at = 0;
      if (lane == leader) at = atomicAdd(&nbuf, (unsigned int)__popcll(bal));
      at = (unsigned int)__shfl((int)at, leader);
      if (listed) buf[at + __popcll(bal & ((1ull << lane) - 1ull))] = (uint32_t)v;
    }
    __syncthreads();
    if (nbuf > kFinalBuf - 256) flush();  // room for the next round's 256 at most
  }
  flush();
}

// The listed vertices' map items, kMapItem per item: prefix sums over the bricks.
__global__ void k_rm_items(BrickFrame f)
{
  if (threadIdx.x != 0) return;
  int total = 0;
  for (int b = 0; b < f.nbricks; ++b) {
    f.rm_item_off[b] = total;
    total += (int)((f.rm_count[b] + kMapItem - 1) / kMapItem);
  }
  f.rm_item_off[f.nbricks] = total;
}

// Bounding sphere (centre c, radius R) of the map vertices [first, last] of a brick: the
// vertices of an x-fastest id range lie in the index box spanned by its rows and slices.
__device__ __forceinline__ void item_sphere(const MapParams& m, int64_t first, int64_t last, float* c,
                                            float& R)
{
  const int64_t n0 = m.mp[0], n01 = m.mp[0] * m.mp[1];
  int64_t lo[3] = { first % n0, (first / n0) % m.mp[1], first / n01 };
  int64_t hi[3] = { last % n0, (last / n0) % m.mp[1], last / n01 };
  if (lo[2] != hi[2]) {
    lo[0] = 0, hi[0] = m.mp[0] - 1, lo[1] = 0, hi[1] = m.mp[1] - 1;
  } else if (lo[1] != hi[1]) {
    lo[0] = 0, hi[0] = m.mp[0] - 1;
  }
  float r2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float pl = m.morg[a] + m.msp[a] * (float)lo[a];
    const float ph = m.morg[a] + m.msp[a] * (float)hi[a];
    c[a] = 0.5f * (pl + ph);
    r2 += 0.25f * (ph - pl) * (ph - pl);
  }
  R = sqrtf(r2) * 1.0001f + 1e-6f;
}

// Conservative test: can a light segment ending in the sphere (c, R) meet the sphere (s, r)?
// Point light at o: the segments lie in the cone from o through the item sphere, no longer
// than |c - o| + R.  Directional light: they lie in the capsule swept from the item sphere
// up-light by dir_far.
__device__ __forceinline__ bool may_cross(const MapParams& m, const float* c, float R, const float* s,
                                          float r)
{
  if (m.light_kind == 1) {
    // capsule axis from c to c - dir_far * dir
    float w[3], t = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      w[a] = s[a] - c[a];
      t -= w[a] * m.light[a];
    }
    t = fminf(fmaxf(t, 0.0f), m.dir_far);
    float d2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float q = w[a] + t * m.light[a];
      d2 += q * q;
    }
    const float lim = (R + r) * 1.001f + 1e-5f;
    return d2 <= lim * lim;
  }
  float ax[3], d[3], la2 = 0.0f, ld2 = 0.0f, dot = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ax[a] = c[a] - m.light[a];
    d[a] = s[a] - m.light[a];
    la2 += ax[a] * ax[a];
    ld2 += d[a] * d[a];
    dot += ax[a] * d[a];
  }
  const float la = sqrtf(la2), ld = sqrtf(ld2);
  if (la <= R * 1.001f || ld <= r * 1.001f) return true;  // the light inside either sphere
  if (ld - r > (la + R) * 1.001f) return false;           // beyond every segment's end
  const float cos_phi = fminf(fmaxf(dot / (la * ld), -1.0f), 1.0f);
  const float phi = acosf(cos_phi);
  const float half = asinf(fminf(R / la, 1.0f)) + asinf(fminf(r / ld, 1.0f));
  return phi <= half * 1.001f + 2e-3f;
}

// Does triangle (A, B, C) meet the box [bx0, bx1] x [bx2, bx3] x [bx4, bx5]?  The separating
// axis test of Akenine-Moeller: the box's face normals, the triangle's normal and the nine
// cross products of box axes and triangle edges.  In FP64 on FP32 inputs, so that a test
// against the inflated FP32 boxes (box32) keeps every box the exact segment test can hit.
__device__ __forceinline__ bool tri_box(const float* bx, const float* A, const float* B, const float* C)
{
  double h[3], v0[3], v1[3], v2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double lo = bx[2 * a], hi = bx[2 * a + 1], c = 0.5 * (lo + hi);
    h[a] = 0.5 * (hi - lo);
    v0[a] = (double)A[a] - c;
    v1[a] = (double)B[a] - c;
    v2[a] = (double)C[a] - c;
    if (fmin(v0[a], fmin(v1[a], v2[a])) > h[a] || fmax(v0[a], fmax(v1[a], v2[a])) < -h[a]) return false;
  }
  const double* vs[3] = { v0, v1, v2 };
  double e[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    e[0][a] = v1[a] - v0[a];
    e[1][a] = v2[a] - v1[a];
    e[2][a] = v0[a] - v2[a];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      // axis = unit(i) x e_k: components on the two other axes u = i+1, w = i+2
      const int u = (i + 1) % 3, w = (i + 2) % 3;
      const double au = -e[k][w], aw = e[k][u];
      double pmin = 1e300, pmax = -1e300;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double pq = au * vs[q][u] + aw * vs[q][w];
        pmin = fmin(pmin, pq);
        pmax = fmax(pmax, pq);
      }
      const double r = h[u] * fabs(au) + h[w] * fabs(aw);
      if (pmin > r || pmax < -r) return false;
    }
  }
  const double n[3] = { e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                        e[0][0] * e[1][1] - e[0][1] * e[1][0] };
  const double d = n[0] * v0[0] + n[1] * v0[1] + n[2] * v0[2];
  return fabs(d) <= h[0] * fabs(n[0]) + h[1] * fabs(n[1]) + h[2] * fabs(n[2]);
}

// Row r of brick bd's map walk: the r-th row in light-sheet order, or x-fastest row r.
__device__ __forceinline__ void row_of(const BrickFrame& f, const BrickDesc& bd, int64_t r, int64_t& j,
                                       int64_t& k)
{
  if (f.light_rows) {
    const uint32_t jk = f.rows[bd.row_off + r];
    j = jk & 0xffffu;
    k = jk >> 16;
  } else {
    // (a brick's map holds fewer than 2^31 vertices: 32-bit divisions)
    const uint32_t mp1 = (uint32_t)bd.m.mp[1], q = (uint32_t)r / mp1;
    j = (int64_t)((uint32_t)r - q * mp1);
    k = (int64_t)q;
  }
}

// Can a light segment of a map vertex in x-run [xa, xb] of row (j, k) of brick map m meet box bx?
// The vertices lie on the segment P0 P1 of the row, so their light segments lie in the
// triangle (light, P0, P1) for a point light, or in the parallelogram (P0, P1, O1, O0) of their
// up-light origins O = P - dir_far dir for a directional light.
__device__ __forceinline__ bool run_may_cross(const MapParams& m, int64_t j, int64_t k, int64_t xa,
                                              int64_t xb, const float* bx)
{
  const float y = m.morg[1] + m.msp[1] * (float)j, z = m.morg[2] + m.msp[2] * (float)k;
  const float P0[3] = { m.morg[0] + m.msp[0] * (float)xa, y, z };
  const float P1[3] = { m.morg[0] + m.msp[0] * (float)xb, y, z };
  if (m.light_kind != 1) return tri_box(bx, m.light, P0, P1);
  float O0[3], O1[3];
  light_origin(m, P0, O0);
  light_origin(m, P1, O1);
  return tri_box(bx, P0, P1, O1) || tri_box(bx, P0, O1, O0);
}

// ---- phase 2: non-local hits (TransmittanceMap.h:342-453) + fetch (:505-508) + composite
// (LightedVolumeRenderer.cxx:1214-1235), one thread per map vertex ------------------------------
// Per work item the wave first keeps the bricks whose bounding sphere may meet a light segment
// of the item (conservative; the per-vertex FP32 slab cull and the exact FP64 test follow).
constexpr int kMaxCandidates = 1000;

// One light segment (o -> p) against brick j: the conservative FP32 slab cull on the brick's
// inflated box, then the exact test (FP64 slab terms); key = (exit distance bits, brick).
__device__ __forceinline__ bool exchange_hit(const BrickFrame& f, int j, int b, const float* o,
                                             const Segment& sg, const float* inv, const float* lift,
                                             unsigned long long& key)
{
  if (j == b) return false;
  // j is the same for every lane and the brick table is read-only during the frame: reading
  // it through the constant address space makes these scalar loads, issued together
  typedef const __attribute__((address_space(4))) BrickDesc ConstBrick;
  ConstBrick* bk = (ConstBrick*)(const void*)(f.bricks + j);
  float bx[6];
  double bb[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    bx[k] = bk->box32[k];
    bb[k] = bk->g.bounds[k];
  }
  // On an axis the segment does not move along, inv[a] = 2^126 and lift[a] = 2: with o in
  // the closed slab the interval then starts at or below 0 and (lifted) ends above 1; with o
  // outside it lies beyond [0, 1], or (o within about 2^-126 of the slab) is kept.  The cull
  // never drops a hit of the exact test and never meets a NaN.
  float t0 = 0.0f, t1 = 1.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float ta = (bx[2 * a] - o[a]) * inv[a], tb = (bx[2 * a + 1] - o[a]) * inv[a];
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb) + lift[a]);
  }
  if (!(t0 <= t1 + 1e-3f)) return false;
  float tmin, tmax;
  if (!segment_box(sg, bb, (float)b, tmin, tmax)) return false;
  // tmax >= tmin >= 0; adding +0 turns a -0 into +0 so the bit order is the float order
  key = ((unsigned long long)f32_bits(tmax + 0.0f) << 32) | (unsigned long long)(unsigned)j;
  return true;
}

// Hits kept sorted per vertex in a lane-private LDS column (slot k of lane l: the key's exit
// distance bits at hit_t[k][l], its brick at hit_b[k][l]); a path crossing more bricks is
// composited by repeated selection of the next key over the candidates (same order, no
// storage).  The item's candidates are ordered by their distance along the light first, so
// hits mostly arrive in order and are appended.  Candidate and hit bricks are 16-bit (at most
// 4096 bricks): with a list of 1000 candidates a wave's LDS is 8160 B, so a CU holds 20 waves,
// as many as 96 VGPRs allow (five per SIMD).
constexpr int kRegHits = 16;

// Ordering key of a candidate brick: distance of its centre from the point light, or its
// centre's position along the directional light's travel direction.
__device__ __forceinline__ float cand_order(const MapParams& m, const float* bx)
{
  const float s[3] = { 0.5f * (bx[0] + bx[1]), 0.5f * (bx[2] + bx[3]), 0.5f * (bx[4] + bx[5]) };
  if (m.light_kind == 1) return s[0] * m.light[0] + s[1] * m.light[1] + s[2] * m.light[2];
  const float d[3] = { s[0] - m.light[0], s[1] - m.light[1], s[2] - m.light[2] };
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

template <int MDT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_bricks_exchange(BrickFrame f)
{
  __shared__ ItemQueue queue;
  __shared__ uint16_t cand[kMaxCandidates];
  __shared__ uint32_t hit_t[kRegHits][64];
  __shared__ uint16_t hit_b[kRegHits][64];
  const int lane = threadIdx.x;
  const int chunk = 1;
  queue_init(&queue);
  for (;;) {
    const int item = next_item(f.queue + 1 * kQueueSlots, &queue, f.n_map_items, chunk, f.xcd_ranges);
    if (item >= f.n_map_items) break;
    const int b = item_brick(f, item);
    const int64_t item_first = (int64_t)(item - f.map_item_off[b]) * kMapItem;
    const BrickDesc& bd = f.bricks[b];
    // candidate bricks of this item: those the light segments of its vertices may cross.  An
    // item of a few map rows is tested run by run against the triangles its light segments lie
    // in; one of many short rows against the cone (or capsule) through its bounding sphere.
    int ncand = 0;
    {
      const int64_t first = item_first;
      const int64_t last = first + kMapItem - 1 < bd.nverts - 1 ? first + kMapItem - 1 : bd.nverts - 1;
      const int64_t mp0 = bd.m.mp[0], r0 = first / mp0, r1 = last / mp0;
      const bool runs = r1 - r0 < 4;
      float c[3], R;
      // rows in light-sheet order are no longer neighbours: the whole brick bounds the item
      item_sphere(bd.m, f.light_rows ? 0 : first, f.light_rows ? bd.nverts - 1 : last, c, R);
      for (int j0 = 0; j0 < f.nbricks; j0 += 64) {
        const int j = j0 + lane;
        bool keep = false;
        if (j < f.nbricks && j != b) {
          const float* bx = f.bricks[j].box32;
          if (runs) {
            for (int64_t r = r0; r <= r1 && !keep; ++r) {
              int64_t rj, rk;
              row_of(f, bd, r, rj, rk);
              keep = run_may_cross(bd.m, rj, rk, r == r0 ? first % mp0 : 0, r == r1 ? last % mp0 : mp0 - 1, bx);
            }
          } else {
            const float s[3] = { 0.5f * (bx[0] + bx[1]), 0.5f * (bx[2] + bx[3]), 0.5f * (bx[4] + bx[5]) };
            const float e[3] = { bx[1] - bx[0], bx[3] - bx[2], bx[5] - bx[4] };
            const float rs = 0.5f * sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * 1.0001f + 1e-6f;
            keep = may_cross(bd.m, c, R, s, rs);
          }
        }
        const unsigned long long mask = __ballot(keep);
        const int pos = ncand + __popcll(mask & ((1ull << lane) - 1ull));
        if (keep && pos < kMaxCandidates) cand[pos] = (uint16_t)j;
        ncand += __popcll(mask);
      }
      __syncthreads();
      if (ncand <= 64) {
        // rank sort of up to 64 candidates by cand_order (ties by list position)
        const int j = lane < ncand ? cand[lane] : 0;
        const float key = lane < ncand ? cand_order(bd.m, f.bricks[j].box32) : 0.0f;
        int rank = 0;
        for (int k = 0; k < ncand; ++k) {
          const float kk = __shfl(key, k);
          rank += (kk < key || (kk == key && k < lane)) ? 1 : 0;
        }
        __syncthreads();
        if (lane < ncand) cand[rank] = (uint16_t)j;
        __syncthreads();
      }
    }
    const bool all = ncand > kMaxCandidates;  // too many to list: test every brick
    const int nloop = all ? f.nbricks : ncand;
    for (int sub = 0; sub < 4; ++sub) {
      const int64_t qv = item_first + sub * 64 + lane;
      if (qv >= bd.nverts) continue;
      // (a brick's map holds fewer than 2^31 vertices: a 32-bit division)
      const int64_t mp0 = bd.m.mp[0];
      const int64_t r = (int64_t)((uint32_t)qv / (uint32_t)mp0), i = qv - r * mp0;
      int64_t rj, rk;
      row_of(f, bd, r, rj, rk);
      const int64_t v = i + mp0 * (rj + bd.m.mp[1] * rk);
      // map_vertex's position of vertex v
      float p[3], o[3];
      p[0] = bd.m.morg[0] + bd.m.msp[0] * (float)i;
      p[1] = bd.m.morg[1] + bd.m.msp[1] * (float)rj;
      p[2] = bd.m.morg[2] + bd.m.msp[2] * (float)rk;
      light_origin(bd.m, p, o);
      const float dd[3] = { p[0] - o[0], p[1] - o[1], p[2] - o[2] };
      const float rs = 1.0f / __builtin_sqrtf(dot3(dd, dd));  // vtkm::Normalize
      const float dir[3] = { dd[0] * rs, dd[1] * rs, dd[2] * rs };
      float inv[3], lift[3];  // the cull's slab terms (exchange_hit)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        inv[a] = dd[a] == 0.0f ? 0x1p126f : 1.0f / dd[a];
        lift[a] = dd[a] == 0.0f ? 2.0f : 0.0f;
      }
      Segment sg;  // SegmentAABB's per-segment terms, shared by every candidate brick
      segment_prepare(o, p, sg);
      auto composite = [&](float op, unsigned long long key) {
        const int j = (int)(key & 0xffffffffull);
        const float t = bits_f32((uint32_t)(key >> 32));
        const float q[3] = { o[0] + dir[0] * t, o[1] + dir[1] * t, o[2] + dir[2] * t };
        // the lanes usually fetch from the same brick: its map parameters then come in scalar
        // registers (one load each for the wave, not one per lane)
        const int j0 = __builtin_amdgcn_readfirstlane(j);
        float local;
        if (__ballot(j != j0) == 0)
          local = map_opacity_uniform<MDT>(load_uniform(&f.bricks[j0].m), q);
        else
          local = map_opacity<MDT>(f.bricks[j].m, q);
        return op + (1.0f - op) * local;
      };
      // slot k of this lane's column: a key (exit distance bits << 32 | brick) split in two
      auto put = [&](int k, unsigned long long key) {
        hit_t[k][lane] = (uint32_t)(key >> 32);
        hit_b[k][lane] = (uint16_t)key;
      };
      auto get = [&](int k) {
        return ((unsigned long long)hit_t[k][lane] << 32) | (unsigned long long)hit_b[k][lane];
      };
      unsigned long long top = 0;  // the largest key so far (slot nh - 1)
      int nh = 0;
      bool spill = false;
      for (int q = 0; q < nloop; ++q) {
        unsigned long long key;
        // the same candidate for every lane of the wave
        const int j = __builtin_amdgcn_readfirstlane(all ? q : cand[q]);
        if (!exchange_hit(f, j, b, o, sg, inv, lift, key)) continue;
        if (nh == kRegHits) {
          spill = true;
          break;
        }
        if (nh == 0 || key > top) {
          put(nh, key);  // in order: append
          top = key;
        } else {
          // insertion: the larger keys move up one slot
          int k = nh;
          for (; k > 0; --k) {
            const unsigned long long prev = get(k - 1);
            if (prev < key) break;
            put(k, prev);
          }
          put(k, key);
        }
        ++nh;
      }
      float op = 0.0f;
      if (!spill) {
        for (int k = 0; k < nh; ++k) op = composite(op, get(k));
      } else {
        bool started = false;
        unsigned long long last = 0;
        for (;;) {
          unsigned long long best = ~0ull;
          for (int q = 0; q < nloop; ++q) {
            unsigned long long key;
            const int j = __builtin_amdgcn_readfirstlane(all ? q : cand[q]);
            if (!exchange_hit(f, j, b, o, sg, inv, lift, key)) continue;
            if ((!started || key > last) && key < best) best = key;
          }
          if (best == ~0ull) break;
          op = composite(op, best);
          last = best;
          started = true;
        }
      }
      f.seeds[bd.seed_off + v] = op;
    }
  }
}

// ---- phase 4: CalcRayStart + Sampler per brick over its screen footprint ---------------------
template <int FDT, int MDT, bool USE_MAP, bool FAST, bool BLK>
__global__ void __launch_bounds__(256) k_bricks_march(BrickFrame f, CameraParams cam)
{
  // One workgroup per (brick, 16x16 screen tile) item.  Not persistent: a loop over items
  // keeps the camera and brick parameters live across the march and doubles the registers.
  extern __shared__ float4 lds4[];
  // march_xcd_tiles: workgroup w runs on XCD w mod 8 (workgroups are dealt to the XCDs round
  // robin); it takes the (w / 8)-th item of that XCD's contiguous block of items, so each XCD's
  // L2 serves one screen region of a brick instead of every eighth tile of all of them
  int item = blockIdx.x;
  if (f.march_xcd_tiles) {
    const int n = (int)gridDim.x, xcd = (int)(blockIdx.x & 7u), k = (int)(blockIdx.x >> 3);
    item = xcd * (n >> 3) + min(xcd, n & 7) + k;
  }
  float4* tf = lds4;
  float* xs = (float*)(lds4 + f.bricks[0].s.ntf);
  // the item's brick is the same for the whole workgroup: its parameters are read with
  // scalar loads (constant address space) into scalar registers, not into every lane
  const int b = __builtin_amdgcn_readfirstlane(f.px_item_brick[item]);
  const BrickDesc* bk = f.bricks + b;
  GridParams g = load_uniform(&bk->g);
  const ShadeParams sp = load_uniform(&bk->s);
  MapParams m = load_uniform(&bk->m);
  m.map = bk->map_final;
  if (BLK) {
    // the blocked copies: same points, the block strides in place of the row and slice strides
    g.field = bk->field_blk;
    g.bytes32 = bk->fblk_bytes;
    g.field_stride[1] = bk->fblk_ybs;
    g.field_stride[2] = bk->fblk_zbs;
    m.map = bk->map_blk;
    m.bytes32 = bk->mblk_bytes;
    m.ms[1] = bk->mblk_ybs;
    m.ms[2] = bk->mblk_zbs;
  }
  const int fp0 = bk->fp[0], fp1 = bk->fp[1], fp2 = bk->fp[2], fp3 = bk->fp[3];
  const int64_t img_off = bk->img_off;
  stage_tf(sp, tf);
  stage_coords(g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const int tiles_x = (fp2 + 15) / 16;
  const int tile = f.px_item_tile[item];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lx = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
  const int ly = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
  if (lx >= fp2 || ly >= fp3) return;
  float org[3], dir[3];
  camera_ray(cam, fp0 + lx, fp1 + ly, org, dir);
  float minD = cam.near_d, maxD = cam.far_d, dist = 0.0f;
  calc_ray_start(g, org, dir, minD, dist, maxD);
  float color[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
  shade<FDT, MDT, USE_MAP, false, FAST, BLK>(g, m, sp, tf, x, org, dir, minD, maxD, color);
  const int64_t pix = img_off + (int64_t)ly * fp2 + lx;
  f.partial[pix] = make_float4(color[0], color[1], color[2], color[3]);
  f.partial_depth[pix] = (minD == -1.0f) ? __builtin_inff() : minD;
}

// ---- blocked copies for the march ---------------------------------------------------------
__global__ void __launch_bounds__(256) k_block_extract(const unsigned char* src, int32_t esz, int64_t sx,
                                                       int64_t sxy, int64_t lo0, int64_t lo1, int64_t lo2,
                                                       int64_t n0, int64_t n1, int64_t n2, int64_t ybs, int64_t zbs,
                                                       unsigned char* dst)
{
  const int64_t total = n0 * n1 * n2;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
    const int64_t s = (lo0 + i) + (lo1 + j) * sx + (lo2 + k) * sxy;
    const int64_t d = blocked_index(i, j, k, ybs, zbs);
    if (esz == 4)
      ((uint32_t*)dst)[d] = ((const uint32_t*)src)[s];
    else
      ((uint16_t*)dst)[d] = ((const uint16_t*)src)[s];
  }
}

// map_final -> map_blk in destination order (blockIdx.y: the brick): consecutive threads write
// consecutive elements of the blocked copy, whole lines at a time, and gather their values from
// the x-fastest map (four consecutive x per row). Each workgroup copies one contiguous run of
// blocks, so the 16 blocks that share a row's line are read by one workgroup, on one XCD, from
// its L2 (interleaved runs put them on four XCDs, each fetching the line from HBM).  A thread
// keeps one element of every fourth block, its block coordinates counted up as it goes.
template <int MDT>
__global__ void __launch_bounds__(256) k_block_maps(BrickFrame f)
{
  const BrickDesc* bk = f.bricks + blockIdx.y;
  const int mp0 = (int)bk->m.mp[0], mp1 = (int)bk->m.mp[1], mp2 = (int)bk->m.mp[2];
  const int nbx = (mp0 + 3) / 4, nby = (mp1 + 3) / 4;
  const int64_t nblk = (int64_t)nbx * nby * ((mp2 + 3) / 4);
  const int64_t per = (nblk + gridDim.x - 1) / gridDim.x;
  const int64_t end = std::min<int64_t>(nblk, per * (blockIdx.x + 1));
  int64_t blk = per * blockIdx.x + (threadIdx.x >> 6);
  const int w = (int)threadIdx.x & 63, di = w & 3, dj = (w >> 2) & 3, dk = w >> 4;
  const int64_t q = blk / nbx;
  int bx = (int)(blk - q * nbx), by = (int)(q % nby), bz = (int)(q / nby);
  // sixteen blocks' gathers in flight before their stores
  constexpr int kInFlight = 16;
  typedef typename std::conditional<MDT == kF32, uint32_t, uint16_t>::type Word;
  for (; blk < end; blk += 4 * kInFlight) {
    Word v[kInFlight];
    bool ok[kInFlight];
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      const int i = 4 * bx + di, j = 4 * by + dj, k = 4 * bz + dk;
      ok[u] = blk + 4 * u < end && i < mp0 && j < mp1 && k < mp2;  // not past the run or the padding
      v[u] = ok[u] ? ((const Word*)bk->map_final)[i + (int64_t)mp0 * (j + (int64_t)mp1 * k)] : (Word)0;
      bx += 4;
      while (bx >= nbx) {
        bx -= nbx;
        if (++by == nby) {
          by = 0;
          ++bz;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kInFlight; ++u)
      if (ok[u]) ((Word*)bk->map_blk)[(blk + 4 * u) * 64 + w] = v[u];
  }
}

// ---- phase 5: GlobalComposite, front-to-back over in camera-distance order -------------------
__global__ void __launch_bounds__(256) k_bricks_composite(BrickFrame f, int32_t width,
                                                          int32_t height, ImageOut out)
{
  const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (pix >= (int64_t)width * height) return;
  const int px = (int)(pix % width), py = (int)(pix / width);
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  float depth = __builtin_inff();
  for (int k = 0; k < f.nbricks; ++k) {
    const BrickDesc& bd = f.bricks[f.depth_order[k]];
    const int lx = px - bd.fp[0], ly = py - bd.fp[1];
    if (lx < 0 || ly < 0 || lx >= bd.fp[2] || ly >= bd.fp[3]) continue;  // contributes 0
    const int64_t q = bd.img_off + (int64_t)ly * bd.fp[2] + lx;
    const float4 c = f.partial[q];
    const float ia = 1.0f - a3;
    a0 = a0 + c.x * ia;
    a1 = a1 + c.y * ia;
    a2 = a2 + c.z * ia;
    a3 = a3 + c.w * ia;
    depth = fminf(depth, f.partial_depth[q]);
  }
  if (out.use_bg) {
    const float ia = 1.0f - a3;
    a0 = a0 + out.bg[0] * ia;
    a1 = a1 + out.bg[1] * ia;
    a2 = a2 + out.bg[2] * ia;
    a3 = a3 + out.bg[3] * ia;
  }
  ((float4*)out.rgba)[pix] = make_float4(a0, a1, a2, a3);
  if (out.depth) out.depth[pix] = depth;
}

__global__ void __launch_bounds__(256) k_extract(const unsigned char* src, int32_t eb, int64_t s0,
                                                 int64_t s1, int64_t lo0, int64_t lo1, int64_t lo2,
                                                 int64_t d0, int64_t d1, int64_t d2,
                                                 unsigned char* dst)
{
  const int64_t total = d0 * d1 * d2;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int64_t x = i % d0, y = (i / d0) % d1, z = i / (d0 * d1);
    const int64_t si = ((z + lo2) * s1 + (y + lo1)) * s0 + (x + lo0);
    if (eb == 2)
      ((uint16_t*)dst)[i] = ((const uint16_t*)src)[si];
    else
      ((uint32_t*)dst)[i] = ((const uint32_t*)src)[si];
  }
}

__global__ void k_zero_u32(unsigned int* p, int n)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0u;
}

}  // namespace

#define BRICK_DISPATCH_FIELD(V, BODY)                        \
  switch (V) {                                               \
    case kF32: { constexpr int FDT = kF32; BODY; } break;    \
    case kF16: { constexpr int FDT = kF16; BODY; } break;    \
    case kBF16: { constexpr int FDT = kBF16; BODY; } break;  \
    default: return hipErrorInvalidValue;                    \
  }
#define BRICK_DISPATCH_MAP(V, BODY)                          \
  switch (V) {                                               \
    case kF32: { constexpr int MDT = kF32; BODY; } break;    \
    case kF16: { constexpr int MDT = kF16; BODY; } break;    \
    default: return hipErrorInvalidValue;                    \
  }

hipError_t launch_bricks_map(const BrickFrame& f, int32_t field_dtype, int32_t map_dtype,
                             int32_t ntf, bool seeded, hipStream_t st)
{
  if (f.n_map_items == 0) return hipSuccess;
  const size_t lds = sizeof(float) * (size_t)(ntf + 2 * f.max_coords);
  const dim3 grid((unsigned)f.persistent_groups);
  BRICK_DISPATCH_FIELD(field_dtype, BRICK_DISPATCH_MAP(map_dtype, {
    if (f.map_fast && f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_map<FDT, MDT, true, true>), grid, dim3(256), lds, st, f, seeded ? 1 : 0);
    else if (f.map_fast)
      hipLaunchKernelGGL((k_bricks_map<FDT, MDT, true, false>), grid, dim3(256), lds, st, f, seeded ? 1 : 0);
    else if (f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_map<FDT, MDT, false, true>), grid, dim3(256), lds, st, f, seeded ? 1 : 0);
    else
      hipLaunchKernelGGL((k_bricks_map<FDT, MDT, false, false>), grid, dim3(256), lds, st, f, seeded ? 1 : 0);
  }));
  return hipGetLastError();
}

hipError_t launch_bricks_finalize(const BrickFrame& f, int32_t map_dtype, hipStream_t st)
{
  if (f.nbricks < 1) return hipSuccess;
  const dim3 grid((unsigned)std::max(8, 4096 / f.nbricks), (unsigned)f.nbricks);
  BRICK_DISPATCH_MAP(map_dtype, { hipLaunchKernelGGL((k_bricks_finalize<MDT>), grid, dim3(256), 0, st, f); });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rm_items, dim3(1), dim3(64), 0, st, f);
  return hipGetLastError();
}

hipError_t launch_bricks_exchange(const BrickFrame& f, int32_t map_dtype, int64_t total_verts,
                                  hipStream_t st)
{
  (void)total_verts;
  if (f.n_map_items == 0) return hipSuccess;
  const dim3 grid((unsigned)(f.persistent_groups * 4));
  BRICK_DISPATCH_MAP(map_dtype, {
    hipLaunchKernelGGL((k_bricks_exchange<MDT>), grid, dim3(64), 0, st, f);
  });
  return hipGetLastError();
}

hipError_t launch_bricks_march(const BrickFrame& f, int32_t field_dtype, int32_t map_dtype,
                               int32_t ntf, bool use_map, const CameraParams& c, hipStream_t st)
{
  if (f.n_px_items == 0) return hipSuccess;
  size_t lds = sizeof(float4) * (size_t)ntf + sizeof(float) * 2 * (size_t)f.max_coords;
  // march_groups_per_cu > 0: LDS padded so that at most that many workgroups share a CU (a
  // larger share of the XCD's L2 for each tile's rays)
  if (f.march_groups_per_cu > 0) lds = std::max(lds, (size_t)(160 * 1024) / (size_t)f.march_groups_per_cu - 512);
  const dim3 grid((unsigned)f.n_px_items);
  BRICK_DISPATCH_FIELD(field_dtype, BRICK_DISPATCH_MAP(map_dtype, {
    if (f.march_blocked && use_map)
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, true, true, true>), grid, dim3(256), lds, st, f, c);
    else if (f.march_blocked)
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, false, true, true>), grid, dim3(256), lds, st, f, c);
    else if (use_map && f.march_fast)
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, true, true, false>), grid, dim3(256), lds, st, f, c);
    else if (use_map)
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, true, false, false>), grid, dim3(256), lds, st, f, c);
    else if (f.march_fast)
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, false, true, false>), grid, dim3(256), lds, st, f, c);
    else
      hipLaunchKernelGGL((k_bricks_march<FDT, MDT, false, false, false>), grid, dim3(256), lds, st, f, c);
  }));
  return hipGetLastError();
}

hipError_t launch_bricks_composite(const BrickFrame& f, const CameraParams& c, const ImageOut& o,
                                   hipStream_t st)
{
  const int64_t npix = (int64_t)c.width * c.height;
  hipLaunchKernelGGL(k_bricks_composite, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, f,
                     c.width, c.height, o);
  return hipGetLastError();
}

hipError_t launch_block_extract(const void* src, int32_t elem_bytes, const int64_t src_dims[3],
                                const int64_t lo[3], const int64_t n[3], void* dst, hipStream_t st)
{
  int64_t ybs, zbs;
  blocked_strides(n, ybs, zbs);
  const int64_t total = n[0] * n[1] * n[2];
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536)));
  hipLaunchKernelGGL(k_block_extract, grid, dim3(256), 0, st, (const unsigned char*)src, elem_bytes, src_dims[0],
                     src_dims[0] * src_dims[1], lo[0], lo[1], lo[2], n[0], n[1], n[2], ybs, zbs,
                     (unsigned char*)dst);
  return hipGetLastError();
}

hipError_t launch_block_maps(const BrickFrame& f, int32_t map_dtype, hipStream_t st)
{
  if (f.nbricks < 1) return hipSuccess;
  const dim3 grid((unsigned)std::max(8, 4096 / f.nbricks), (unsigned)f.nbricks);
  if (map_dtype == kF32)
    hipLaunchKernelGGL((k_block_maps<kF32>), grid, dim3(256), 0, st, f);
  else
    hipLaunchKernelGGL((k_block_maps<kF16>), grid, dim3(256), 0, st, f);
  return hipGetLastError();
}

hipError_t launch_extract_brick(const void* src, int32_t elem_bytes, const int64_t src_dims[3],
                                const int64_t lo[3], const int64_t dims[3], void* dst,
                                hipStream_t st)
{
  hipLaunchKernelGGL(k_extract, dim3(1024), dim3(256), 0, st, (const unsigned char*)src,
                     elem_bytes, src_dims[0], src_dims[1], lo[0], lo[1], lo[2], dims[0], dims[1],
                     dims[2], (unsigned char*)dst);
  return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t st)
{
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::max(1, (n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace beams
