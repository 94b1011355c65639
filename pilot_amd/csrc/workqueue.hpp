// Device helpers shared by the brick kernels: the persistent workgroups' work queue and scalar
// (uniform) copies of per-brick parameter blocks.
#pragma once

#include "kernels.hpp"

namespace beams {

// Work queue of a persistent workgroup.  Items are handed out one at a time from up to eight
// counters (one per XCD: workgroups are dealt to the XCDs round-robin), counter x handing out
// items x, x + S, x + 2S, ... with S = min(8, workgroups) counters in use, so that every
// counter has a workgroup behind it whatever the grid: the workgroups always work on the next
// few thousand consecutive items, so the bricks they read stay few and cached (larger chunks,
// or a fixed stride whose workgroups drift apart, spread the reads over many bricks and
// slowed the 1024^3 maps by half), while the device-scope atomics are spread over eight
// addresses instead of one.
//
// With `ranges` set the items are instead cut into S contiguous ranges, counter r handing out
// the items of range r in order: a workgroup starts on its XCD's range and moves on to the
// following ranges once that one is drained, so each XCD's L2 serves one compact region of
// the work while no XCD idles before the end.  Every item is handed out exactly once either way.
struct ItemQueue {
  int next, end, item, range;
};

__device__ __forceinline__ void queue_init(ItemQueue* q)
{
  if (threadIdx.x == 0) {
    q->next = q->end = 0;
    q->range = (int)(blockIdx.x % (gridDim.x < 8u ? gridDim.x : 8u));
  }
}

// The next item of the workgroup (uniform across the group), nitems when the work is done.
__device__ __forceinline__ int next_item(unsigned int* counters, ItemQueue* q, int nitems, int chunk,
                                         int ranges = 0)
{
  __syncthreads();  // every thread is done with the previous item and its LDS
  if (threadIdx.x == 0) {
    if (q->next >= q->end) {
      const unsigned slots = gridDim.x < 8u ? gridDim.x : 8u;
      if (!ranges) {
        const unsigned x = blockIdx.x % slots;
        const int64_t k = (int64_t)atomicAdd(&counters[x], 1u);
        const int64_t first = (k * slots + x) * chunk;
        q->next = first < nitems ? (int)first : nitems;
        q->end = first + chunk < nitems ? (int)(first + chunk) : nitems;
      } else {
        q->next = q->end = nitems;
        // this workgroup's range first, then the following ones (each tried once)
        for (unsigned t = 0; t < slots; ++t) {
          const unsigned r = ((unsigned)q->range + t) % slots;
          const int64_t lo = (int64_t)nitems * r / slots, hi = (int64_t)nitems * (r + 1) / slots;
          const int64_t first = lo + (int64_t)atomicAdd(&counters[r], 1u) * chunk;
          if (first < hi) {
            q->next = (int)first;
            q->end = (int)(first + chunk < hi ? first + chunk : hi);
            q->range = (int)r;
            break;
          }
        }
      }
    }
    q->item = q->next < q->end ? q->next++ : nitems;
  }
  __syncthreads();
  return q->item;
}

// The brick of map work item `item` (uniform): the last b with map_item_off[b] <= item.
// The last b with offs[b] <= item (uniform) for prefix sums offs[0 .. nbricks] that no kernel
// writes while this one runs.
__device__ __forceinline__ int item_brick_in(const int32_t* offs, int nbricks, int item)
{
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  ConstInt* off = (ConstInt*)(const void*)offs;
  item = __builtin_amdgcn_readfirstlane(item);
  int lo = 0, hi = nbricks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

template <class Frame>
__device__ __forceinline__ int item_brick(const Frame& f, int item)
{
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  ConstInt* off = (ConstInt*)(const void*)f.map_item_off;
  item = __builtin_amdgcn_readfirstlane(item);
  int lo = 0, hi = f.nbricks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// A copy of a struct that every lane reads at the same address and that no kernel writes
// during the frame: word by word through the constant address space, i.e. scalar loads into
// scalar registers (the words a kernel never uses are not loaded).
template <class T>
__device__ __forceinline__ T load_uniform(const T* p)
{
  static_assert(sizeof(T) % 4 == 0, "word-sized struct");
  typedef const __attribute__((address_space(4))) uint32_t ConstWord;
  ConstWord* q = (ConstWord*)(const void*)p;
  T out;
  uint32_t* o = (uint32_t*)&out;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = q[i];
  return out;
}

}  // namespace beams
