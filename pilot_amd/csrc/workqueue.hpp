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
struct ItemQueue {
  int next, end, item;
};

__device__ __forceinline__ void queue_init(ItemQueue* q)
{
  if (threadIdx.x == 0) q->next = q->end = 0;
}

// The next item of the workgroup (uniform across the group), nitems when the work is done.
__device__ __forceinline__ int next_item(unsigned int* counters, ItemQueue* q, int nitems, int chunk)
{
  __syncthreads();  // every thread is done with the previous item and its LDS
  if (threadIdx.x == 0) {
    if (q->next >= q->end) {
      const unsigned slots = gridDim.x < 8u ? gridDim.x : 8u;
      const unsigned x = blockIdx.x % slots;
      const int64_t k = (int64_t)atomicAdd(&counters[x], 1u);
      const int64_t first = (k * slots + x) * chunk;
      q->next = first < nitems ? (int)first : nitems;
      q->end = first + chunk < nitems ? (int)(first + chunk) : nitems;
    }
    q->item = q->next < q->end ? q->next++ : nitems;
  }
  __syncthreads();
  return q->item;
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
