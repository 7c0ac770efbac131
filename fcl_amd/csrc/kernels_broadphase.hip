// Broadphase on the device: every pair of objects whose world AABBs overlap
// (AABB::overlap, math/bv/AABB-inl.h: closed intervals on the three axes,
// so touching boxes overlap), the candidate list a BroadPhaseCollisionManager
// hands to its callback (broadphase_collision_manager.h:77-122).
//
// Sort and sweep: the objects are sorted by their box's lower x; the lane of
// object i walks the sorted list from i's position while the next lower x
// does not exceed i's upper x, and keeps the pairs that also overlap in y and
// z (and, with groups, that join objects of different groups).  The pairs
// are then sorted by (lower index, higher index), so the output does not
// depend on the order in which lanes reserved their slots.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace fclx {

namespace {

__global__ void bp_keys(int n, const double* aabbs, double* keys, int32_t* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = aabbs[6 * static_cast<size_t>(i)];
  idx[i] = i;
}

__global__ void bp_zero(unsigned long long* p) { *p = 0; }

// unused pair slots sort after every real (lower, higher) key
__global__ void bp_fill(uint64_t* p, long long n) {
  const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = ~0ULL;
}

__global__ void __launch_bounds__(256) bp_sweep(int n, const double* aabbs, const int32_t* groups,
                                                const int32_t* sorted, uint64_t* out,
                                                long long capacity, unsigned long long* count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = sorted[k];
  const double* a = aabbs + 6 * static_cast<size_t>(i);
  const double ax0 = a[0], ay0 = a[1], az0 = a[2], ax1 = a[3], ay1 = a[4], az1 = a[5];
  const int gi = groups ? groups[i] : 0;
  for (int m = k + 1; m < n; ++m) {
    const int j = sorted[m];
    const double* b = aabbs + 6 * static_cast<size_t>(j);
    if (b[0] > ax1) break;  // sorted by lower x: no later box reaches i
    if (groups && groups[j] == gi) continue;
    if (ay0 > b[4] || b[1] > ay1 || az0 > b[5] || b[2] > az1 || ax0 > b[3]) continue;
    const unsigned long long slot = atomicAdd(count, 1ULL);
    if (static_cast<long long>(slot) < capacity) {
      const uint32_t lo = static_cast<uint32_t>(i < j ? i : j);
      const uint32_t hi = static_cast<uint32_t>(i < j ? j : i);
      out[slot] = (static_cast<uint64_t>(lo) << 32) | hi;
    }
  }
}

__global__ void bp_unpack(const uint64_t* keys, long long n, int32_t* pairs) {
  const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  pairs[2 * k] = static_cast<int32_t>(keys[k] >> 32);
  pairs[2 * k + 1] = static_cast<int32_t>(keys[k] & 0xffffffffu);
}

__global__ void bp_store_count(const unsigned long long* count, long long* num_pairs) {
  *num_pairs = static_cast<long long>(*count);
}

}  // namespace

hipError_t launch_broadphase(int n, const double* aabbs, const int32_t* groups, int32_t* pairs,
                             long long capacity, long long* num_pairs, hipStream_t s) {
  if (n <= 1) {
    bp_zero<<<1, 1, 0, s>>>(reinterpret_cast<unsigned long long*>(num_pairs));
    return hipGetLastError();
  }
  const int block = 256;
  const int grid = (n + block - 1) / block;
  // scratch: keys and indices (sorted and unsorted), the pair keys (twice),
  // the counter, and the radix sorts' temporary storage
  const long long cap = capacity > 0 ? capacity : 0;
  double *keys = nullptr, *keys_sorted = nullptr;
  int32_t *idx = nullptr, *idx_sorted = nullptr;
  uint64_t *pk = nullptr, *pk_sorted = nullptr;
  unsigned long long* count = nullptr;
  void* tmp = nullptr;
  size_t t1 = 0, t2 = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, t1, keys, keys_sorted, idx, idx_sorted,
                                                    n, 0, 64, s);
  if (e == hipSuccess && cap > 0)
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, t2, pk, pk_sorted, static_cast<int>(cap), 0, 64,
                                          s);
  const size_t tb = t1 > t2 ? t1 : t2;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMallocAsync(p, bytes > 0 ? bytes : 8, s);
  };
  alloc(reinterpret_cast<void**>(&keys), n * sizeof(double));
  alloc(reinterpret_cast<void**>(&keys_sorted), n * sizeof(double));
  alloc(reinterpret_cast<void**>(&idx), n * sizeof(int32_t));
  alloc(reinterpret_cast<void**>(&idx_sorted), n * sizeof(int32_t));
  alloc(reinterpret_cast<void**>(&pk), (cap > 0 ? cap : 1) * sizeof(uint64_t));
  alloc(reinterpret_cast<void**>(&pk_sorted), (cap > 0 ? cap : 1) * sizeof(uint64_t));
  alloc(reinterpret_cast<void**>(&count), sizeof(unsigned long long));
  alloc(&tmp, tb);
  if (e == hipSuccess) {
    bp_keys<<<grid, block, 0, s>>>(n, aabbs, keys, idx);
    bp_zero<<<1, 1, 0, s>>>(count);
    if (cap > 0) bp_fill<<<static_cast<unsigned>((cap + block - 1) / block), block, 0, s>>>(pk, cap);
    e = hipcub::DeviceRadixSort::SortPairs(tmp, t1, keys, keys_sorted, idx, idx_sorted, n, 0, 64,
                                           s);
  }
  if (e == hipSuccess) {
    bp_sweep<<<grid, block, 0, s>>>(n, aabbs, groups, idx_sorted, pk, cap, count);
    bp_store_count<<<1, 1, 0, s>>>(count, num_pairs);
    e = hipGetLastError();
  }
  // sort the pairs found (up to the capacity): the count is only known on
  // the device, so the whole capacity is sorted, unused slots last
  if (e == hipSuccess && cap > 0) {
    e = hipcub::DeviceRadixSort::SortKeys(tmp, t2, pk, pk_sorted, static_cast<int>(cap), 0, 64, s);
    if (e == hipSuccess) {
      bp_unpack<<<static_cast<unsigned>((cap + block - 1) / block), block, 0, s>>>(pk_sorted, cap,
                                                                                 pairs);
      e = hipGetLastError();
    }
  }
  for (void* p : {static_cast<void*>(keys), static_cast<void*>(keys_sorted), static_cast<void*>(idx),
                  static_cast<void*>(idx_sorted), static_cast<void*>(pk),
                  static_cast<void*>(pk_sorted), static_cast<void*>(count), tmp})
    if (p) (void)hipFreeAsync(p, s);
  return e;
}

}  // namespace fclx
