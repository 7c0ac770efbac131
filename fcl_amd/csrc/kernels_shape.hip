// Batched primitive-shape collide(): FP64, FCL's GST_INDEP solver (closed
// forms for the five special pairs, GJK + EPA for every other pair).  The
// per-pair bookkeeping follows shape_collision_traversal_node-inl.h:69-134:
// with enable_contact the pair's contacts (deepest first when more than
// num_max_contacts), without it one contact and no geometry.
//
// One pair per lane, in three stages so that the lanes of a wavefront run
// the same code:
//   1. classify + scatter: a counting sort of the pairs into 25 buckets by
//      ordered type pair, the eight closed-form buckets first (malformed
//      pairs get count 0 here and are dropped);
//   2. closed forms over the first buckets; GJK over the rest, where a pair
//      GJK finds inside is queued for EPA when contacts are wanted (without
//      them it collides, and EPA -- whose status is never Failed after GJK's
//      Inside -- would only have supplied the contact);
//   3. EPA over the queue: a persistent grid whose wavefronts take 64 queued
//      pairs at a time; each lane re-runs the pair's GJK (deterministic)
//      and keeps EPA's faces and vertices in a contiguous block of device
//      memory of its own.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/shape64.h"
#include "kernels.h"

namespace fclx {

static_assert(sizeof(fclx_shape) == 40, "fclx_shape layout");

namespace {

constexpr int kBlk = 256;
constexpr int kPerThread = 4;    // pairs per thread in the sorting kernels
constexpr int kTypes = 5;
constexpr int kBuckets = 26;     // 25 ordered type pairs + malformed

// counters
enum : int {
  kCtrCount = 0,                 // per bucket
  kCtrCursor = kCtrCount + 32,   // per bucket: next free position
  kCtrStart = kCtrCursor + 32,   // per bucket: first position
  kCtrClosed = kCtrStart + 32,   // closed-form pairs (positions [0, closed))
  kCtrValid,                     // well-formed pairs (positions [0, valid))
  kCtrQueue,                     // pairs queued for EPA
  kCtrGrab,                      // EPA queue entries taken
  kNumCtr
};

constexpr bool closed_pair(int t1, int t2) {
  return t1 == t2 ? (t1 == shp::kSphere || t1 == shp::kBox)
                  : ((t1 == shp::kSphere &&
                      (t2 == shp::kBox || t2 == shp::kCapsule || t2 == shp::kCylinder)) ||
                     (t2 == shp::kSphere &&
                      (t1 == shp::kBox || t1 == shp::kCapsule || t1 == shp::kCylinder)));
}

struct BucketMap {
  int of[kBuckets];  // bucket of key t1 * 5 + t2 (25: malformed)
  int closed;        // buckets [0, closed) are the closed-form ones
};

constexpr BucketMap make_bucket_map() {
  BucketMap b{};
  int k = 0;
  for (int key = 0; key < 25; ++key)
    if (closed_pair(key / kTypes, key % kTypes)) b.of[key] = k++;
  b.closed = k;
  for (int key = 0; key < 25; ++key)
    if (!closed_pair(key / kTypes, key % kTypes)) b.of[key] = k++;
  b.of[25] = 25;
  return b;
}

constexpr BucketMap kMap = make_bucket_map();
static_assert(kMap.closed == 8, "eight ordered closed-form type pairs");
__constant__ BucketMap d_map = make_bucket_map();

// A shape record is usable when its type is known and a convex names a
// non-empty vertex range inside the pool.
__device__ __forceinline__ bool load_shape(const ShapeLaunch& L, int id, shp::Shape& s) {
  if (id < 0 || id >= L.num_shapes) return false;
  const fclx_shape& r = L.shapes[id];
  s.type = r.type;
  s.nv = r.num_vertices;
  s.dims[0] = r.dims[0];
  s.dims[1] = r.dims[1];
  s.dims[2] = r.dims[2];
  s.verts = nullptr;
  switch (r.type) {
    case shp::kBox:
    case shp::kSphere:
    case shp::kCapsule:
    case shp::kCylinder:
      return true;
    case shp::kConvex:
      if (r.num_vertices < 1 || r.vertex_offset < 0 ||
          r.vertex_offset + r.num_vertices > L.num_vertices || L.vertices == nullptr)
        return false;
      s.verts = L.vertices + 3 * r.vertex_offset;
      return true;
  }
  return false;
}

__device__ __forceinline__ void write_contacts(const ShapeLaunch& L, int q, const shp::Contact* c,
                                               int n) {
  L.counts[q] = n;
  if (!L.contact) return;
  const size_t base = static_cast<size_t>(q) * L.cap;
  for (int i = 0; i < n; ++i) {
    const size_t k = base + i;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      L.contact_pos[3 * k + j] = c[i].pos.v[j];
      L.contact_normal[3 * k + j] = c[i].normal.v[j];
    }
    L.contact_depth[k] = c[i].depth;
  }
}

__global__ void shape_zero(int32_t* ctr) {
  for (int i = threadIdx.x; i < kNumCtr; i += blockDim.x) ctr[i] = 0;
}

// Validates every pair, records its bucket and counts the buckets.
__global__ void __launch_bounds__(kBlk) shape_classify(ShapeLaunch L, uint8_t* keys,
                                                       int32_t* ctr) {
  __shared__ int hist[kBuckets];
  for (int i = threadIdx.x; i < kBuckets; i += kBlk) hist[i] = 0;
  __syncthreads();
  const long long first = static_cast<long long>(blockIdx.x) * kBlk * kPerThread;
  for (int k = 0; k < kPerThread; ++k) {
    const long long q = first + k * kBlk + threadIdx.x;
    if (q >= L.num_pairs) break;
    shp::Shape s1, s2;
    int b;
    if (load_shape(L, L.pair_shapes[2 * q], s1) && load_shape(L, L.pair_shapes[2 * q + 1], s2)) {
      b = d_map.of[s1.type * kTypes + s2.type];
    } else {
      b = kBuckets - 1;
      L.counts[q] = 0;
      atomicOr(L.status, kStatusBadQuery);
    }
    keys[q] = static_cast<uint8_t>(b);
    atomicAdd(&hist[b], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBuckets; i += kBlk)
    if (hist[i]) atomicAdd(&ctr[kCtrCount + i], hist[i]);
}

__global__ void shape_scan(int32_t* ctr) {
  if (threadIdx.x != 0) return;
  int run = 0;
  for (int b = 0; b < kBuckets; ++b) {
    if (b == kMap.closed) ctr[kCtrClosed] = run;
    ctr[kCtrStart + b] = run;
    ctr[kCtrCursor + b] = run;
    if (b == kBuckets - 1) ctr[kCtrValid] = run;
    run += ctr[kCtrCount + b];
  }
}

// Positions within a bucket are reserved per block; their order inside the
// bucket is arbitrary (every pair's result is independent of it).
__global__ void __launch_bounds__(kBlk) shape_scatter(int n, const uint8_t* keys, int32_t* ctr,
                                                      int32_t* perm) {
  __shared__ int hist[kBuckets];
  __shared__ int base[kBuckets];
  for (int i = threadIdx.x; i < kBuckets; i += kBlk) hist[i] = 0;
  __syncthreads();
  const long long first = static_cast<long long>(blockIdx.x) * kBlk * kPerThread;
  int key[kPerThread], rank[kPerThread];
  for (int k = 0; k < kPerThread; ++k) {
    const long long q = first + k * kBlk + threadIdx.x;
    key[k] = -1;
    if (q < n) {
      key[k] = keys[q];
      rank[k] = atomicAdd(&hist[key[k]], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBuckets; i += kBlk)
    if (hist[i]) base[i] = atomicAdd(&ctr[kCtrCursor + i], hist[i]);
  __syncthreads();
  for (int k = 0; k < kPerThread; ++k)
    if (key[k] >= 0 && key[k] < kBuckets - 1)
      perm[base[key[k]] + rank[k]] = static_cast<int32_t>(first + k * kBlk + threadIdx.x);
}

__global__ void __launch_bounds__(kBlk) shape_closed(ShapeLaunch L, const int32_t* perm,
                                                     const int32_t* ctr) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= ctr[kCtrClosed]) return;
  const int q = perm[i];
  shp::Shape s1, s2;
  load_shape(L, L.pair_shapes[2 * q], s1);
  load_shape(L, L.pair_shapes[2 * q + 1], s2);
  const shp::Xf a = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
  const shp::Xf b = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
  shp::Contact c[4];
  int n = 0;
  const bool hit = shp::closed_form(s1, a, s2, b, L.contact != 0, c, n);
  write_contacts(L, q, c, shp::keep_contacts(hit, c, n, L.cap, L.contact != 0));
}

template <int kWaves>
__global__ void __launch_bounds__(kBlk, kWaves) shape_gjk(ShapeLaunch L, const int32_t* perm,
                                                          int32_t* ctr, int32_t* queue) {
  const int i = ctr[kCtrClosed] + blockIdx.x * kBlk + threadIdx.x;
  bool push = false;
  int q = -1;
  if (i < ctr[kCtrValid]) {
    q = perm[i];
    shp::Shape s1, s2;
    load_shape(L, L.pair_shapes[2 * q], s1);
    load_shape(L, L.pair_shapes[2 * q + 1], s2);
    const shp::Xf a = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
    const shp::Xf b = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
    shp::MDiff m;
    shp::make_mdiff(s1, a, s2, b, m);
    shp::GjkState g;
    const bool inside = shp::gjk_evaluate(g, m, shp::neg(shp::mk(1, 0, 0)), L.tol) ==
                        shp::kGjkInside;
    if (!inside) {
      L.counts[q] = 0;
    } else if (!L.contact) {
      L.counts[q] = 1;  // num_max_contacts >= 1
    } else {
      push = true;
    }
  }
  const unsigned long long mask = __ballot(push);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[kCtrQueue], __popcll(mask));
    base = __shfl(base, leader);
    if (push) queue[base + __popcll(mask & ((1ull << lane) - 1))] = q;
  }
}

template <int kWaves>
__global__ void __launch_bounds__(kBlk, kWaves) shape_epa(ShapeLaunch L, const int32_t* queue,
                                                          int32_t* ctr, shp::EpaMem* mem) {
  const int lane = threadIdx.x & 63;
  shp::EpaMem* my = mem + static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  const int total = ctr[kCtrQueue];
  while (true) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&ctr[kCtrGrab], 64);
    base = __shfl(base, 0);
    if (base >= total) break;
    const int i = base + lane;
    if (i < total) {
      const int q = queue[i];
      shp::Shape s1, s2;
      load_shape(L, L.pair_shapes[2 * q], s1);
      load_shape(L, L.pair_shapes[2 * q + 1], s2);
      const shp::Xf a = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
      const shp::Xf b = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
      shp::MDiff m;
      shp::make_mdiff(s1, a, s2, b, m);
      shp::GjkState g;
      shp::gjk_evaluate(g, m, shp::neg(shp::mk(1, 0, 0)), L.tol);
      const shp::Contact c = shp::epa_contact(my, g, m, a, L.tol);
      write_contacts(L, q, &c, 1);
    }
  }
}

}  // namespace

struct ShapeWorkspace {
  int device = 0;
  long long cap = 0;
  int32_t* perm = nullptr;
  int32_t* queue = nullptr;
  uint8_t* keys = nullptr;
  int32_t* ctr = nullptr;
  shp::EpaMem* epa = nullptr;
  int epa_blocks = 0;
  int epa_blocks_per_cu = 0;  // 0: as many as are resident
  int waves = 4;              // register budget of the GJK and EPA kernels (waves per SIMD)
  hipEvent_t done = nullptr;
  bool pending = false;       // done marks the last batch's end
};

ShapeWorkspace* shape_workspace_create(int device) {
  auto* w = new ShapeWorkspace();
  w->device = device;
  return w;
}

static void release_buffers(ShapeWorkspace* w) {
  (void)hipFree(w->perm);
  (void)hipFree(w->queue);
  (void)hipFree(w->keys);
  w->perm = w->queue = nullptr;
  w->keys = nullptr;
  w->cap = 0;
}

void shape_workspace_destroy(ShapeWorkspace* w) {
  if (!w) return;
  if (w->pending) (void)hipEventSynchronize(w->done);
  release_buffers(w);
  (void)hipFree(w->ctr);
  (void)hipFree(w->epa);
  if (w->done) (void)hipEventDestroy(w->done);
  delete w;
}

void shape_workspace_set_epa_blocks_per_cu(ShapeWorkspace* w, int blocks) {
  w->epa_blocks_per_cu = blocks;
}

void shape_workspace_set_waves(ShapeWorkspace* w, int waves) { w->waves = waves; }

namespace {

// GJK and EPA kernels for a register budget of 2, 3 or 4 waves per SIMD
struct ShapeKernels {
  const void* epa;
  void (*launch_gjk)(int, hipStream_t, const ShapeLaunch&, const int32_t*, int32_t*, int32_t*);
  void (*launch_epa)(int, hipStream_t, const ShapeLaunch&, const int32_t*, int32_t*,
                     shp::EpaMem*);
};

template <int W>
ShapeKernels kernels_for() {
  return {reinterpret_cast<const void*>(&shape_epa<W>),
          [](int grid, hipStream_t s, const ShapeLaunch& L, const int32_t* perm, int32_t* ctr,
             int32_t* queue) { shape_gjk<W><<<grid, kBlk, 0, s>>>(L, perm, ctr, queue); },
          [](int grid, hipStream_t s, const ShapeLaunch& L, const int32_t* queue, int32_t* ctr,
             shp::EpaMem* mem) { shape_epa<W><<<grid, kBlk, 0, s>>>(L, queue, ctr, mem); }};
}

ShapeKernels select_kernels(int waves) {
  switch (waves) {
    case 2:
      return kernels_for<2>();
    case 3:
      return kernels_for<3>();
    default:
      return kernels_for<4>();
  }
}

}  // namespace

size_t shape_epa_lane_bytes() { return sizeof(shp::EpaMem); }

hipError_t shape_collide_run(ShapeWorkspace* w, const ShapeLaunch& L, hipStream_t s) {
  if (L.num_pairs <= 0) return hipSuccess;
  hipError_t e;
  if (!w->done && (e = hipEventCreateWithFlags(&w->done, hipEventDisableTiming)) != hipSuccess)
    return e;
  if (!w->ctr && (e = hipMalloc(&w->ctr, sizeof(int32_t) * kNumCtr)) != hipSuccess) return e;
  // the EPA grid: resident blocks on every CU (or the set cap per CU)
  int cus = 0, resident = 0;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w->device)) !=
      hipSuccess)
    return e;
  const ShapeKernels K = select_kernels(w->waves);
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, K.epa, kBlk, 0)) !=
      hipSuccess)
    return e;
  resident = std::max(1, resident);
  if (w->epa_blocks_per_cu > 0) resident = std::min(resident, w->epa_blocks_per_cu);
  const int epa_blocks = cus * resident;
  const bool grow = L.num_pairs > w->cap;
  if ((grow || epa_blocks != w->epa_blocks) && w->pending) {
    if ((e = hipEventSynchronize(w->done)) != hipSuccess) return e;
    w->pending = false;
  }
  if (grow) {
    release_buffers(w);
    const long long cap = std::max<long long>(L.num_pairs, 1 << 16);
    if ((e = hipMalloc(&w->perm, sizeof(int32_t) * cap)) != hipSuccess ||
        (e = hipMalloc(&w->queue, sizeof(int32_t) * cap)) != hipSuccess ||
        (e = hipMalloc(&w->keys, cap)) != hipSuccess) {
      release_buffers(w);
      return e;
    }
    w->cap = cap;
  }
  if (epa_blocks != w->epa_blocks) {
    (void)hipFree(w->epa);
    w->epa = nullptr;
    w->epa_blocks = 0;
    if ((e = hipMalloc(&w->epa, sizeof(shp::EpaMem) * static_cast<size_t>(epa_blocks) * kBlk)) !=
        hipSuccess)
      return e;
    w->epa_blocks = epa_blocks;
  }
  // a batch on another stream may still be using the buffers
  if (w->pending && (e = hipStreamWaitEvent(s, w->done, 0)) != hipSuccess) return e;

  const int n = L.num_pairs;
  const int sort_grid = static_cast<int>((n + kBlk * kPerThread - 1) / (kBlk * kPerThread));
  const int grid = (n + kBlk - 1) / kBlk;
  shape_zero<<<1, 64, 0, s>>>(w->ctr);
  shape_classify<<<sort_grid, kBlk, 0, s>>>(L, w->keys, w->ctr);
  shape_scan<<<1, 64, 0, s>>>(w->ctr);
  shape_scatter<<<sort_grid, kBlk, 0, s>>>(n, w->keys, w->ctr, w->perm);
  shape_closed<<<grid, kBlk, 0, s>>>(L, w->perm, w->ctr);
  K.launch_gjk(grid, s, L, w->perm, w->ctr, w->queue);
  if (L.contact) K.launch_epa(w->epa_blocks, s, L, w->queue, w->ctr, w->epa);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipEventRecord(w->done, s)) != hipSuccess) return e;
  w->pending = true;
  return hipSuccess;
}

}  // namespace fclx
