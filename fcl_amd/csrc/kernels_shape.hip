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
//   3. EPA over the queue (GJK's final simplex per pair): a persistent grid
//      in which every lane runs its own pair one EPA loop pass at a time and
//      takes the next pair of its wavefront's share when done, so lanes do
//      not wait for the slowest pair of their wavefront; a lane keeps EPA's
//      faces and vertices in a contiguous block of device memory of its own.
// Same-address atomics are kept few: one per block per counter in the
// sorting and GJK kernels (the GJK kernel's on per-XCD counters), one per 64
// pairs in EPA.
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
  kCtrQueue,                     // pairs queued for EPA, per XCD sub-queue (8)
  kCtrGrab = kCtrQueue + 8,      // EPA queue entries handed to wavefronts
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
      if (L.passes) L.passes[2 * q] = L.passes[2 * q + 1] = 0;
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
  if (L.passes) L.passes[2 * q] = L.passes[2 * q + 1] = 0;
}

// The EPA queue is split into 8 sub-queues, one per XCD (blocks are
// dispatched to XCDs round-robin by index): a block reserves its entries
// with one atomic on its XCD's counter.  Sub-queue x holds the pairs of the
// blocks with index = x (mod 8), so sub_cap = ceil(grid / 8) * kBlk bounds
// it.
template <int kWaves>
__global__ void __launch_bounds__(kBlk, kWaves) shape_gjk(ShapeLaunch L, const int32_t* perm,
                                                          int32_t* ctr, shp::EpaSeed* seeds,
                                                          int sub_cap) {
  __shared__ int wave_count[kBlk / 64];
  __shared__ int block_base;
  const int i = ctr[kCtrClosed] + blockIdx.x * kBlk + threadIdx.x;
  bool push = false;
  int q = -1, passes = 0;
  shp::Gjk g;
  if (i < ctr[kCtrValid]) {
    q = perm[i];
    shp::Shape s1, s2;
    load_shape(L, L.pair_shapes[2 * q], s1);
    load_shape(L, L.pair_shapes[2 * q + 1], s2);
    const shp::Xf a = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
    const shp::Xf b = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
    shp::MDiff m;
    shp::make_mdiff(s1, a, s2, b, m);
    const bool inside = shp::gjk_evaluate(g, m, L.tol, &passes) == shp::kGjkInside;
    if (!inside) {
      L.counts[q] = 0;
    } else if (!L.contact) {
      L.counts[q] = 1;  // num_max_contacts >= 1
    } else {
      push = true;
    }
    if (L.passes && !push) {
      L.passes[2 * q] = passes;
      L.passes[2 * q + 1] = 0;
    }
  }
  if (!L.contact) return;  // uniform over the grid
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long mask = __ballot(push);
  if (lane == 0) wave_count[wave] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int k = 0; k < kBlk / 64; ++k) total += wave_count[k];
    block_base = total ? atomicAdd(&ctr[kCtrQueue + (blockIdx.x & 7)], total) : 0;
  }
  __syncthreads();
  if (push) {
    int off = block_base;
    for (int k = 0; k < wave; ++k) off += wave_count[k];
    off += __popcll(mask & ((1ull << lane) - 1));
    shp::EpaSeed& sd = seeds[static_cast<size_t>(blockIdx.x & 7) * sub_cap + off];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sd.w[k] = g.w[k];
      sd.d[k] = g.d[k];
    }
    sd.q = q;
    sd.rank = g.rank;
    sd.gjk_passes = passes;
  }
}

// Idle lanes of a wavefront take new pairs once this many are idle (or all).
constexpr int kRefill = 16;

// EPA over the queued pairs (the sub-queues, concatenated).  A wavefront
// takes 64 entries at a time with one atomic; every lane runs its own pair
// one loop pass per trip and takes the wavefront's next entry when done, so
// lanes do not wait for the slowest pair of their wavefront.
template <int kWaves>
__global__ void __launch_bounds__(kBlk, kWaves) shape_epa(ShapeLaunch L,
                                                          const shp::EpaSeed* seeds,
                                                          int32_t* ctr, int sub_cap,
                                                          shp::EpaLane* lanes,
                                                          shp::EpaWave* waves) {
  const int lane = threadIdx.x & 63;
  const size_t gid = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  int sub_end[8];  // running totals of the sub-queue sizes
  int total = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    total += ctr[kCtrQueue + x];
    sub_end[x] = total;
  }
  const int wid = static_cast<int>(gid / 64);
  int cursor = 0, stop = 0;  // the wavefront's entries not yet started
  bool exhausted = false;
  // the horizon walk's stack in LDS, one column per thread (a frame is
  // popped right after the walk returns from a child, at LDS latency)
  __shared__ int16_t xs_lds[shp::kNfEpa][kBlk];
  shp::Epa E;
  E.M = lanes + gid;
  E.S = waves + wid;
  E.lane = lane;
  E.xs = &xs_lds[0][threadIdx.x];
  E.xs_stride = kBlk;
  shp::MDiff m;
  int q = -1, gjk_passes = 0;
  bool active = false;
  while (true) {
    bool finish = false, started = false;
    const unsigned long long idle = __ballot(!active);
    const int nidle = __popcll(idle);
    if (cursor >= stop && !exhausted && (nidle == 64 || nidle >= kRefill)) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&ctr[kCtrGrab], 64);
      base = __shfl(base, 0);
      cursor = base < total ? base : total;
      stop = base + 64 < total ? base + 64 : total;
      exhausted = base + 64 >= total;
    }
    if (cursor < stop && (nidle == 64 || nidle >= kRefill)) {
      const int v = cursor + __popcll(idle & ((1ull << lane) - 1));
      cursor = cursor + nidle < stop ? cursor + nidle : stop;
      if (!active && v < stop) {
        int x = 0, before = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k)
          if (v >= sub_end[k]) {
            x = k + 1;
            before = sub_end[k];
          }
        const shp::EpaSeed& sd = seeds[static_cast<size_t>(x) * sub_cap + (v - before)];
        q = sd.q;
        gjk_passes = sd.gjk_passes;
        shp::Shape s1, s2;
        load_shape(L, L.pair_shapes[2 * q], s1);
        load_shape(L, L.pair_shapes[2 * q + 1], s2);
        shp::make_mdiff(s1, shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q)), s2,
                        shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q)), m);
        shp::Gjk g;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          g.w[k] = sd.w[k];
          g.d[k] = sd.d[k];
        }
        g.rank = sd.rank;
        active = shp::epa_start(E, g, m, L.tol);
        finish = !active;
        started = true;
      }
    }
    if (active && !started) {
      active = shp::epa_step(E, m, L.tol);
      finish = !active;
    }
    if (finish) {
      const shp::Contact c =
          shp::epa_finish(E, m, shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q)));
      write_contacts(L, q, &c, 1);
      if (L.passes) {
        L.passes[2 * q] = gjk_passes;
        L.passes[2 * q + 1] = E.fallback ? 0 : E.it;
      }
    }
    if (exhausted && cursor >= stop && !__ballot(active)) break;
  }
}

}  // namespace

struct ShapeWorkspace {
  int device = 0;
  long long cap = 0;
  int32_t* perm = nullptr;
  uint8_t* keys = nullptr;
  shp::EpaSeed* seeds = nullptr;  // EPA queue (allocated with enable_contact)
  long long seed_cap = 0;
  int32_t* ctr = nullptr;
  shp::EpaLane* epa = nullptr;
  shp::EpaWave* epa_waves = nullptr;
  int epa_blocks = 0;
  int epa_blocks_per_cu = 0;  // 0: as many as are resident
  int waves = 2;              // register budget of the GJK and EPA kernels (waves per SIMD)
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
  (void)hipFree(w->keys);
  (void)hipFree(w->seeds);
  w->perm = nullptr;
  w->keys = nullptr;
  w->seeds = nullptr;
  w->cap = 0;
  w->seed_cap = 0;
}

void shape_workspace_destroy(ShapeWorkspace* w) {
  if (!w) return;
  if (w->pending) (void)hipEventSynchronize(w->done);
  release_buffers(w);
  (void)hipFree(w->ctr);
  (void)hipFree(w->epa);
  (void)hipFree(w->epa_waves);
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
  void (*launch_gjk)(int, hipStream_t, const ShapeLaunch&, const int32_t*, int32_t*,
                     shp::EpaSeed*, int);
  void (*launch_epa)(int, hipStream_t, const ShapeLaunch&, const shp::EpaSeed*, int32_t*,
                     int, shp::EpaLane*, shp::EpaWave*);
};

template <int W>
ShapeKernels kernels_for() {
  return {reinterpret_cast<const void*>(&shape_epa<W>),
          [](int grid, hipStream_t s, const ShapeLaunch& L, const int32_t* perm, int32_t* ctr,
             shp::EpaSeed* seeds, int sub_cap) {
            shape_gjk<W><<<grid, kBlk, 0, s>>>(L, perm, ctr, seeds, sub_cap);
          },
          [](int grid, hipStream_t s, const ShapeLaunch& L, const shp::EpaSeed* seeds,
             int32_t* ctr, int sub_cap, shp::EpaLane* lanes, shp::EpaWave* waves) {
            shape_epa<W><<<grid, kBlk, 0, s>>>(L, seeds, ctr, sub_cap, lanes, waves);
          }};
}

ShapeKernels select_kernels(int waves) {
  switch (waves) {
    case 3:
      return kernels_for<3>();
    case 4:
      return kernels_for<4>();
    default:
      return kernels_for<2>();
  }
}

}  // namespace

size_t shape_epa_lane_bytes() { return sizeof(shp::EpaLane) + sizeof(shp::EpaWave) / 64; }

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
  const int n = L.num_pairs;
  const int grid = (n + kBlk - 1) / kBlk;
  const int sub_cap = (grid + 7) / 8 * kBlk;
  const bool grow = n > w->cap || (L.contact && 8LL * sub_cap > w->seed_cap);
  if ((grow || epa_blocks != w->epa_blocks) && w->pending) {
    if ((e = hipEventSynchronize(w->done)) != hipSuccess) return e;
    w->pending = false;
  }
  if (grow) {
    release_buffers(w);
    const long long cap = std::max<long long>(n, 1 << 16);
    const long long seed_cap = L.contact ? 8LL * sub_cap : 0;
    if ((e = hipMalloc(&w->perm, sizeof(int32_t) * cap)) != hipSuccess ||
        (e = hipMalloc(&w->keys, cap)) != hipSuccess ||
        (seed_cap && (e = hipMalloc(&w->seeds, sizeof(shp::EpaSeed) * seed_cap)) != hipSuccess)) {
      release_buffers(w);
      return e;
    }
    w->cap = cap;
    w->seed_cap = seed_cap;
  }
  if (epa_blocks != w->epa_blocks) {
    (void)hipFree(w->epa);
    (void)hipFree(w->epa_waves);
    w->epa = nullptr;
    w->epa_waves = nullptr;
    w->epa_blocks = 0;
    const size_t lanes = static_cast<size_t>(epa_blocks) * kBlk;
    if ((e = hipMalloc(&w->epa, sizeof(shp::EpaLane) * lanes)) != hipSuccess ||
        (e = hipMalloc(&w->epa_waves, sizeof(shp::EpaWave) * (lanes / 64))) != hipSuccess)
      return e;
    w->epa_blocks = epa_blocks;
  }
  // a batch on another stream may still be using the buffers
  if (w->pending && (e = hipStreamWaitEvent(s, w->done, 0)) != hipSuccess) return e;

  const int sort_grid = static_cast<int>((n + kBlk * kPerThread - 1) / (kBlk * kPerThread));
  shape_zero<<<1, 64, 0, s>>>(w->ctr);
  shape_classify<<<sort_grid, kBlk, 0, s>>>(L, w->keys, w->ctr);
  shape_scan<<<1, 64, 0, s>>>(w->ctr);
  shape_scatter<<<sort_grid, kBlk, 0, s>>>(n, w->keys, w->ctr, w->perm);
  shape_closed<<<grid, kBlk, 0, s>>>(L, w->perm, w->ctr);
  K.launch_gjk(grid, s, L, w->perm, w->ctr, w->seeds, sub_cap);
  if (L.contact)
    K.launch_epa(w->epa_blocks, s, L, w->seeds, w->ctr, sub_cap, w->epa, w->epa_waves);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipEventRecord(w->done, s)) != hipSuccess) return e;
  w->pending = true;
  return hipSuccess;
}

}  // namespace fclx
