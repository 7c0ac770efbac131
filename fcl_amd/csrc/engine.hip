// Fast collide engine.
//
// Exhaustive contacts (num_max_contacts > kOrderedCapMax):
//   1. setup: per-query FP64 and FP32 relative transforms, margins, and the
//      root BVTT task (q, 0, 0) of every valid query;
//   2. breadth-first front expansion: one lane per BVTT task, FP32
//      conservative OBB cull (FP64 only in the uncertain band), children
//      appended with one atomic per wavefront (ballot + popcount prefix);
//      leaf-pair candidates go to a separate buffer.  This splits a large
//      query over many wavefronts before the depth-first phase starts;
//   3. leaf kernel: FP64 triangle SAT over the candidates of phase 2;
//   4. persistent depth-first phase: every wavefront owns a LIFO task stack
//      in HBM; each step pops up to 64 tasks (one per lane), tests them, and
//      pushes the children of overlapping internal pairs with one prefix sum.
//      Leaf candidates collect in an LDS queue and run 64 at a time, so the
//      FP64 triangle tests execute with the wavefront converged.  Idle
//      wavefronts refill from the front of phase 2;
//   5. any query whose contact count reached the cap is re-run by the plain
//      depth-first kernel and its records replaced (the reference stops at
//      the cap-th contact in depth-first order).
// Without early stop the set of BVTT nodes visited is order-independent, so
// the contact set and the per-query test counts equal the reference's.
//
// Small caps (<= kOrderedCapMax, e.g. first contact): per-lane depth-first
// traversal in the reference's order with FP32 culling, lanes refilling with
// new queries as they finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include <vector>

#include "device/cull32.h"
#include "device/exact64.h"
#include "engine.h"

namespace fclx {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Per-wavefront LDS queues of the depth-first kernel, flushed half at a
// time to the deferred buffers.
constexpr int kLeafQ = 256;       // leaf pairs
constexpr int kUncQ = 128;        // uncertain tasks
constexpr int kQueueCap[2] = {kLeafQ, kUncQ};
// Each deferred buffer is split into kSub sub-buffers, one per XCD under the
// hardware's round-robin block placement (blockIdx % kSub), each with its
// reservation counter on a cache line of its own.
constexpr int kSub = 8;
constexpr unsigned long long kNoTask = ~0ULL;  // padding the consumers skip

// Device counters of the exhaustive path (ws->counters[...]).
enum : int {
  kCtrSetupOut = 0,   // root tasks
  kCtrExpandOut = 1,  // breadth-first level output
  kCtrLeafc = 2,      // breadth-first leaf candidates
  kCtrHead = 3,       // depth-first refill cursor
  kCtrFlagged = 4,    // queries that reached the cap
  // 5..7: cap fix-up
  kCtrWork = 8,       // 8 .. 8+kWork*: work counters
  // the counters the depth-first kernel touches often sit on 128-B lines
  // of their own (16 counters apart), so atomics on one do not queue
  // behind another
  kCtrRoundBegin = 128,
  kCtrResv = 128,     // reservation of sub-buffer x of buffer w: kCtrResv + 16 (w kSub + x)
  kCtrDump = kCtrResv + 16 * 2 * kSub,  // dump counter of buffer w: kCtrDump + 16 w
  kCtrStop = kCtrDump + 32,             // polled by every wavefront
  kCtrIdle = kCtrStop + 16,
  kCtrNext = kCtrIdle + 1,              // next round's front
  kCtrRoundEnd = kCtrNext + 1,          // one past the last counter a round resets
  kNumCounters = kCtrRoundEnd + 14
};


// FP64 relative pose of a query plus its meshes' triangle-major vertex
// arrays: everything a triangle test needs, in one 128-B line.
struct alignas(128) QueryXf64 {
  double R[9];
  double T[3];
  const double* tv1;
  const double* tv2;
};

// task = (query:22 | b1:21 | b2:21)
__device__ __forceinline__ unsigned long long pack_task(int q, int b1, int b2) {
  return (static_cast<unsigned long long>(q) << 42) |
         (static_cast<unsigned long long>(b1) << 21) | static_cast<unsigned long long>(b2);
}
__device__ __forceinline__ void unpack_task(unsigned long long t, int& q, int& b1, int& b2) {
  q = static_cast<int>(t >> 42);
  b1 = static_cast<int>((t >> 21) & 0x1FFFFF);
  b2 = static_cast<int>(t & 0x1FFFFF);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long lanes_below() {
  return (1ULL << lane_id()) - 1ULL;
}

// One atomic per wavefront: returns this lane's slot among the lanes with
// pred set (slot values base .. base+popcount-1, in lane order).
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long* counter,
                                                           bool pred, int per_lane) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0) return 0;
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader)
    base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)) * per_lane);
  base = __shfl(base, leader);
  return base + static_cast<unsigned long long>(__popcll(mask & lanes_below())) * per_lane;
}

__device__ __forceinline__ bool query_ok(const MeshDev* meshes, int n, int i1, int i2) {
  if (i1 < 0 || i2 < 0 || i1 >= n || i2 >= n) return false;
  const MeshDev& a = meshes[i1];
  const MeshDev& b = meshes[i2];
  return a.num_bvs > 0 && b.num_bvs > 0 && a.bv_type == 0 && b.bv_type == 0;
}

__device__ __forceinline__ V3 vtx(const MeshDev& m, int i) {
  const double* p = m.verts + 3 * static_cast<size_t>(i);
  return v3(p[0], p[1], p[2]);
}

// R = R1^T R2, T = R1^T (t2 - t1)
__device__ __forceinline__ void rel_pose64(const double* tf1, const double* tf2, double* R,
                                           double* T) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = (tf1[i] * tf2[j] + tf1[3 + i] * tf2[3 + j]) + tf1[6 + i] * tf2[6 + j];
  const V3 d = v3(tf2[9] - tf1[9], tf2[10] - tf1[10], tf2[11] - tf1[11]);
  T[0] = dot(col(tf1, 0), d);
  T[1] = dot(col(tf1, 1), d);
  T[2] = dot(col(tf1, 2), d);
}

__device__ __forceinline__ void make_xf32(const double* R, const double* T, double cmax1,
                                          double cmax2, QueryXf32& X) {
#pragma unroll
  for (int i = 0; i < 9; ++i) X.R[i] = static_cast<float>(R[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) X.T[i] = static_cast<float>(T[i]);
  const double st = cmax1 + (fabs(T[0]) + fabs(T[1]) + fabs(T[2])) + cmax2;
  // margin rounded up: the float conversion may round down by <= 2^-24
  X.ms = static_cast<float>(static_cast<double>(kCullKS) * static_cast<double>(kCullU) * st *
                            (1.0 + 1.0 / 1048576.0));
}

// Node and size arrays of a query's two meshes, read in parallel with the
// query's transform so no mesh-descriptor load sits on the critical path.
struct QueryPtr {
  const ObbNode32* n1;
  const ObbNode32* n2;
  const double* s1;
  const double* s2;
};

// Up to kSub + 1 ranges of tasks read as one sequence (the sub-buffers of a
// deferred buffer and its dump region); consumers skip kNoTask entries.
struct TaskRanges {
  const unsigned long long* p[kSub + 1];
  long long start[kSub + 2];  // start[r] = entries before range r; start[count] = total
  int count;
};

__device__ __forceinline__ unsigned long long range_task(const TaskRanges& R, long long i) {
  int r = 0;
  while (r + 1 < R.count && i >= R.start[r + 1]) ++r;
  return R.p[r][i - R.start[r]];
}

struct EngineArgs {
  CollideLaunch L;
  QueryXf32* qx32;
  QueryPtr* qptr;
  QueryXf64* qx64;
  int32_t* qflag;
  unsigned long long* nflag;
  unsigned long long* front_in;
  unsigned long long* front_out;
  unsigned long long* n_out;
  unsigned long long* leafc;
  unsigned long long* n_leafc;
  long long n_in;
  unsigned long long* head;       // DFS refill cursor
  unsigned long long* stacks;
  int stack_cap;
  int stack_top;                  // DFS stack entries kept in LDS (128..kStackTop)
  unsigned long long* work;  // kWork* counters
  unsigned long long* leaf_work;  // work counter the leaf kernel adds to
  TaskRanges in;                  // input of the leaf and uncertain kernels
  unsigned long long* wave_prof;  // profiling: (start, end, bv tests) per DFS wavefront, or null
  // depth-first rounds: tasks that need FP64 work are deferred to two
  // buffers (0: overlapping leaf pairs as (q, tri1, tri2), 1: tasks the
  // FP32 cull left uncertain).  Buffer w holds kSub sub-buffers of
  // defer_sub[w] entries; a flush reserves room in its block's sub-buffer
  // and, when that does not fit, pads the part inside it with kNoTask,
  // fails, and stops the round.  Per-wavefront leftovers go to the dump
  // region after the sub-buffers.
  unsigned long long* defer[2];
  long long defer_sub[2];
  unsigned long long* ctr;  // the engine's counter array (kCtr* indices)
  unsigned long long* stop;   // nonzero: every wavefront dumps its state and exits
  unsigned long long* idle;   // wavefronts that ran out of work
  unsigned long long idle_limit;  // idle count that stops the round (0: never)
  unsigned stop_poll;             // depth-first steps between polls of the stop flag
};

enum : int { kWorkExpandBV = 0, kWorkLeafKernel, kWorkDfsBV, kWorkDfsLeaf, kWorkFallback,
             kWorkDfsInline, kWorkDfsSteps };

// add popcount(pred) over the wavefront to *counter with one atomic
__device__ __forceinline__ void wave_count(unsigned long long* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m != 0 && lane_id() == __ffsll(static_cast<long long>(m)) - 1)
    atomicAdd(counter, static_cast<unsigned long long>(__popcll(m)));
}

// ---------------------------------------------------------------- setup
__global__ void __launch_bounds__(kBlock) engine_setup(EngineArgs A) {
  const CollideLaunch& L = A.L;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = false;
  if (q < L.num_queries) {
    const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
    valid = query_ok(L.meshes, L.num_meshes, i1, i2);
    L.counts[q] = 0;
    A.qflag[q] = 0;
    if (L.bv_tests) L.bv_tests[q] = 0;
    if (L.leaf_tests) L.leaf_tests[q] = 0;
    QueryXf32 X;
    if (valid) {
      QueryXf64 x;
      rel_pose64(L.tf1 + 12 * static_cast<size_t>(q), L.tf2 + 12 * static_cast<size_t>(q), x.R,
                 x.T);
      x.tv1 = L.meshes[i1].tri_verts;
      x.tv2 = L.meshes[i2].tri_verts;
      A.qx64[q] = x;
      make_xf32(x.R, x.T, L.meshes[i1].cmax, L.meshes[i2].cmax, X);
      X.m1 = i1;
      X.m2 = i2;
      X.valid = 1;
      A.qptr[q] = QueryPtr{L.meshes[i1].obb32, L.meshes[i2].obb32, L.meshes[i1].size64,
                           L.meshes[i2].size64};
    } else {
      X = QueryXf32{};
      X.valid = 0;
      atomicOr(L.status, kStatusBadQuery);
    }
    A.qx32[q] = X;
  }
  const unsigned long long slot = wave_reserve(A.n_out, valid, 1);
  if (valid) A.front_out[slot] = pack_task(q, 0, 0);
}

// Outcome of testing one BVTT task.
struct Visit {
  bool push = false;   // overlapping internal pair: children c0, c1
  bool leafc = false;  // overlapping leaf pair: triangle test pending, c0 = (q, tri1, tri2)
  bool fallback = false;
  unsigned long long c0 = 0, c1 = 0;
};

// BV test of task (q,b1,b2): FP32 cull, FP64 only when unsure; then the
// reference's descent rule (firstOverSecond) for the children.  The query
// record, both nodes and both sizes are loaded before any is used.
template <bool kCount = true>
__device__ __forceinline__ Visit visit_task(const EngineArgs& A, unsigned long long task) {
  int q, b1, b2;
  unpack_task(task, q, b1, b2);
  const QueryPtr P = A.qptr[q];
  const QueryXf32& X = A.qx32[q];
  const ObbNode32& n1 = P.n1[b1];
  const ObbNode32& n2 = P.n2[b2];
  if (kCount && A.L.bv_tests)
    atomicAdd(reinterpret_cast<unsigned long long*>(&A.L.bv_tests[q]), 1ULL);
  Visit v;
  int r = obb_cull32(X, n1, n2);
  if (r == kCullUnsure) {
    v.fallback = true;
    const QueryXf64& x = A.qx64[q];
    M3 R;
#pragma unroll
    for (int i = 0; i < 9; ++i) R.m[i] = x.R[i];
    const MeshDev& m1 = A.L.meshes[X.m1];
    const MeshDev& m2 = A.L.meshes[X.m2];
    r = obb_overlap64(R, v3(x.T[0], x.T[1], x.T[2]), m1.obb64[b1], m2.obb64[b2]) ? kCullOverlap
                                                                                   : kCullDisjoint;
  }
  if (r == kCullOverlap) {
    const int fc1 = n1.first_child, fc2 = n2.first_child;
    if (fc1 < 0 && fc2 < 0) {
      v.leafc = true;
      v.c0 = pack_task(q, -(fc1 + 1), -(fc2 + 1));
    } else {
      v.push = true;
      if ((fc2 < 0) || (fc1 >= 0 && (P.s1[b1] > P.s2[b2]))) {
        v.c0 = pack_task(q, fc1, b2);
        v.c1 = pack_task(q, fc1 + 1, b2);
      } else {
        v.c0 = pack_task(q, b1, fc2);
        v.c1 = pack_task(q, b1, fc2 + 1);
      }
    }
  }
  return v;
}

// Per-query contact counts of the lanes that hit: one atomic per distinct
// query among them (the candidates of one flush nearly always share their
// query); flags a query once its count reaches the cap.  Call with only the
// hitting lanes active.
__device__ __forceinline__ void count_hits(const EngineArgs& A, int q) {
  const CollideLaunch& L = A.L;
  unsigned long long todo = __ballot(true);
  while (todo != 0) {
    const int leader = __ffsll(static_cast<long long>(todo)) - 1;
    const int ql = __shfl(q, leader);
    const unsigned long long same = __ballot(q == ql);
    if (lane_id() == leader) {
      const int k = __popcll(same);
      const int old = atomicAdd(&L.counts[ql], k);
      if (old + k >= L.cap) {
        if (atomicExch(&A.qflag[ql], 1) == 0) atomicAdd(A.nflag, 1ULL);
      }
    }
    todo &= ~same;
  }
}

// Record intersecting pairs (one output reservation per wavefront) and
// count them; for kernels that see few hits.
__device__ __forceinline__ void record_hits(const EngineArgs& A, bool hit, int q, int id1,
                                            int id2) {
  const CollideLaunch& L = A.L;
  const unsigned long long slot = wave_reserve(reinterpret_cast<unsigned long long*>(L.num_pairs),
                                               hit, 1);
  if (!hit) return;
  if (static_cast<long long>(slot) < L.pair_capacity) {
    L.pairs[3 * slot] = q;
    L.pairs[3 * slot + 1] = id1;
    L.pairs[3 * slot + 2] = id2;
  } else if (L.pairs) {  // no buffer: counts only
    atomicOr(L.status, kStatusPairOverflow);
  }
  count_hits(A, q);
}

// Triangle test of a leaf candidate (q, tri1, tri2).
__device__ __forceinline__ bool leaf_hit(const EngineArgs& A, unsigned long long task, int& q,
                                         int& id1, int& id2) {
  unpack_task(task, q, id1, id2);
  const QueryXf64& x = A.qx64[q];
  const double* p = x.tv1 + kTriVertsStride * static_cast<size_t>(id1);
  const double* r = x.tv2 + kTriVertsStride * static_cast<size_t>(id2);
  const V3 P1 = v3(p[0], p[1], p[2]), P2 = v3(p[3], p[4], p[5]), P3 = v3(p[6], p[7], p[8]);
  const V3 Q1 = v3(r[0], r[1], r[2]), Q2 = v3(r[3], r[4], r[5]), Q3 = v3(r[6], r[7], r[8]);
  M3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = x.R[i];
  if (A.L.leaf_tests) atomicAdd(reinterpret_cast<unsigned long long*>(&A.L.leaf_tests[q]), 1ULL);
  return tri_intersect64(P1, P2, P3, Q1, Q2, Q3, R, v3(x.T[0], x.T[1], x.T[2]));
}

// ------------------------------------------------------- front expansion
__global__ void __launch_bounds__(kBlock) engine_expand(EngineArgs A) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = i < A.n_in;
  unsigned long long task = 0;
  Visit v;
  if (active) {
    task = A.front_in[i];
    v = visit_task(A, task);
  }
  wave_count(A.work + kWorkExpandBV, active);
  wave_count(A.work + kWorkFallback, v.fallback);
  const unsigned long long so = wave_reserve(A.n_out, v.push, 2);
  if (v.push) {
    A.front_out[so] = v.c0;
    A.front_out[so + 1] = v.c1;
  }
  const unsigned long long sl = wave_reserve(A.n_leafc, v.leafc, 1);
  if (v.leafc) A.leafc[sl] = v.c0;
}

// Triangle tests of leaf candidates leafc[0, n_in).  Persistent blocks walk
// 256-candidate tiles; hits are staged in LDS and written out in runs of
// up to kLeafStage records with one reservation on the global record
// counter, which every block shares (same-address atomics serialise in L2,
// so one per wavefront would throttle the whole kernel).
constexpr int kLeafStage = 2048;
__global__ void __launch_bounds__(kBlock) engine_leaves(EngineArgs A) {
  __shared__ int32_t stage[3 * kLeafStage];
  __shared__ unsigned n_stage;
  __shared__ unsigned long long base;
  const CollideLaunch& L = A.L;
  if (threadIdx.x == 0) n_stage = 0;
  __syncthreads();
  auto flush = [&]() {  // block-uniform call
    __syncthreads();
    const unsigned n = n_stage;
    if (threadIdx.x == 0 && n > 0)
      base = atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs),
                       static_cast<unsigned long long>(n));
    __syncthreads();
    bool over = false;
    for (unsigned r = threadIdx.x; r < n; r += kBlock) {
      const long long o = static_cast<long long>(base + r);
      if (o < L.pair_capacity) {
        L.pairs[3 * o] = stage[3 * r];
        L.pairs[3 * o + 1] = stage[3 * r + 1];
        L.pairs[3 * o + 2] = stage[3 * r + 2];
      } else {
        over = true;
      }
    }
    if (over && L.pairs) atomicOr(L.status, kStatusPairOverflow);
    __syncthreads();
    if (threadIdx.x == 0) n_stage = 0;
    __syncthreads();
  };
  const long long total = A.in.start[A.in.count];
  const long long tiles = (total + kBlock - 1) / kBlock;
  unsigned long long tests = 0;  // this wavefront's triangle tests
  for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const long long i = t * kBlock + threadIdx.x;
    bool hit = false;
    int q = 0, id1 = 0, id2 = 0;
    const unsigned long long task = i < total ? range_task(A.in, i) : kNoTask;
    if (task != kNoTask) hit = leaf_hit(A, task, q, id1, id2);
    tests += __popcll(__ballot(task != kNoTask));
    const unsigned long long m = __ballot(hit);
    unsigned off = 0;
    if (lane_id() == 0 && m != 0) off = atomicAdd(&n_stage, static_cast<unsigned>(__popcll(m)));
    off = __shfl(off, 0);
    if (hit) {
      const unsigned r = off + static_cast<unsigned>(__popcll(m & lanes_below()));
      stage[3 * r] = q;
      stage[3 * r + 1] = id1;
      stage[3 * r + 2] = id2;
      count_hits(A, q);
    }
    __syncthreads();
    const bool full = n_stage > kLeafStage - kBlock;
    __syncthreads();  // every thread has read n_stage before the next tile adds to it
    if (full) flush();
  }
  flush();
  if (lane_id() == 0 && tests) atomicAdd(A.leaf_work, tests);
}

// Deferred uncertain tasks: the FP64 test decides (the FP32 pass already
// counted the BV test); overlapping leaf pairs run their triangle test here,
// children of overlapping internal pairs go to the next round's front.
__global__ void __launch_bounds__(kBlock) engine_uncertain(EngineArgs A) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long task = i < A.in.start[A.in.count] ? range_task(A.in, i) : kNoTask;
  Visit v;
  if (task != kNoTask) v = visit_task<false>(A, task);
  const unsigned long long so = wave_reserve(A.n_out, v.push, 2);
  if (v.push) {
    A.front_out[so] = v.c0;
    A.front_out[so + 1] = v.c1;
  }
  bool hit = false;
  int q = 0, id1 = 0, id2 = 0;
  if (v.leafc) hit = leaf_hit(A, v.c0, q, id1, id2);
  wave_count(A.leaf_work, v.leafc);
  record_hits(A, hit, q, id1, id2);
}

// ----------------------------------------------- persistent depth-first
// firstOverSecond's size comparison (bvh_collision_traversal_node-inl.h:
// 77-89; OBBRSS::size = |OBB extents|^2, OBB-inl.h:206-209) from the FP32
// extents already loaded with the nodes.  Each FP32 sum of squares is
// within 5u (u = 2^-24) of the FP64 size, relative; outside a 50u band the
// FP32 order is the FP64 order, inside it (ties included) and for sizes
// near the FP32 underflow the FP64 sizes decide.  Saves the two scattered
// 8-B size loads of almost every descent.
__device__ __forceinline__ bool first_larger(const ObbNode32& n1, const ObbNode32& n2,
                                             const double* s1, const double* s2) {
  const float a = fmaf(n1.extent[2], n1.extent[2],
                       fmaf(n1.extent[1], n1.extent[1], n1.extent[0] * n1.extent[0]));
  const float b = fmaf(n2.extent[2], n2.extent[2],
                       fmaf(n2.extent[1], n2.extent[1], n2.extent[0] * n2.extent[0]));
  constexpr float kBand = 50.0f * 5.9604644775390625e-08f;
  if (a > 1e-30f && b > 1e-30f) {
    if (a > b * (1.0f + kBand)) return true;
    if (a < b * (1.0f - kBand)) return false;
  }
  return *s1 > *s2;
}

// Outcome of the FP32 pass over one task.
enum : int { kStepNone = 0, kStepPush = 1, kStepLeaf = 2, kStepUncertain = 3 };

// A 64-B node record in four 16-B loads issued together (its child link
// with the rest, not in a later load once the cull has passed).
__device__ __forceinline__ ObbNode32 load_node32(const ObbNode32* p) {
  const uint4* w = reinterpret_cast<const uint4*>(p);
  const uint4 a = w[0], b = w[1], c = w[2], d = w[3];
  ObbNode32 n;
  uint4* o = reinterpret_cast<uint4*>(&n);
  o[0] = a;
  o[1] = b;
  o[2] = c;
  o[3] = d;
  return n;
}

// The query's 64-B FP32 transform record, likewise in four 16-B loads.
__device__ __forceinline__ QueryXf32 load_xf32(const QueryXf32* p) {
  static_assert(sizeof(QueryXf32) == 64, "four 16-B words");
  const uint4* w = reinterpret_cast<const uint4*>(p);
  const uint4 a = w[0], b = w[1], c = w[2], d = w[3];
  QueryXf32 x;
  uint4* o = reinterpret_cast<uint4*>(&x);
  o[0] = a;
  o[1] = b;
  o[2] = c;
  o[3] = d;
  return x;
}

// FP32-only BV test of task (q,b1,b2) and the descent rule; nothing FP64 is
// inlined here, which keeps the depth-first kernel's register count low.
// kStats: per-query BV-test counts are requested (the branch and its
// pointer are compiled out of the kernel that runs without them).
template <bool kStats>
__device__ __forceinline__ int step32(const EngineArgs& A, unsigned long long task,
                                      unsigned long long& c0, unsigned long long& c1) {
  int q, b1, b2;
  unpack_task(task, q, b1, b2);
  const QueryPtr P = A.qptr[q];
  const QueryXf32 X = load_xf32(A.qx32 + q);
  const ObbNode32 n1 = load_node32(P.n1 + b1);
  const ObbNode32 n2 = load_node32(P.n2 + b2);
  if constexpr (kStats) {
    if (A.L.bv_tests) atomicAdd(reinterpret_cast<unsigned long long*>(&A.L.bv_tests[q]), 1ULL);
  }
  const int r = obb_cull32(X, n1, n2);
  if (r == kCullDisjoint) return kStepNone;
  if (r == kCullUnsure) return kStepUncertain;
  const int fc1 = n1.first_child, fc2 = n2.first_child;
  if (fc1 < 0 && fc2 < 0) {
    c0 = pack_task(q, -(fc1 + 1), -(fc2 + 1));  // leaf candidate: (q, tri1, tri2)
    return kStepLeaf;
  }
  if ((fc2 < 0) || (fc1 >= 0 && first_larger(n1, n2, P.s1 + b1, P.s2 + b2))) {
    c0 = pack_task(q, fc1, b2);
    c1 = pack_task(q, fc1 + 1, b2);
  } else {
    c0 = pack_task(q, b1, fc2);
    c1 = pack_task(q, b1, fc2 + 1);
  }
  return kStepPush;
}

// Move the top k entries of a wavefront's LDS queue to sub-buffer x of
// deferred buffer w.  Returns false (and moves nothing) when the
// sub-buffer has no room for them.
__device__ __forceinline__ bool flush_deferred(const EngineArgs& A, int w, int x,
                                               const unsigned long long* q, int& n, int k) {
  unsigned long long s = 0;
  if (lane_id() == 0)
    s = atomicAdd(A.ctr + kCtrResv + 16 * (w * kSub + x), static_cast<unsigned long long>(k));
  s = __shfl(s, 0);
  const long long cap = A.defer_sub[w];
  unsigned long long* dst = A.defer[w] + static_cast<size_t>(x) * cap;
  if (static_cast<long long>(s + k) > cap) {
    for (long long i = static_cast<long long>(s) + lane_id(); i < cap; i += kWave)
      dst[i] = kNoTask;
    return false;
  }
  for (int i = lane_id(); i < k; i += kWave) dst[s + i] = q[n - k + i];
  __builtin_amdgcn_wave_barrier();
  n -= k;
  return true;
}

// A stopped or finishing wavefront's queue leftovers, into the dump region
// after the sub-buffers (room for a full queue per wavefront).
__device__ __forceinline__ void dump_deferred(const EngineArgs& A, int w,
                                              const unsigned long long* q, int n) {
  if (n == 0) return;
  unsigned long long s = 0;
  if (lane_id() == 0) s = atomicAdd(A.ctr + kCtrDump + 16 * w, static_cast<unsigned long long>(n));
  s = __shfl(s, 0);
  unsigned long long* dst = A.defer[w] + static_cast<size_t>(kSub) * A.defer_sub[w];
  for (int i = lane_id(); i < n; i += kWave) dst[s + i] = q[i];
}

// One depth-first round.  Every wavefront owns a LIFO stack; each step pops
// up to 64 tasks (one per lane), runs the FP32 test, pushes the children of
// overlapping internal pairs with one prefix sum and queues overlapping leaf
// pairs and uncertain tasks in LDS.  Idle wavefronts refill from the front.
// The round stops early when a deferred buffer is full or when idle_limit
// wavefronts have run out of work; then every wavefront appends its stack to
// front_out and its queue leftovers to the dump regions, and the host runs
// the FP64 kernels and the next round.
//
// kInline = false: the stack lives in device memory, and leaf pairs and
// uncertain tasks move 64 at a time to the deferred buffers for the FP64
// kernels.
//
// kInline = true: the wavefront runs the FP64 triangle tests of its leaf
// queue itself, 64 at a time, while the other wavefronts of the CU wait on
// their node loads; hits are staged in LDS and written out in runs of up to
// kHitStage records with one reservation on the record counter (a
// reservation per step would serialise on that counter).  Only the queue's
// remainder at the end of a round is deferred to engine_leaves.  The top
// of the stack, up to A.stack_top <= kStackTop entries, lives in LDS (a ring
// indexed by stack position): a pop reads LDS instead of waiting on device
// memory, and only the deep part of the stack is written out (down to half
// the limit at a time, read back when the LDS part runs low).
constexpr int kHitStage = 256;
constexpr int kStackTop = 512;
template <int kWavesPerSimd, bool kInline, bool kStats = true>
__global__ void __launch_bounds__(kBlock, kWavesPerSimd) engine_dfs(EngineArgs A) {
  // with the triangle tests inline a queue never holds two full steps
  constexpr int kLq = kInline ? 2 * kWave : kLeafQ;
  constexpr int kTop = kInline ? kStackTop : 1;
  __shared__ unsigned long long lq_all[kWavesPerBlock][kLq];
  __shared__ unsigned long long uq_all[kWavesPerBlock][kUncQ];
  __shared__ int32_t hits_all[kInline ? kWavesPerBlock : 1][kInline ? 3 * kHitStage : 1];
  __shared__ unsigned long long top_all[kInline ? kWavesPerBlock : 1][kTop];
  const int wave = threadIdx.x / kWave;
  const int lane = lane_id();
  const int xcd = static_cast<int>(blockIdx.x % kSub);
  unsigned long long* lq = lq_all[wave];
  unsigned long long* uq = uq_all[wave];
  unsigned long long* stack =
      A.stacks + (static_cast<size_t>(blockIdx.x) * kWavesPerBlock + wave) * A.stack_cap;
  unsigned long long* top = top_all[kInline ? wave : 0];
  // stack positions [0, base) are in device memory, [base, sp) in LDS
  // (kInline; otherwise base stays 0 and every position is in device memory)
  int sp = 0, base = 0;       // wavefront-uniform
  int nl = 0, nu = 0;         // queue fill, wavefront-uniform
  bool input_left = true;
  bool stopped = false;
  unsigned long long n_bv = 0, n_unc = 0;  // wavefront-uniform
  unsigned long long n_leaf = 0;           // inline triangle tests
  unsigned long long n_steps = 0;          // BV-test steps (n_bv / n_steps: lanes used)
  int nh = 0;                              // staged hit records (kInline)
  int32_t* hits = hits_all[kInline ? wave : 0];
  const CollideLaunch& L = A.L;
  auto slot = [&](int pos) -> unsigned long long& {
    if constexpr (kInline) return top[pos & (kTop - 1)];
    else return stack[pos];
  };
  // staged records to the caller's buffer: one reservation for all of them
  auto flush_hits = [&]() {
    if constexpr (kInline) {
      if (nh == 0) return;
      unsigned long long b = 0;
      if (lane == 0)
        b = atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs),
                      static_cast<unsigned long long>(nh));
      b = __shfl(b, 0);
      bool over = false;
      for (int r = lane; r < nh; r += kWave) {
        const long long o = static_cast<long long>(b) + r;
        if (o < L.pair_capacity) {
          L.pairs[3 * o] = hits[3 * r];
          L.pairs[3 * o + 1] = hits[3 * r + 1];
          L.pairs[3 * o + 2] = hits[3 * r + 2];
        } else {
          over = true;
        }
      }
      if (__ballot(over) != 0 && lane == 0 && L.pairs) atomicOr(L.status, kStatusPairOverflow);
      nh = 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  };
  const unsigned long long t_start = A.wave_prof ? wall_clock64() : 0;
  unsigned step = 0;

  while (true) {
    if ((++step % A.stop_poll) == 0) {
      const int stop_seen =
          __hip_atomic_load(A.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 1 : 0;
      if (__builtin_amdgcn_readfirstlane(stop_seen)) {  // one decision per wavefront
        stopped = true;
        break;
      }
    }
    if constexpr (kInline) {
      // a full step of queued leaf pairs: their FP64 triangle tests
      if (nl >= kWave) {
        nl -= kWave;
        int q = 0, id1 = 0, id2 = 0;
        const bool hit = leaf_hit(A, lq[nl + lane], q, id1, id2);
        n_leaf += kWave;
        const unsigned long long hm = __ballot(hit);
        if (hit) {
          const int r = nh + __popcll(hm & lanes_below());
          hits[3 * r] = q;
          hits[3 * r + 1] = id1;
          hits[3 * r + 2] = id2;
          count_hits(A, q);
        }
        nh += __popcll(hm);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (nh > kHitStage - kWave) flush_hits();
        continue;
      }
    }
    // room for one full step in both queues
    const bool full = (nl >= kLq / 2 && !flush_deferred(A, 0, xcd, lq, nl, kLq / 2)) ||
                      (nu >= kUncQ / 2 && !flush_deferred(A, 1, xcd, uq, nu, kUncQ / 2));
    if (full) {
      if (lane == 0) __hip_atomic_store(A.stop, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      stopped = true;
      break;
    }
    if (sp < kWave && input_left) {
      // sp - base <= sp < 64: the new entries fit the LDS part
      const int want = kWave - sp;
      unsigned long long start = 0;
      if (lane == 0) start = atomicAdd(A.head, static_cast<unsigned long long>(want));
      start = __shfl(start, 0);
      const long long avail = A.n_in - static_cast<long long>(start);
      const int got = avail <= 0 ? 0 : static_cast<int>(avail < want ? avail : want);
      if (got < want) input_left = false;
      if (lane < got) slot(sp + lane) = A.front_in[start + lane];
      sp += got;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (sp == 0) break;
    const int n = sp < kWave ? sp : kWave;
    if constexpr (kInline) {
      // the popped entries must be in LDS: read back up to half the limit
      // (the LDS part holds fewer than 64 entries here, so with the limit
      // in [128, kTop] that fits the ring and covers the pop)
      if (sp - n < base) {
        const int half = A.stack_top / 2;
        const int r = base < half ? base : half;
        for (int i = lane; i < r; i += kWave) slot(base - r + i) = stack[base - r + i];
        base -= r;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    sp -= n;
    const bool active = lane < n;
    unsigned long long task = 0, c0 = 0, c1 = 0;
    if (active) task = slot(sp + lane);
    n_bv += n;
    ++n_steps;
    const int r = active ? step32<kStats>(A, task, c0, c1) : kStepNone;
    // children onto the stack (one prefix sum)
    const unsigned long long pm = __ballot(r == kStepPush);
    const int npush = 2 * __popcll(pm);
    if (sp + npush > A.stack_cap) {
      if (lane == 0) atomicOr(A.L.status, kStatusQueueOverflow);
      break;  // never write past the stack; the overflow is reported
    }
    if constexpr (kInline) {
      // keep [base, sp + npush) within the limit: write the deepest LDS
      // entries out, 64 at a time, until at most half the limit is in use
      // after the push (or all of them: then npush <= 128 entries remain)
      if (sp + npush - base > A.stack_top) {
        int w = sp + npush - base - A.stack_top / 2;
        w = (w + kWave - 1) & ~(kWave - 1);
        if (w > sp - base) w = sp - base;
        for (int i = lane; i < w; i += kWave) stack[base + i] = slot(base + i);
        base += w;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    if (r == kStepPush) {
      const int o = sp + 2 * __popcll(pm & lanes_below());
      slot(o) = c0;
      slot(o + 1) = c1;
    }
    sp += npush;
    // leaf pairs and uncertain tasks into the LDS queues
    const unsigned long long lm = __ballot(r == kStepLeaf);
    const unsigned long long um = __ballot(r == kStepUncertain);
    if (r == kStepLeaf) lq[nl + __popcll(lm & lanes_below())] = c0;
    if (r == kStepUncertain) uq[nu + __popcll(um & lanes_below())] = task;
    __builtin_amdgcn_wave_barrier();
    nl += __popcll(lm);
    nu += __popcll(um);
    n_unc += __popcll(um);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  flush_hits();
  if (stopped) {
    // hand the stack to the next round
    if (sp > 0) {
      unsigned long long s = 0;
      if (lane == 0) s = atomicAdd(A.n_out, static_cast<unsigned long long>(sp));
      s = __shfl(s, 0);
      for (int i = lane; i < sp; i += kWave) A.front_out[s + i] = i < base ? stack[i] : slot(i);
    }
    dump_deferred(A, 0, lq, nl);
    dump_deferred(A, 1, uq, nu);
  } else {
    if (nl > 0 && !flush_deferred(A, 0, xcd, lq, nl, nl)) dump_deferred(A, 0, lq, nl);
    if (nu > 0 && !flush_deferred(A, 1, xcd, uq, nu, nu)) dump_deferred(A, 1, uq, nu);
    if (lane == 0) {
      const unsigned long long idle = atomicAdd(A.idle, 1ULL) + 1;
      if (A.idle_limit != 0 && idle == A.idle_limit)
        __hip_atomic_store(A.stop, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (A.wave_prof && lane == 0) {
    unsigned long long* w =
        A.wave_prof + 3 * (static_cast<size_t>(blockIdx.x) * kWavesPerBlock + wave);
    w[0] = t_start;
    w[1] = wall_clock64();
    w[2] = n_bv;
  }
  if (lane == 0) {
    atomicAdd(A.work + kWorkDfsBV, n_bv);
    if (n_unc) atomicAdd(A.work + kWorkFallback, n_unc);
    if (n_leaf) atomicAdd(A.work + kWorkDfsInline, n_leaf);
    if (n_steps) atomicAdd(A.work + kWorkDfsSteps, n_steps);
  }
}

// ---------------------------------- fix-up of queries that reached the cap
// list[s] = q for every flagged query; qflag[q] becomes s + 1
__global__ void engine_collect_flagged(int32_t* qflag, int nq, int32_t* list,
                                       unsigned long long* n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = q < nq && qflag[q] != 0;
  const unsigned long long s = wave_reserve(n, f, 1);
  if (f) {
    list[s] = q;
    qflag[q] = static_cast<int32_t>(s) + 1;
  }
}

// positions of the records of flagged queries (one atomic per wavefront)
__global__ void engine_find_flagged(const int32_t* rec, long long nrec, const int32_t* qflag,
                                    long long* pos, unsigned long long* n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool f = i < nrec && qflag[rec[3 * i]] != 0;
  const unsigned long long s = wave_reserve(n, f, 1);
  if (f) pos[s] = i;
}

// segment (flagged-query index) of each flagged record
__global__ void engine_segment_of(const int32_t* rec, const long long* pos, long long n,
                                  const int32_t* qflag, uint32_t* seg) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) seg[i] = static_cast<uint32_t>(qflag[rec[3 * pos[i]]] - 1);
}

// flagged positions below `limit` (holes to refill)
__global__ void engine_holes(const long long* pos, long long n, long long limit, long long* out,
                             unsigned long long* cnt) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool h = i < n && pos[i] < limit;
  const unsigned long long s = wave_reserve(cnt, h, 1);
  if (h) out[s] = pos[i];
}

// unflagged positions in [from, nrec) (records to move into holes)
__global__ void engine_tail_keep(const int32_t* rec, long long from, long long nrec,
                                 const int32_t* qflag, long long* out, unsigned long long* cnt) {
  const long long i = from + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool k = i < nrec && qflag[rec[3 * i]] == 0;
  const unsigned long long s = wave_reserve(cnt, k, 1);
  if (k) out[s] = i;
}

// stage the records that will fill the holes: truncated lists first
// (segment-major, depth-first order), then the kept tail records
__global__ void engine_stage_fill(const int32_t* rec, const long long* sorted,
                                  const long long* seg_off, const long long* out_off,
                                  const int32_t* take, int nseg, const long long* tail,
                                  long long ntail, long long ntake, int32_t* stage) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < ntake) {
    // segment of output slot i: last seg with out_off[seg] <= i
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (out_off[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const long long j = i - out_off[lo];
    if (j < take[lo]) {
      const int32_t* r = rec + 3 * sorted[seg_off[lo] + j];
      stage[3 * i] = r[0];
      stage[3 * i + 1] = r[1];
      stage[3 * i + 2] = r[2];
    }
  } else if (i < ntake + ntail) {
    const int32_t* r = rec + 3 * tail[i - ntake];
    stage[3 * i] = r[0];
    stage[3 * i + 1] = r[1];
    stage[3 * i + 2] = r[2];
  }
}

__global__ void engine_fill_holes(const int32_t* stage, const long long* holes, long long n,
                                  int32_t* rec) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long o = holes[i];
  rec[3 * o] = stage[3 * i];
  rec[3 * o + 1] = stage[3 * i + 1];
  rec[3 * o + 2] = stage[3 * i + 2];
}

// Depth-first position of a leaf pair: the BVTT path from (0,0) to the
// pair's leaf nodes, one bit per descent (0 = first child), most significant
// first.  The descent rule is the traversal's (firstOverSecond); the child
// holding a triangle is the one whose primitive range contains the
// triangle's build position.  Distinct leaf pairs have paths that differ
// before either ends, so keys order pairs exactly as depth-first search.
__global__ void engine_leaf_keys(const int32_t* rec, const long long* grouped, long long n,
                                 const CollideLaunch L, unsigned long long* keys) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* r = rec + 3 * grouped[i];
  const int q = r[0];
  const MeshDev& m1 = L.meshes[L.query_meshes[2 * q]];
  const MeshDev& m2 = L.meshes[L.query_meshes[2 * q + 1]];
  const int p1 = m1.prim_pos[r[1]], p2 = m2.prim_pos[r[2]];
  int x1 = 0, x2 = 0, len = 0;
  unsigned long long key = 0;
  while (true) {
    const int fc1 = m1.obb64[x1].first_child, fc2 = m2.obb64[x2].first_child;
    if (fc1 < 0 && fc2 < 0) break;
    int bit;
    if ((fc2 < 0) || (fc1 >= 0 && (m1.size64[x1] > m2.size64[x2]))) {
      bit = p1 < m1.span[2 * fc1] + m1.span[2 * fc1 + 1] ? 0 : 1;
      x1 = fc1 + bit;
    } else {
      bit = p2 < m2.span[2 * fc2] + m2.span[2 * fc2 + 1] ? 0 : 1;
      x2 = fc2 + bit;
    }
    if (len < 64) key |= static_cast<unsigned long long>(bit) << (63 - len);
    ++len;
  }
  keys[i] = key;
}


__global__ void engine_gather_counts(const int32_t* list, int n, const int32_t* counts,
                                     int32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = counts[list[i]];
}

__global__ void engine_set_counts(const int32_t* list, const int32_t* take, int n,
                                  int32_t* counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counts[list[i]] = take[i];
}

// copy records of unflagged queries to out (unordered), count in n
__global__ void engine_drop_flagged(const int32_t* rec, long long nrec, const int32_t* qflag,
                                    int32_t* out, unsigned long long* n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool keep = i < nrec && qflag[rec[3 * i]] == 0;
  const unsigned long long s = wave_reserve(n, keep, 1);
  if (keep) {
    out[3 * s] = rec[3 * i];
    out[3 * s + 1] = rec[3 * i + 1];
    out[3 * s + 2] = rec[3 * i + 2];
  }
}

// gather the flagged queries' inputs into a compact batch
__global__ void engine_gather_flagged(const int32_t* list, int n, const int32_t* qm,
                                      const double* tf1, const double* tf2, int32_t* qm_o,
                                      double* tf1_o, double* tf2_o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int q = list[i];
  qm_o[2 * i] = qm[2 * q];
  qm_o[2 * i + 1] = qm[2 * q + 1];
  for (int k = 0; k < 12; ++k) {
    tf1_o[12 * i + k] = tf1[12 * static_cast<size_t>(q) + k];
    tf2_o[12 * i + k] = tf2[12 * static_cast<size_t>(q) + k];
  }
}

// scatter the re-run results back: counts/stats per query, records appended
__global__ void engine_scatter_flagged(const int32_t* list, int n, const int32_t* counts_s,
                                       const long long* bvt_s, const long long* lft_s,
                                       int32_t* counts, long long* bvt, long long* lft) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int q = list[i];
  counts[q] = counts_s[i];
  if (bvt) bvt[q] = bvt_s[i];
  if (lft) lft[q] = lft_s[i];
}

__global__ void engine_append_records(const int32_t* rec, long long nrec, const int32_t* list,
                                      int32_t* out, long long out_base, long long cap,
                                      int32_t* status) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  const long long o = out_base + i;
  if (o >= cap) {
    atomicOr(status, kStatusPairOverflow);
    return;
  }
  out[3 * o] = list[rec[3 * i]];
  out[3 * o + 1] = rec[3 * i + 1];
  out[3 * o + 2] = rec[3 * i + 2];
}

__global__ void set_u64(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void zero_i32_grid(int32_t* p, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void zero_u64(unsigned long long* p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0;
}

// ------------------------------------------------- booleans (any hit)
// Whether each query's pair collides (count 1 or 0, no records).  That
// answer does not depend on the order in which the BVTT is visited, only on
// whether some leaf pair intersects, but the work it takes does: a pair
// that collides is answered at its first intersecting leaf pair, which a
// narrow depth-first descent reaches soonest, while a wavefront needs 64
// tasks per step.  So a wavefront keeps kSlots queries in flight, each with
// a LIFO stack of its own in device memory, and a step tests up to
// 64 / kSlots tasks from the top of each stack (up to `width` when other
// stacks run short): every query descends nearly depth first while the
// lanes stay busy.  A slot whose query is answered takes the next one from
// the launch's query counter, kSlots queries per atomic.
constexpr int kAnyWaves = kBlock / kWave;

// A slot's query in LDS as 16-byte chunks, chunk-major: lanes that read the
// same field of different slots read neighbouring words (no bank conflicts).
template <typename T, int kSlots>
struct SlotRecords {
  static constexpr int kChunks = sizeof(T) / 16;
  static_assert(sizeof(T) % 16 == 0, "records are whole 16-byte chunks");
  uint4 c[kChunks][kSlots];
  __device__ __forceinline__ void put(int slot, const T& v) {
    const uint4* p = reinterpret_cast<const uint4*>(&v);
#pragma unroll
    for (int k = 0; k < kChunks; ++k) c[k][slot] = p[k];
  }
  __device__ __forceinline__ T get(int slot) const {
    union {
      uint4 u[kChunks];
      T v;
    } r;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) r.u[k] = c[k][slot];
    return r.v;
  }
};

// FP64 test of a BV pair the FP32 cull left unsure; c0/c1 carry `tag` in
// the query field.
__device__ __forceinline__ Visit visit_any(const CollideLaunch& L, const QueryXf32& X,
                                           const QueryPtr& P, const QueryXf64& x, int tag,
                                           int b1, int b2) {
  const ObbNode32& n1 = P.n1[b1];
  const ObbNode32& n2 = P.n2[b2];
  Visit v;
  M3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = x.R[i];
  if (obb_overlap64(R, v3(x.T[0], x.T[1], x.T[2]), L.meshes[X.m1].obb64[b1],
                    L.meshes[X.m2].obb64[b2])) {
    const int fc1 = n1.first_child, fc2 = n2.first_child;
    if (fc1 < 0 && fc2 < 0) {
      v.leafc = true;
      v.c0 = pack_task(tag, -(fc1 + 1), -(fc2 + 1));
    } else {
      v.push = true;
      if ((fc2 < 0) || (fc1 >= 0 && (P.s1[b1] > P.s2[b2]))) {
        v.c0 = pack_task(tag, fc1, b2);
        v.c1 = pack_task(tag, fc1 + 1, b2);
      } else {
        v.c0 = pack_task(tag, b1, fc2);
        v.c1 = pack_task(tag, b1, fc2 + 1);
      }
    }
  }
  return v;
}

__device__ __forceinline__ bool leaf_any(const QueryXf64& x, int id1, int id2) {
  const double* p = x.tv1 + kTriVertsStride * static_cast<size_t>(id1);
  const double* r = x.tv2 + kTriVertsStride * static_cast<size_t>(id2);
  const V3 P1 = v3(p[0], p[1], p[2]), P2 = v3(p[3], p[4], p[5]), P3 = v3(p[6], p[7], p[8]);
  const V3 Q1 = v3(r[0], r[1], r[2]), Q2 = v3(r[3], r[4], r[5]), Q3 = v3(r[6], r[7], r[8]);
  M3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = x.R[i];
  return tri_intersect64(P1, P2, P3, Q1, Q2, Q3, R, v3(x.T[0], x.T[1], x.T[2]));
}

// Outcome of one BV task after the FP32 cull: nothing, two children, a leaf
// pair for the triangle test, or a pair for the FP64 BV test.
enum : int { kAnyNone = 0, kAnyPush, kAnyLeaf, kAnyUnsure };

__device__ __forceinline__ int cull_any(const QueryXf32& X, const QueryPtr& P, int tag, int b1,
                                        int b2, unsigned long long& c0, unsigned long long& c1) {
  const ObbNode32 n1 = load_node32(P.n1 + b1);
  const ObbNode32 n2 = load_node32(P.n2 + b2);
  const int r = obb_cull32(X, n1, n2);
  if (r == kCullUnsure) return kAnyUnsure;
  if (r == kCullDisjoint) return kAnyNone;
  const int fc1 = n1.first_child, fc2 = n2.first_child;
  if (fc1 < 0 && fc2 < 0) {
    c0 = pack_task(tag, -(fc1 + 1), -(fc2 + 1));
    return kAnyLeaf;
  }
  if ((fc2 < 0) || (fc1 >= 0 && first_larger(n1, n2, P.s1 + b1, P.s2 + b2))) {
    c0 = pack_task(tag, fc1, b2);
    c1 = pack_task(tag, fc1 + 1, b2);
  } else {
    c0 = pack_task(tag, b1, fc2);
    c1 = pack_task(tag, b1, fc2 + 1);
  }
  return kAnyPush;
}

// A wavefront's steps come in three kinds, one kind per step for all its
// lanes: FP32 BV tests from the slots' stacks; up to 64 FP64 triangle tests
// from the leaf queue; up to 64 FP64 BV tests of the pairs the FP32 cull
// left unsure.  The queues (LDS, 128 entries) are worked off once they hold
// leaf_batch (<= 64) leaf pairs / 64 unsure pairs, or once the stacks
// would fill less than half the lanes, so that no lane waits on another's
// FP64 work and no slot waits long on its queued pairs.  Queue entries
// carry (generation, slot): a slot whose query has a hit is reused at once
// and the entries its query left behind are dropped when they come up; a
// slot whose query has none is reused once its stack is empty and none of
// its entries is queued.
constexpr int kAnyQueue = 128;
constexpr int kAnySlotBits = 5;  // slots per wavefront <= 32
constexpr int kAnyXcds = 8;      // query ranges (one per XCD)
constexpr int kAnyAreaHead = 16 * kAnyXcds;  // area: range counters, 128 B apart; then the stacks
constexpr long long kAnySmallBatch = 65536;  // bool_queries_per_wave 0: 4 slots per wavefront
                                             // below this many queries, else 8
constexpr int kAnySlotCap = 1024;            // stack entries per slot (at least)
constexpr int kAnyGenMask = (1 << (22 - kAnySlotBits)) - 1;

template <int kSlots>
struct AnyWave {
  SlotRecords<QueryXf32, kSlots> x32;
  SlotRecords<QueryPtr, kSlots> ptr;
  SlotRecords<QueryXf64, kSlots> x64;
  int sp[kSlots];     // stack depth per slot
  int pend[kSlots];   // queued leaf and unsure entries of the slot's query
  int found[kSlots];  // 0 open, 1 hit, 2 abandoned (stack full; reported)
  int gen[kSlots];    // generation of the slot's query
  int offs[kSlots];   // first lane of each slot in a BV step
  int base[kSlots];   // stack position of that lane's task
  unsigned long long lq[kAnyQueue];
  unsigned long long uq[kAnyQueue];
  unsigned char lane_slot[kWave];
  // exhaustive mode: the slot's query, its contacts, BV tests, leaf tests
  int qid[kSlots];
  int cnt[kSlots];
  unsigned long long nbv[kSlots];
  unsigned long long nlf[kSlots];
};

// Inclusive sum over lanes 0..n-1 (n <= 32; valid in those lanes) with DPP
// row shifts, and lane n-1's value on the scalar unit.
template <int n>
__device__ __forceinline__ int scan_lanes(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
  if (n > 2) v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
  if (n > 4) v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
  if (n > 8) v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
  if (n > 16) {
    const int row0 = __builtin_amdgcn_readlane(v, 15);
    if (lane_id() >= 16) v += row0;
  }
  return v;
}
template <int n>
__device__ __forceinline__ int lane_total(int incl) {
  return __builtin_amdgcn_readlane(incl, n - 1);
}

// kAll = false: booleans (a query ends at its first intersecting leaf
// pair).  kAll = true: every intersecting leaf pair of every query is
// recorded (one reservation per wavefront in the caller's pair buffer), with
// the query's count and num_bv_tests / num_leaf_tests; nothing stops a
// query early, so the BVTT nodes visited are the reference's, and a query
// whose count reaches num_max_contacts is flagged (qflag, nflag) for the
// depth-first prefix fix-up.
template <int kSlots, int kWavesPerSimd, bool kAll>
__global__ void __launch_bounds__(kBlock, kWavesPerSimd) engine_anyhit(CollideLaunch L,
                                                                       unsigned long long* area,
                                                                       unsigned long long* stk,
                                                                       int slot_cap, int wide_cap,
                                                                       int width,
                                                                       int leaf_batch, int refill,
                                                                       unsigned long long* prof,
                                                                       unsigned long long* wprof,
                                                                       int32_t* qflag,
                                                                       unsigned long long* nflag) {
  static_assert(kSlots <= (1 << kAnySlotBits) && kSlots <= 32, "slot index must fit its field");
  __shared__ AnyWave<kSlots> waves_s[kAnyWaves];
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  AnyWave<kSlots>& W = waves_s[wave];
  const long long gw = static_cast<long long>(blockIdx.x) * kAnyWaves + wave;
  // Queries are split into one contiguous range per XCD (workgroups go to
  // the XCDs round-robin), each with its counter on a line of its own:
  // neighbouring queries (consecutive waypoints of a trajectory) touch the
  // same part of the scene, so an XCD's L2 holds fewer BVH nodes.  A
  // wavefront whose range is used up takes from the next XCD's.
  int xcd = static_cast<int>(blockIdx.x & (kAnyXcds - 1));
  int xcd_left = kAnyXcds;
  unsigned long long* const stacks = stk + gw * kSlots * static_cast<long long>(slot_cap);
  const bool slot_lane = lane < kSlots;
  unsigned long long* const mine = stacks + static_cast<long long>(slot_lane ? lane : 0) * slot_cap;
  const int nq = L.num_queries;
  int q = -1;          // slot lanes: the query of this lane's slot
  int spv = 0;         // slot lanes: its stack depth (W.sp mirrors it)
  int nx = 0, ne = 0;  // queries taken from the counter, not yet in a slot
  bool exhausted = false;  // every query of the batch is taken
  bool check = true;       // some slot may be answered, or free
  int nl = 0, nu = 0;
  unsigned long long st[3] = {0, 0, 0}, ln[3] = {0, 0, 0}, skipped = 0, rounds = 0;
  int sp_max = 0;
  const unsigned long long t_start = prof ? wall_clock64() : 0;
  if (slot_lane) W.gen[lane] = 0;
  // Starts query qq in this lane's slot (false: a bad query).
  auto start = [&](int qq) {
    const int i1 = L.query_meshes[2 * qq], i2 = L.query_meshes[2 * qq + 1];
    if (L.bv_tests) L.bv_tests[qq] = 0;
    if (L.leaf_tests) L.leaf_tests[qq] = 0;
    if (!query_ok(L.meshes, L.num_meshes, i1, i2)) {
      atomicOr(L.status, kStatusBadQuery);
      return false;
    }
    QueryXf64 x;
    rel_pose64(L.tf1 + 12 * static_cast<size_t>(qq), L.tf2 + 12 * static_cast<size_t>(qq), x.R,
               x.T);
    const MeshDev& m1 = L.meshes[i1];
    const MeshDev& m2 = L.meshes[i2];
    x.tv1 = m1.tri_verts;
    x.tv2 = m2.tri_verts;
    QueryXf32 X;
    make_xf32(x.R, x.T, m1.cmax, m2.cmax, X);
    X.m1 = i1;
    X.m2 = i2;
    X.valid = 1;
    W.x32.put(lane, X);
    W.ptr.put(lane, QueryPtr{m1.obb32, m2.obb32, m1.size64, m2.size64});
    W.x64.put(lane, x);
    mine[0] = pack_task(0, 0, 0);
    W.sp[lane] = 1;
    W.found[lane] = 0;
    W.pend[lane] = 0;
    W.gen[lane] = (W.gen[lane] + 1) & kAnyGenMask;
    if constexpr (kAll) {
      W.qid[lane] = qq;
      W.cnt[lane] = 0;
      W.nbv[lane] = 0;
      W.nlf[lane] = 0;
    }
    return true;
  };
  for (;;) {
    if (check) {
      check = false;
      // retire answered queries: a hit, or nothing left on the stack or queued
      if (slot_lane && q >= 0) {
        spv = W.sp[lane];
        const int f = W.found[lane];
        if (f != 0 || (spv == 0 && W.pend[lane] == 0)) {
          if constexpr (kAll) {
            const int c = W.cnt[lane];
            L.counts[q] = c;
            if (L.bv_tests) L.bv_tests[q] = static_cast<long long>(W.nbv[lane]);
            if (L.leaf_tests) L.leaf_tests[q] = static_cast<long long>(W.nlf[lane]);
            if (c >= L.cap) {
              qflag[q] = 1;
              atomicAdd(nflag, 1ULL);
            }
          } else {
            L.counts[q] = f == 1 ? 1 : 0;
          }
          q = -1;
          spv = 0;
        }
      }
      const int tasks = lane_total<kSlots>(scan_lanes<kSlots>(spv));
      unsigned long long freem = __ballot(slot_lane && q < 0);
      if (!exhausted && freem != 0 && (__popcll(freem) >= refill || tasks < kWave)) {
        ++rounds;
        while (freem != 0) {
          if (nx == ne) {
            const long long lo = static_cast<long long>(nq) * xcd / kAnyXcds;
            const long long hi = static_cast<long long>(nq) * (xcd + 1) / kAnyXcds;
            unsigned long long b = 0;
            if (lane == 0) b = atomicAdd(area + 16 * xcd, static_cast<unsigned long long>(kSlots));
            b = __shfl(b, 0);
            if (static_cast<long long>(b) >= hi - lo) {
              xcd = (xcd + 1) & (kAnyXcds - 1);
              if (--xcd_left == 0) {
                exhausted = true;
                break;
              }
              continue;
            }
            nx = static_cast<int>(lo + static_cast<long long>(b));
            ne = hi - nx < kSlots ? static_cast<int>(hi) : nx + kSlots;
          }
          const int r = __popcll(freem & lanes_below());
          const int n = __popcll(freem) < ne - nx ? __popcll(freem) : ne - nx;
          if (slot_lane && q < 0 && r < n) {
            q = nx + r;
            if (start(q)) {
              spv = 1;
            } else {
              L.counts[q] = 0;
              q = -1;
            }
          }
          nx += n;
          freem = __ballot(slot_lane && q < 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      if (exhausted && __ballot(slot_lane && q >= 0) == 0) break;
    }

    // tasks per slot for a BV step: an even share of the lanes first, then
    // what is left to the slots with deeper stacks, up to `width`.  A step
    // that takes t tasks grows the stack by at most t, so a slot keeps its
    // depth under wide_cap with wide steps and goes one task at a time above
    // it: the stack then holds at most two entries per level above that
    // point, which the margin between wide_cap and slot_cap covers.
    // (once every query is taken, lanes would idle: up to all 64 per query)
    const int lim = wide_cap - spv > 1 ? wide_cap - spv : 1;
    const int wmax = exhausted ? kWave : width;
    const int cap_w = wmax < lim ? wmax : lim;
    int take = spv < kWave / kSlots ? spv : kWave / kSlots;
    take = take < cap_w ? take : cap_w;
    int incl = scan_lanes<kSlots>(take);
    int total = lane_total<kSlots>(incl);
    if (total < kWave) {
      const int e = (spv < cap_w ? spv : cap_w) - take;
      const int ei = scan_lanes<kSlots>(e);
      const int room = kWave - total - (ei - e);
      take += room <= 0 ? 0 : (room < e ? room : e);
      incl = scan_lanes<kSlots>(take);
      total = lane_total<kSlots>(incl);
    }
    const int excl = incl - take;

    if (nl >= leaf_batch || (total == 0 && nl > 0)) {
      // FP64 triangle tests
      const int n = nl < kWave ? nl : kWave;
      nl -= n;
      bool skip = false, hit = false;
      int hq = 0, h1 = 0, h2 = 0;
      if (lane < n) {
        int tag, t1, t2;
        unpack_task(W.lq[nl + lane], tag, t1, t2);
        const int slot = tag & ((1 << kAnySlotBits) - 1);
        skip = (tag >> kAnySlotBits) != W.gen[slot] || W.found[slot] != 0;
        if (!skip) {
          hit = leaf_any(W.x64.get(slot), t1, t2);
          if constexpr (kAll) {
            atomicAdd(&W.nlf[slot], 1ULL);
            if (hit) {
              atomicAdd(&W.cnt[slot], 1);
              hq = W.qid[slot];
              h1 = t1;
              h2 = t2;
            }
          } else if (hit) {
            W.found[slot] = 1;
          }
          atomicSub(&W.pend[slot], 1);
        }
      }
      if constexpr (kAll) {
        // the hits' records, one reservation for the wavefront
        const unsigned long long at = wave_reserve(
            reinterpret_cast<unsigned long long*>(L.num_pairs), hit, 1);
        if (hit) {
          if (static_cast<long long>(at) < L.pair_capacity) {
            L.pairs[3 * at] = hq;
            L.pairs[3 * at + 1] = h1;
            L.pairs[3 * at + 2] = h2;
          } else if (L.pairs) {
            atomicOr(L.status, kStatusPairOverflow);
          }
        }
      }
      ++st[1];
      ln[1] += n;
      skipped += __popcll(__ballot(skip));
      check = true;
    } else if (nu >= kWave || (total == 0 && nu > 0)) {
      // FP64 BV tests of unsure pairs: children to their slot's stack
      const int n = nu < kWave ? nu : kWave;
      nu -= n;
      Visit v;
      int slot = 0;
      bool skip = false;
      if (lane < n) {
        int tag, b1, b2;
        unpack_task(W.uq[nu + lane], tag, b1, b2);
        slot = tag & ((1 << kAnySlotBits) - 1);
        skip = (tag >> kAnySlotBits) != W.gen[slot] || W.found[slot] != 0;
        if (!skip) {
          v = visit_any(L, W.x32.get(slot), W.ptr.get(slot), W.x64.get(slot), tag, b1, b2);
          atomicSub(&W.pend[slot], 1);
        }
      }
      if (v.push) {
        const int o = atomicAdd(&W.sp[slot], 2);
        if (o + 2 <= slot_cap) {
          unsigned long long* stk = stacks + static_cast<long long>(slot) * slot_cap;
          stk[o] = v.c0;
          stk[o + 1] = v.c1;
        } else {
          W.found[slot] = 2;
          atomicOr(L.status, kStatusStackOverflow);
        }
      }
      const unsigned long long lm = __ballot(v.leafc);
      if (v.leafc) {
        W.lq[nl + __popcll(lm & lanes_below())] = v.c0;
        atomicAdd(&W.pend[slot], 1);
      }
      nl += __popcll(lm);
      ++st[2];
      ln[2] += n;
      skipped += __popcll(__ballot(skip));
      check = true;  // stacks grew through W.sp
    } else if (total > 0) {
      // FP32 BV tests from the tops of the stacks
      if (slot_lane) {
        W.offs[lane] = excl;
        W.base[lane] = spv - take;
        for (int j = 0; j < take; ++j) W.lane_slot[excl + j] = static_cast<unsigned char>(lane);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      int kind = kAnyNone, slot = 0, first = 0, b1 = 0, b2 = 0, tag = 0, at = 0;
      unsigned long long c0 = 0, c1 = 0;
      if (lane < total) {
        slot = W.lane_slot[lane];
        first = W.offs[slot];
        at = W.base[slot];
        tag = (W.gen[slot] << kAnySlotBits) | slot;
        int dummy;
        unpack_task(stacks[static_cast<long long>(slot) * slot_cap + at + lane - first], dummy, b1,
                    b2);
        kind = cull_any(W.x32.get(slot), W.ptr.get(slot), tag, b1, b2, c0, c1);
      }
      const unsigned long long pm = __ballot(kind == kAnyPush);
      const unsigned long long lm = __ballot(kind == kAnyLeaf);
      const unsigned long long um = __ballot(kind == kAnyUnsure);
      if (kind == kAnyPush) {
        const unsigned long long same = ~((1ULL << first) - 1ULL);
        const int o = at + 2 * __popcll(pm & lanes_below() & same);
        if (o + 2 <= slot_cap) {
          unsigned long long* stk = stacks + static_cast<long long>(slot) * slot_cap;
          stk[o] = c0;
          stk[o + 1] = c1;
        }
      }
      if (kind == kAnyLeaf) W.lq[nl + __popcll(lm & lanes_below())] = c0;
      if (kind == kAnyUnsure) W.uq[nu + __popcll(um & lanes_below())] = pack_task(tag, b1, b2);
      nl += __popcll(lm);
      nu += __popcll(um);
      bool emptied = false;
      if (slot_lane && take > 0) {
        const unsigned long long own = (take >= kWave ? ~0ULL : ((1ULL << take) - 1ULL)) << excl;
        const int nsp = spv - take + 2 * __popcll(pm & own);
        if (nsp > slot_cap) {
          W.found[lane] = 2;
          atomicOr(L.status, kStatusStackOverflow);
          emptied = true;
        } else {
          spv = nsp;
          W.sp[lane] = nsp;
          emptied = nsp == 0;
          sp_max = nsp > sp_max ? nsp : sp_max;
        }
        W.pend[lane] += __popcll((lm | um) & own);
        if constexpr (kAll) W.nbv[lane] += static_cast<unsigned long long>(take);
      }
      if (__ballot(emptied) != 0) check = true;
      ++st[0];
      ln[0] += total;
    } else {
      check = true;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (prof && lane == 0) {
    for (int k = 0; k < 3; ++k) {
      atomicAdd(prof + k, st[k]);
      atomicAdd(prof + 3 + k, ln[k]);
    }
    atomicAdd(prof + 6, skipped);
    atomicAdd(prof + 7, rounds);
  }
  if (prof) {
    for (int d = 1; d < kSlots; d <<= 1) {
      const int o = __shfl_xor(sp_max, d);
      sp_max = o > sp_max ? o : sp_max;
    }
  }
  if (prof && lane == 0) {
    atomicMax(prof + 8, static_cast<unsigned long long>(sp_max));
    unsigned long long* w = wprof + 3 * gw;
    w[0] = t_start;
    w[1] = wall_clock64();
    w[2] = rounds;
  }
}

// ---------------------------------------------- ordered, small caps
template <int STACK>
__global__ void __launch_bounds__(kBlock) ordered_collide(CollideLaunch L,
                                                          unsigned long long* next) {
  // per-lane state
  int q = -1;
  int count = 0;
  long long bvt = 0, lft = 0;
  int sp = 0, b1 = 0, b2 = 0;
  bool overflow = false;
  M3 R;
  V3 T;
  QueryXf32 X;
  MeshDev m1{}, m2{};
  int2 stack[STACK];
  while (true) {
    const bool need = q == -1;  // -2: the batch is exhausted for this lane
    const unsigned long long nm = __ballot(need);
    if (nm != 0) {
      const int leader = __ffsll(static_cast<long long>(nm)) - 1;
      unsigned long long base = 0;
      if (lane_id() == leader) base = atomicAdd(next, static_cast<unsigned long long>(__popcll(nm)));
      base = __shfl(base, leader);
      if (need) {
        const long long cand = static_cast<long long>(base + __popcll(nm & lanes_below()));
        q = cand < L.num_queries ? static_cast<int>(cand) : -2;
        if (q >= 0) {
          const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
          count = 0;
          bvt = lft = 0;
          sp = 0;
          b1 = b2 = 0;
          if (!query_ok(L.meshes, L.num_meshes, i1, i2)) {
            L.counts[q] = 0;
            if (L.bv_tests) L.bv_tests[q] = 0;
            if (L.leaf_tests) L.leaf_tests[q] = 0;
            atomicOr(L.status, kStatusBadQuery);
            q = -1;  // fetch again next round
          } else {
            m1 = L.meshes[i1];
            m2 = L.meshes[i2];
            double Rd[9], Td[3];
            rel_pose64(L.tf1 + 12 * static_cast<size_t>(q), L.tf2 + 12 * static_cast<size_t>(q),
                       Rd, Td);
#pragma unroll
            for (int k = 0; k < 9; ++k) R.m[k] = Rd[k];
            T = v3(Td[0], Td[1], Td[2]);
            make_xf32(Rd, Td, m1.cmax, m2.cmax, X);
          }
        }
      }
    }
    if (__ballot(q != -2) == 0) break;
    if (q < 0) continue;
    // one BVTT node visit, as in the plain kernel
    const int fc1 = m1.obb32[b1].first_child, fc2 = m2.obb32[b2].first_child;
    ++bvt;
    bool descend = false;
    int r = obb_cull32(X, m1.obb32[b1], m2.obb32[b2]);
    if (r == kCullUnsure) r = obb_overlap64(R, T, m1.obb64[b1], m2.obb64[b2]) ? 1 : 0;
    if (r == kCullOverlap) {
      if (fc1 < 0 && fc2 < 0) {
        ++lft;
        const int id1 = -(fc1 + 1), id2 = -(fc2 + 1);
        const int32_t* t1 = m1.tris + 3 * static_cast<size_t>(id1);
        const int32_t* t2 = m2.tris + 3 * static_cast<size_t>(id2);
        if (tri_intersect64(vtx(m1, t1[0]), vtx(m1, t1[1]), vtx(m1, t1[2]), vtx(m2, t2[0]),
                            vtx(m2, t2[1]), vtx(m2, t2[2]), R, T)) {
          if (count < L.cap) {
            const long long slot = static_cast<long long>(
                atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs), 1ULL));
            if (slot < L.pair_capacity) {
              L.pairs[3 * slot] = q;
              L.pairs[3 * slot + 1] = id1;
              L.pairs[3 * slot + 2] = id2;
            } else if (L.pairs) {  // no buffer: counts only
              atomicOr(L.status, kStatusPairOverflow);
            }
            ++count;
          }
        }
      } else {
        descend = true;
      }
    }
    bool done = false;
    if (descend) {
      if (sp >= STACK) {
        overflow = true;
        done = true;
      } else if ((fc2 < 0) || (fc1 >= 0 && (m1.size64[b1] > m2.size64[b2]))) {
        stack[sp++] = make_int2(fc1 + 1, b2);
        b1 = fc1;
      } else {
        stack[sp++] = make_int2(b1, fc2 + 1);
        b2 = fc2;
      }
    } else if (sp == 0 || (count > 0 && count >= L.cap)) {
      done = true;
    } else {
      const int2 e = stack[--sp];
      b1 = e.x;
      b2 = e.y;
    }
    if (done) {
      L.counts[q] = count;
      if (L.bv_tests) L.bv_tests[q] = bvt;
      if (L.leaf_tests) L.leaf_tests[q] = lft;
      if (overflow) atomicOr(L.status, kStatusStackOverflow);
      overflow = false;
      q = -1;
    }
  }
}

int device_cus() {
  static int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 256;
    return n > 0 ? n : 256;
  }();
  return cus;
}

template <typename T>
hipError_t grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap) return hipSuccess;
  size_t ncap = std::max(need, cap * 2);
  (void)hipFree(p);
  p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * ncap);
  if (e == hipSuccess) cap = ncap;
  return e;
}

#define ENG_TRY(x)                       \
  do {                                   \
    hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) return e_;     \
  } while (0)

// grow() for a buffer whose first `used` elements must survive: the new
// allocation receives them (stream-ordered) before the old one is freed
template <typename T>
hipError_t grow_keep(T*& p, size_t& cap, size_t need, size_t used, hipStream_t s) {
  if (need <= cap) return hipSuccess;
  const size_t ncap = std::max(need, cap * 2);
  T* np = nullptr;
  ENG_TRY(hipMalloc(reinterpret_cast<void**>(&np), sizeof(T) * ncap));
  hipError_t e = hipSuccess;
  if (used > 0) {
    e = hipMemcpyAsync(np, p, sizeof(T) * used, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    (void)hipFree(np);
    return e;
  }
  (void)hipFree(p);
  p = np;
  cap = ncap;
  return hipSuccess;
}

}  // namespace

struct EngineWorkspace {
  QueryXf32* qx32 = nullptr;
  size_t qx32_cap = 0;
  QueryPtr* qptr = nullptr;
  size_t qptr_cap = 0;
  QueryXf64* qx64 = nullptr;
  size_t qx64_cap = 0;
  int32_t* qflag = nullptr;
  size_t qflag_cap = 0;
  unsigned long long* front[2] = {nullptr, nullptr};
  size_t front_cap[2] = {0, 0};
  unsigned long long* leafc = nullptr;
  size_t leafc_cap = 0;
  unsigned long long* stacks = nullptr;
  size_t stacks_cap = 0;
  unsigned long long* defer[2] = {nullptr, nullptr};
  size_t defer_cap[2] = {0, 0};
  unsigned long long* counters = nullptr;  // kNumCounters device counters
  unsigned long long* host_counters = nullptr;
  bool profile = false;
  // profiling: event pairs around launches, summed per phase at sync points
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<int, size_t>> spans;  // (phase, index of the start event)
  size_t ev_used = 0;
  unsigned long long* wave_prof = nullptr;
  size_t wave_prof_cap = 0;
  double busy_sum = 0, busy_den = 0;
  EngineReport report{};
  EngineTuning tune{};
  // work counters of the small-cap kernel: batches on different streams may
  // run at once, so each launch takes the next slot of a ring (one cache
  // line each) and waits for the slot's previous launch to end
  static constexpr int kOrdSlots = 16;
  // stack areas of the boolean kernel, a ring guarded like the counters
  static constexpr int kAnySlots = 8;
  unsigned long long* any_stk[kAnySlots] = {};
  size_t any_cap[kAnySlots] = {};
  hipEvent_t any_ev[kAnySlots] = {};
  bool any_used[kAnySlots] = {};
  int any_next = 0;
  // work area of the exhaustive persistent form (one batch at a time)
  unsigned long long* all_stk = nullptr;
  size_t all_cap = 0;
  // its launch sequence captured in a HIP graph, replayed while the launch
  // arguments stay the same (the key); captured on a private stream
  hipGraphExec_t all_graph = nullptr;
  std::vector<unsigned long long> all_key;
  hipStream_t cap_stream = nullptr;
  unsigned long long* ord_ctr = nullptr;  // kOrdSlots x 16
  hipEvent_t ord_ev[kOrdSlots] = {};
  bool ord_used[kOrdSlots] = {};
  int ord_next = 0;
};

EngineTuning& engine_tuning(EngineWorkspace* ws) { return ws->tune; }

EngineWorkspace* engine_workspace_create() {
  auto* ws = new EngineWorkspace();
  bool ok = hipMalloc(reinterpret_cast<void**>(&ws->counters),
                      kNumCounters * sizeof(unsigned long long)) == hipSuccess &&
            hipHostMalloc(reinterpret_cast<void**>(&ws->host_counters),
                          kNumCounters * sizeof(unsigned long long)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->ord_ctr),
                      EngineWorkspace::kOrdSlots * 16 * sizeof(unsigned long long)) ==
                hipSuccess;
  for (int i = 0; ok && i < EngineWorkspace::kOrdSlots; ++i)
    ok = hipEventCreateWithFlags(&ws->ord_ev[i], hipEventDisableTiming) == hipSuccess;
  for (int i = 0; ok && i < EngineWorkspace::kAnySlots; ++i)
    ok = hipEventCreateWithFlags(&ws->any_ev[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    engine_workspace_destroy(ws);
    return nullptr;
  }
  return ws;
}

void engine_workspace_destroy(EngineWorkspace* ws) {
  if (!ws) return;
  (void)hipFree(ws->qx32);
  (void)hipFree(ws->qptr);
  (void)hipFree(ws->qx64);
  (void)hipFree(ws->qflag);
  (void)hipFree(ws->front[0]);
  (void)hipFree(ws->front[1]);
  (void)hipFree(ws->leafc);
  (void)hipFree(ws->stacks);
  (void)hipFree(ws->counters);
  (void)hipFree(ws->wave_prof);
  (void)hipFree(ws->defer[0]);
  (void)hipFree(ws->defer[1]);
  if (ws->host_counters) (void)hipHostFree(ws->host_counters);
  for (int i = 0; i < EngineWorkspace::kOrdSlots; ++i) {
    if (ws->ord_used[i]) (void)hipEventSynchronize(ws->ord_ev[i]);
    if (ws->ord_ev[i]) (void)hipEventDestroy(ws->ord_ev[i]);
  }
  (void)hipFree(ws->ord_ctr);
  for (int i = 0; i < EngineWorkspace::kAnySlots; ++i) {
    if (ws->any_used[i]) (void)hipEventSynchronize(ws->any_ev[i]);
    if (ws->any_ev[i]) (void)hipEventDestroy(ws->any_ev[i]);
    (void)hipFree(ws->any_stk[i]);
  }
  (void)hipFree(ws->all_stk);
  if (ws->all_graph) (void)hipGraphExecDestroy(ws->all_graph);
  if (ws->cap_stream) (void)hipStreamDestroy(ws->cap_stream);
  for (auto& e : ws->ev)
    if (e) (void)hipEventDestroy(e);
  delete ws;
}

hipError_t engine_set_profiling(EngineWorkspace* ws, bool on) {
  if (!ws) return hipErrorInvalidValue;
  ws->profile = on;
  return hipSuccess;
}

const EngineReport& engine_report(const EngineWorkspace* ws) { return ws->report; }

namespace {

hipError_t read_counter(EngineWorkspace* ws, int idx, unsigned long long& v, hipStream_t s) {
  ENG_TRY(hipMemcpyAsync(ws->host_counters + idx, ws->counters + idx, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, s));
  ENG_TRY(hipStreamSynchronize(s));
  v = ws->host_counters[idx];
  return hipSuccess;
}

enum : int { kPhaseSetup = 0, kPhaseExpand, kPhaseLeaves, kPhaseDfs, kPhaseFixup };

// Profiling clock: an event pair around each timed launch sequence, summed
// per phase once the stream has passed them.  No-ops unless profiling.
hipError_t clock_begin(EngineWorkspace* ws, hipStream_t s, size_t& idx) {
  if (!ws->profile) return hipSuccess;
  while (ws->ev.size() < ws->ev_used + 2) {
    hipEvent_t e;
    ENG_TRY(hipEventCreate(&e));
    ws->ev.push_back(e);
  }
  idx = ws->ev_used;
  ws->ev_used += 2;
  return hipEventRecord(ws->ev[idx], s);
}

void clock_end(EngineWorkspace* ws, hipStream_t s, size_t idx, int phase) {
  if (!ws->profile) return;
  (void)hipEventRecord(ws->ev[idx + 1], s);
  ws->spans.emplace_back(phase, idx);
}

hipError_t clock_collect(EngineWorkspace* ws) {
  if (!ws->profile) return hipSuccess;
  EngineReport& r = ws->report;
  double* out[5] = {&r.setup_ms, &r.expand_ms, &r.leaves_ms, &r.dfs_ms, &r.fixup_ms};
  for (const auto& sp : ws->spans) {
    ENG_TRY(hipEventSynchronize(ws->ev[sp.second + 1]));
    float ms = 0;
    ENG_TRY(hipEventElapsedTime(&ms, ws->ev[sp.second], ws->ev[sp.second + 1]));
    *out[sp.first] += ms;
  }
  ws->spans.clear();
  ws->ev_used = 0;
  return hipSuccess;
}

// Load balance of the depth-first kernel from its per-wavefront (start, end,
// BV tests) records: mean busy time of the wavefronts that did work over the
// kernel's span, and when 50 % / 90 % / all of them had finished.
hipError_t summarize_waves(EngineWorkspace* ws, size_t nwaves, hipStream_t s) {
  std::vector<unsigned long long> h(3 * nwaves);
  ENG_TRY(hipMemcpyAsync(h.data(), ws->wave_prof, sizeof(unsigned long long) * h.size(),
                         hipMemcpyDeviceToHost, s));
  ENG_TRY(hipStreamSynchronize(s));
  int dev = 0, khz = 0;
  ENG_TRY(hipGetDevice(&dev));
  ENG_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return hipSuccess;
  const double ms_per_tick = 1.0 / khz;
  unsigned long long t0 = ~0ULL;
  std::vector<unsigned long long> ends;
  double busy = 0;
  for (size_t w = 0; w < nwaves; ++w) {
    if (h[3 * w + 2] == 0) continue;
    t0 = std::min(t0, h[3 * w]);
    ends.push_back(h[3 * w + 1]);
    busy += static_cast<double>(h[3 * w + 1] - h[3 * w]);
  }
  EngineReport& r = ws->report;
  if (ends.empty()) return hipSuccess;
  std::sort(ends.begin(), ends.end());
  const double span = static_cast<double>(ends.back() - t0);
  ws->busy_sum += busy;
  ws->busy_den += span * static_cast<double>(ends.size());
  r.dfs_wave_busy = ws->busy_den > 0 ? ws->busy_sum / ws->busy_den : 1.0;
  if (r.dfs_waves > 0) return hipSuccess;  // percentiles describe the first round
  r.dfs_waves = static_cast<int>(ends.size());
  auto at = [&](double f) {
    const size_t i = std::min(ends.size() - 1, static_cast<size_t>(f * (ends.size() - 1)));
    return static_cast<double>(ends[i] - t0) * ms_per_tick;
  };
  r.dfs_end_p50_ms = at(0.5);
  r.dfs_end_p90_ms = at(0.9);
  r.dfs_end_max_ms = at(1.0);
  return hipSuccess;
}

inline unsigned grid_for(long long n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

// Persistent grid of the leaf kernel: every block resident, none idle.
unsigned leaf_grid(long long n) {
  static const unsigned resident = [] {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, engine_leaves, kBlock, 0) !=
            hipSuccess ||
        occ < 1)
      occ = 1;
    return static_cast<unsigned>(device_cus()) * static_cast<unsigned>(occ);
  }();
  return std::max(1u, std::min(grid_for(n), resident));
}

// Re-run the flagged queries with the plain depth-first kernel.
hipError_t fixup_flagged(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                         unsigned long long nflag, hipStream_t s) {
  const int nq = L.num_queries;
  int32_t* list = nullptr;
  int32_t* qm = nullptr;
  double* tf = nullptr;
  int32_t* cnt = nullptr;
  long long* st = nullptr;
  int32_t* rec = nullptr;
  int32_t* kept = nullptr;
  long long* nrun = nullptr;
  hipError_t e = hipSuccess;
  const long long cap_rec = static_cast<long long>(nflag) * L.cap;
  unsigned long long nkept = 0;
  do {
    if ((e = hipMalloc(&list, sizeof(int32_t) * nflag)) != hipSuccess) break;
    if ((e = hipMalloc(&qm, sizeof(int32_t) * 2 * nflag)) != hipSuccess) break;
    if ((e = hipMalloc(&tf, sizeof(double) * 24 * nflag)) != hipSuccess) break;
    if ((e = hipMalloc(&cnt, sizeof(int32_t) * nflag)) != hipSuccess) break;
    if ((e = hipMalloc(&st, sizeof(long long) * 2 * nflag)) != hipSuccess) break;
    if ((e = hipMalloc(&rec, sizeof(int32_t) * 3 * std::max<long long>(cap_rec, 1))) != hipSuccess)
      break;
    if ((e = hipMalloc(&nrun, sizeof(long long))) != hipSuccess) break;
    set_u64<<<1, 1, 0, s>>>(ws->counters + 5, 0);
    engine_collect_flagged<<<grid_for(nq), kBlock, 0, s>>>(ws->qflag, nq, list, ws->counters + 5);
    engine_gather_flagged<<<grid_for(nflag), kBlock, 0, s>>>(list, static_cast<int>(nflag),
                                                             L.query_meshes, L.tf1, L.tf2, qm, tf,
                                                             tf + 12 * nflag);
    set_u64<<<1, 1, 0, s>>>(reinterpret_cast<unsigned long long*>(nrun), 0);
    CollideLaunch P = L;
    P.num_queries = static_cast<int32_t>(nflag);
    P.query_meshes = qm;
    P.tf1 = tf;
    P.tf2 = tf + 12 * nflag;
    P.counts = cnt;
    P.pairs = rec;
    P.pair_capacity = cap_rec;
    P.num_pairs = nrun;
    P.bv_tests = st;
    P.leaf_tests = st + nflag;
    if ((e = launch_collide_plain(P, stack_need, s)) != hipSuccess) break;
  } while (false);
  if (e == hipSuccess) {
    // records produced by the fast engine
    unsigned long long produced = 0;
    e = hipMemcpyAsync(&produced, L.num_pairs, sizeof(produced), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const long long have = std::min<long long>(static_cast<long long>(produced), L.pair_capacity);
    if (e == hipSuccess && have > 0 && L.pairs) {
      e = hipMalloc(&kept, sizeof(int32_t) * 3 * have);
      if (e == hipSuccess) {
        set_u64<<<1, 1, 0, s>>>(ws->counters + 6, 0);
        engine_drop_flagged<<<grid_for(have), kBlock, 0, s>>>(L.pairs, have, ws->qflag, kept,
                                                              ws->counters + 6);
        e = read_counter(ws, 6, nkept, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(L.pairs, kept, sizeof(int32_t) * 3 * nkept, hipMemcpyDeviceToDevice,
                             s);
      }
    }
    unsigned long long nplain = 0;
    if (e == hipSuccess) {
      e = hipMemcpyAsync(&nplain, nrun, sizeof(nplain), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e == hipSuccess) {
      engine_scatter_flagged<<<grid_for(nflag), kBlock, 0, s>>>(
          list, static_cast<int>(nflag), cnt, st, st + nflag, L.counts, L.bv_tests,
          L.leaf_tests);
      const unsigned long long dropped = static_cast<unsigned long long>(have) - nkept;
      const unsigned long long total = produced - dropped + nplain;
      if (nplain > 0 && L.pairs)
        engine_append_records<<<grid_for(nplain), kBlock, 0, s>>>(
            rec, static_cast<long long>(nplain), list, L.pairs, static_cast<long long>(nkept),
            L.pair_capacity, L.status);
      set_u64<<<1, 1, 0, s>>>(reinterpret_cast<unsigned long long*>(L.num_pairs), total);
      e = hipStreamSynchronize(s);
    }
  }
  (void)hipFree(list);
  (void)hipFree(qm);
  (void)hipFree(tf);
  (void)hipFree(cnt);
  (void)hipFree(st);
  (void)hipFree(rec);
  (void)hipFree(kept);
  (void)hipFree(nrun);
  return e;
}

// Device buffers of one fix-up, freed together.
struct Scratch {
  std::vector<void*> ptrs;
  template <typename T>
  hipError_t alloc(T*& p, size_t n) {
    p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1));
    if (e == hipSuccess) ptrs.push_back(p);
    return e;
  }
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// Truncate each flagged query's recorded contact set to its first `cap`
// pairs in depth-first order and rewrite the output buffer in place.  Work is
// proportional to the flagged records F (one scan of the buffer finds them):
// path keys (engine_leaf_keys), a radix sort by key then a stable one by
// query, and the freed slots below the new end are refilled with the
// truncated lists and with the unflagged records from past the new end.
hipError_t fixup_keys(EngineWorkspace* ws, const CollideLaunch& L, unsigned long long nflag,
                      hipStream_t s) {
  const int nq = L.num_queries;
  const int nf = static_cast<int>(nflag);
  unsigned long long produced = 0;
  ENG_TRY(hipMemcpyAsync(&produced, L.num_pairs, sizeof(produced), hipMemcpyDeviceToHost, s));
  ENG_TRY(hipStreamSynchronize(s));
  Scratch S;
  int32_t *list, *seg_count, *take;
  long long *seg_off, *out_off;
  ENG_TRY(S.alloc(list, nf));
  ENG_TRY(S.alloc(seg_count, nf));
  ENG_TRY(S.alloc(take, nf));
  ENG_TRY(S.alloc(seg_off, nf));
  ENG_TRY(S.alloc(out_off, nf));
  set_u64<<<1, 1, 0, s>>>(ws->counters + 5, 0);
  engine_collect_flagged<<<grid_for(nq), kBlock, 0, s>>>(ws->qflag, nq, list, ws->counters + 5);
  engine_gather_counts<<<grid_for(nf), kBlock, 0, s>>>(list, nf, L.counts, seg_count);
  std::vector<int32_t> h_count(nf), h_take(nf);
  std::vector<long long> h_off(nf), h_out(nf);
  ENG_TRY(hipMemcpyAsync(h_count.data(), seg_count, sizeof(int32_t) * nf, hipMemcpyDeviceToHost,
                         s));
  ENG_TRY(hipStreamSynchronize(s));
  long long F = 0, T = 0;
  for (int i = 0; i < nf; ++i) {
    h_off[i] = F;
    F += h_count[i];
    h_take[i] = std::min(h_count[i], L.cap);
    h_out[i] = T;
    T += h_take[i];
  }
  if (static_cast<long long>(produced) > L.pair_capacity || !L.pairs) {
    // records were dropped (capacity status already set): report capped
    // counts and leave the buffer for the caller's retry
    ENG_TRY(hipMemcpyAsync(take, h_take.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
    engine_set_counts<<<grid_for(nf), kBlock, 0, s>>>(list, take, nf, L.counts);
    ENG_TRY(hipStreamSynchronize(s));
    return hipGetLastError();
  }
  if (F >= (1LL << 31)) return hipErrorInvalidValue;  // radix sort item count is int
  const long long N = static_cast<long long>(produced);
  const long long Nn = N - F + T;  // new end of the buffer
  ENG_TRY(hipMemcpyAsync(seg_off, h_off.data(), sizeof(long long) * nf, hipMemcpyHostToDevice,
                         s));
  ENG_TRY(hipMemcpyAsync(out_off, h_out.data(), sizeof(long long) * nf, hipMemcpyHostToDevice,
                         s));
  ENG_TRY(hipMemcpyAsync(take, h_take.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
  long long *pos, *pos_a, *pos_b, *holes, *tail;
  unsigned long long *keys, *keys_sorted;
  uint32_t *seg, *seg_sorted;
  ENG_TRY(S.alloc(pos, F));
  ENG_TRY(S.alloc(pos_a, F));
  ENG_TRY(S.alloc(pos_b, F));
  ENG_TRY(S.alloc(keys, F));
  ENG_TRY(S.alloc(keys_sorted, F));
  ENG_TRY(S.alloc(seg, F));
  ENG_TRY(S.alloc(seg_sorted, F));
  ENG_TRY(S.alloc(holes, F));
  ENG_TRY(S.alloc(tail, F));
  // 1. where the flagged records are
  set_u64<<<1, 1, 0, s>>>(ws->counters + 6, 0);
  engine_find_flagged<<<grid_for(N), kBlock, 0, s>>>(L.pairs, N, ws->qflag, pos, ws->counters + 6);
  unsigned long long found = 0;
  ENG_TRY(read_counter(ws, 6, found, s));
  if (static_cast<long long>(found) != F) return hipErrorUnknown;  // counts disagree with records
  // 2. their depth-first keys, ordered within each query
  if (F > 0) {
    engine_leaf_keys<<<grid_for(F), kBlock, 0, s>>>(L.pairs, pos, F, L, keys);
    int seg_bits = 1;
    while ((1 << seg_bits) < nf) ++seg_bits;
    size_t b1 = 0, b2 = 0;
    ENG_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, keys, keys_sorted, pos, pos_a,
                                               static_cast<int>(F), 0, 64, s));
    ENG_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b2, seg, seg_sorted, pos_a, pos_b,
                                               static_cast<int>(F), 0, seg_bits, s));
    unsigned char* tmp;
    ENG_TRY(S.alloc(tmp, std::max(b1, b2)));
    size_t bt = std::max(b1, b2);
    ENG_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, bt, keys, keys_sorted, pos, pos_a,
                                               static_cast<int>(F), 0, 64, s));
    // segment ids follow the key order: regather them for the stable sort
    engine_segment_of<<<grid_for(F), kBlock, 0, s>>>(L.pairs, pos_a, F, ws->qflag, seg);
    bt = std::max(b1, b2);
    ENG_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, bt, seg, seg_sorted, pos_a, pos_b,
                                               static_cast<int>(F), 0, seg_bits, s));
  }
  // 3. slots freed below the new end, and the unflagged records beyond it
  set_u64<<<1, 1, 0, s>>>(ws->counters + 6, 0);
  set_u64<<<1, 1, 0, s>>>(ws->counters + 7, 0);
  if (F > 0) engine_holes<<<grid_for(F), kBlock, 0, s>>>(pos, F, Nn, holes, ws->counters + 6);
  if (N > Nn)
    engine_tail_keep<<<grid_for(N - Nn), kBlock, 0, s>>>(L.pairs, Nn, N, ws->qflag, tail,
                                                         ws->counters + 7);
  unsigned long long nholes = 0, ntail = 0;
  ENG_TRY(read_counter(ws, 6, nholes, s));
  ENG_TRY(read_counter(ws, 7, ntail, s));
  if (static_cast<long long>(nholes) != T + static_cast<long long>(ntail))
    return hipErrorUnknown;  // bookkeeping invariant of the in-place rewrite
  // 4. stage every record that moves, then write it into a freed slot; the
  // slots are taken in ascending order so each truncated list stays in
  // depth-first order in the buffer
  if (nholes > 0) {
    long long* holes_sorted;
    ENG_TRY(S.alloc(holes_sorted, nholes));
    int pos_bits = 1;
    while (pos_bits < 63 && (1LL << pos_bits) < Nn) ++pos_bits;
    size_t bh = 0;
    ENG_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bh, holes, holes_sorted,
                                              static_cast<int>(nholes), 0, pos_bits, s));
    unsigned char* tmp_h;
    ENG_TRY(S.alloc(tmp_h, bh));
    ENG_TRY(hipcub::DeviceRadixSort::SortKeys(tmp_h, bh, holes, holes_sorted,
                                              static_cast<int>(nholes), 0, pos_bits, s));
    int32_t* stage;
    ENG_TRY(S.alloc(stage, 3 * nholes));
    engine_stage_fill<<<grid_for(nholes), kBlock, 0, s>>>(L.pairs, pos_b, seg_off, out_off, take,
                                                          nf, tail, static_cast<long long>(ntail),
                                                          T, stage);
    engine_fill_holes<<<grid_for(nholes), kBlock, 0, s>>>(stage, holes_sorted,
                                                          static_cast<long long>(nholes), L.pairs);
  }
  engine_set_counts<<<grid_for(nf), kBlock, 0, s>>>(list, take, nf, L.counts);
  set_u64<<<1, 1, 0, s>>>(reinterpret_cast<unsigned long long*>(L.num_pairs),
                          static_cast<unsigned long long>(Nn));
  ENG_TRY(hipStreamSynchronize(s));
  return hipGetLastError();
}

// Launch shape of the boolean kernel for a batch: slots per wavefront (K),
// register budget (waves per SIMD), step width, stack entries per slot and
// the grid.
struct AnyShape {
  int K, wps, width, slot_cap, wide_cap;
  unsigned grid;
  // 8-byte words of the work area: range counters, then one stack per slot
  size_t words() const {
    return kAnyAreaHead + static_cast<size_t>(grid) * kAnyWaves * K * slot_cap;
  }
};

AnyShape anyhit_shape(const EngineTuning& tu, long long nq, int stack_need) {
  AnyShape a;
  // slots per wavefront: the tunable (8 by default: under serving load,
  // where batches of several streams share the device, 8 beat 4 for small
  // batches too), or with 0 fewer for small batches, so that a batch alone
  // still spreads over many wavefronts
  a.K = tu.bool_queries_per_wave > 0 ? tu.bool_queries_per_wave : nq < kAnySmallBatch ? 4 : 8;
  a.wps = tu.bool_waves_per_simd;
  a.width = std::max(tu.bool_width, kWave / a.K);
  // as many wavefronts as fit at once (slots take queries from a counter),
  // fewer for small batches
  const long long nw = std::max<long long>(
      1, std::min<long long>((nq + a.K - 1) / a.K, static_cast<long long>(device_cus()) * 4 * a.wps));
  a.grid = static_cast<unsigned>((nw + kAnyWaves - 1) / kAnyWaves);
  // A slot's stack stays sorted by BVTT level (the children of a step's
  // tasks go on top, in order), and a step takes at most `width` tasks from
  // its top, so it never holds more than 2 width entries of one level;
  // children of unsure pairs come back later, hence the queue-sized margin.
  a.slot_cap = std::max(kAnySlotCap, 4 * (stack_need + 2) + 2 * kAnyQueue);
  a.wide_cap = a.slot_cap - (stack_need + 2) - 2 * kAnyQueue;
  return a;
}

// The boolean kernel over `area` (a.words() words; its range counters are
// cleared here by a kernel, not a memset, so the sequence can be captured).
// qflag / nflag: the exhaustive mode's flags (null for booleans).
void anyhit_launch(const AnyShape& a, const EngineTuning& tu, const CollideLaunch& L,
                   unsigned long long* area, unsigned long long* prof,
                   unsigned long long* wave_prof, hipStream_t s, int32_t* qflag = nullptr,
                   unsigned long long* nflag = nullptr) {
  zero_u64<<<1, kAnyAreaHead, 0, s>>>(area, kAnyAreaHead);
  const int lb = tu.bool_leaf_batch;
  const int rf = tu.bool_refill;
  const int K = a.K, wps = a.wps;
#define FCLX_ANYHIT(k, w)                                                                     \
  (qflag ? engine_anyhit<k, w, true><<<a.grid, kBlock, 0, s>>>(                               \
               L, area, area + kAnyAreaHead, a.slot_cap, a.wide_cap, a.width, lb, rf, prof,   \
               wave_prof, qflag, nflag)                                                       \
         : engine_anyhit<k, w, false><<<a.grid, kBlock, 0, s>>>(                              \
               L, area, area + kAnyAreaHead, a.slot_cap, a.wide_cap, a.width, lb, rf, prof,   \
               wave_prof, qflag, nflag))
  if (K == 4)
    wps >= 4 ? FCLX_ANYHIT(4, 4) : wps == 3 ? FCLX_ANYHIT(4, 3) : FCLX_ANYHIT(4, 2);
  else if (K == 8)
    wps >= 4 ? FCLX_ANYHIT(8, 4) : wps == 3 ? FCLX_ANYHIT(8, 3) : FCLX_ANYHIT(8, 2);
  else if (K == 16)
    wps >= 4 ? FCLX_ANYHIT(16, 4) : wps == 3 ? FCLX_ANYHIT(16, 3) : FCLX_ANYHIT(16, 2);
  else
    wps >= 3 ? FCLX_ANYHIT(32, 3) : FCLX_ANYHIT(32, 2);
#undef FCLX_ANYHIT
}

}  // namespace

hipError_t engine_anyhit_collide_area(const EngineTuning& tu, const CollideLaunch& L,
                                      int stack_need, unsigned long long* area,
                                      size_t area_words, hipStream_t s) {
  if (!area) return hipErrorInvalidValue;
  if (L.num_queries == 0) return hipSuccess;
  AnyShape a = anyhit_shape(tu, L.num_queries, stack_need);
  // whole workgroups of stacks that fit the area (slots take the queries
  // from the counters, so fewer wavefronts only take longer)
  const size_t per_block = static_cast<size_t>(kAnyWaves) * a.K * a.slot_cap;
  const size_t fit = area_words > kAnyAreaHead ? (area_words - kAnyAreaHead) / per_block : 0;
  if (fit == 0) return hipErrorInvalidValue;
  a.grid = static_cast<unsigned>(std::min<size_t>(a.grid, fit));
  anyhit_launch(a, tu, L, area, nullptr, nullptr, s);
  return hipGetLastError();
}

size_t engine_anyhit_min_words(const EngineTuning& tu, long long num_queries, int stack_need) {
  AnyShape a = anyhit_shape(tu, num_queries, stack_need);
  a.grid = 1;
  return a.words();
}

hipError_t engine_anyhit_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                                 hipStream_t s) {
  if (!ws) return hipErrorInvalidValue;
  const long long nq = L.num_queries;
  if (nq == 0) return hipSuccess;
  const int cus = device_cus();
  const AnyShape a = anyhit_shape(ws->tune, nq, stack_need);
  const int K = a.K, wps = a.wps, slot_cap = a.slot_cap;
  const unsigned grid = a.grid;
  const size_t need = a.words();
  // the next area in turn whose last batch has finished (an area still in
  // use makes this batch wait on the device for that batch, on whatever
  // stream it runs); if every area is busy, the next in turn
  int slot = ws->any_next;
  for (int i = 0; i < EngineWorkspace::kAnySlots; ++i) {
    const int c = (ws->any_next + i) % EngineWorkspace::kAnySlots;
    if (!ws->any_used[c] || hipEventQuery(ws->any_ev[c]) == hipSuccess) {
      slot = c;
      break;
    }
  }
  ws->any_next = (slot + 1) % EngineWorkspace::kAnySlots;
  if (need > ws->any_cap[slot]) {
    // Every area of the ring grows at once, to what any batch with these
    // meshes and settings can need (all the wavefronts that fit, 8 slots
    // each at least): freeing device memory waits for the whole device, so
    // it should not happen again while batches of other streams run.
    const size_t full = std::max(
        need, kAnyAreaHead + static_cast<size_t>(cus) * 4 * std::max(wps, 4) * std::max(K, 8) *
                                 static_cast<size_t>(slot_cap));
    for (int i = 0; i < EngineWorkspace::kAnySlots; ++i) {
      if (ws->any_cap[i] >= full) continue;
      if (ws->any_used[i]) ENG_TRY(hipEventSynchronize(ws->any_ev[i]));
      (void)hipFree(ws->any_stk[i]);
      ws->any_stk[i] = nullptr;
      ws->any_cap[i] = 0;
      ENG_TRY(hipMalloc(reinterpret_cast<void**>(&ws->any_stk[i]),
                        full * sizeof(unsigned long long)));
      ws->any_cap[i] = full;
    }
  }
  if (ws->any_used[slot]) ENG_TRY(hipStreamWaitEvent(s, ws->any_ev[slot], 0));
  unsigned long long* area = ws->any_stk[slot];
  unsigned long long* prof = nullptr;
  if (ws->profile) {  // counters of this batch (the batch is waited for)
    prof = ws->counters + kCtrWork;
    zero_u64<<<1, 64, 0, s>>>(prof, 9);
    ENG_TRY(grow(ws->wave_prof, ws->wave_prof_cap, 3 * static_cast<size_t>(grid) * kAnyWaves));
  }
  anyhit_launch(a, ws->tune, L, area, prof, ws->wave_prof, s);
  ENG_TRY(hipGetLastError());
  ENG_TRY(hipEventRecord(ws->any_ev[slot], s));
  ws->any_used[slot] = true;
  if (prof) {
    unsigned long long h[9];
    const size_t nwv = static_cast<size_t>(grid) * kAnyWaves;
    std::vector<unsigned long long> wv(3 * nwv);
    ENG_TRY(hipMemcpyAsync(h, prof, sizeof(h), hipMemcpyDeviceToHost, s));
    ENG_TRY(hipMemcpyAsync(wv.data(), ws->wave_prof, sizeof(unsigned long long) * wv.size(),
                           hipMemcpyDeviceToHost, s));
    ENG_TRY(hipStreamSynchronize(s));
    int dev = 0, khz = 1;
    ENG_TRY(hipGetDevice(&dev));
    ENG_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    EngineReport& r = ws->report;
    unsigned long long t0 = ~0ULL;
    double life = 0;
    std::vector<unsigned long long> ends(nwv);
    for (size_t w = 0; w < nwv; ++w) {
      t0 = std::min(t0, wv[3 * w]);
      ends[w] = wv[3 * w + 1];
      life += static_cast<double>(wv[3 * w + 1] - wv[3 * w]);
    }
    std::sort(ends.begin(), ends.end());
    auto at = [&](double f) {
      return static_cast<double>(ends[std::min(nwv - 1, static_cast<size_t>(f * (nwv - 1)))] - t0) /
             khz;
    };
    r.bool_wave_ms = life / static_cast<double>(nwv) / khz;
    r.bool_end_p50_ms = at(0.5);
    r.bool_end_p90_ms = at(0.9);
    r.bool_span_ms = at(1.0);
    for (int k = 0; k < 3; ++k) {
      r.bool_steps[k] = static_cast<long long>(h[k]);
      r.bool_lanes[k] = static_cast<long long>(h[3 + k]);
    }
    r.bool_skipped = static_cast<long long>(h[6]);
    r.bool_rounds = static_cast<long long>(h[7]);
    r.bool_stack_max = static_cast<long long>(h[8]);
  }
  return hipSuccess;
}

// Exhaustive collide as one persistent launch fed by the query counters:
// the exhaustive form of the boolean kernel (per-query slots, FP32 cull,
// FP64 leaf and uncertain steps inline, every hit recorded), then the
// depth-first prefix fix-up of the queries that reached the cap (only when
// some did; it needs the host).
hipError_t engine_persistent_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                                     hipStream_t s) {
  const int nq = L.num_queries;
  unsigned long long* C = ws->counters;
  EngineReport& rep = ws->report;
  rep = EngineReport{};
  ws->spans.clear();
  ws->ev_used = 0;
  ENG_TRY(grow(ws->qflag, ws->qflag_cap, nq));
  const AnyShape a = anyhit_shape(ws->tune, nq, stack_need);
  ENG_TRY(grow(ws->all_stk, ws->all_cap, a.words()));
  unsigned long long* prof = nullptr;
  if (ws->profile) {
    prof = C + kCtrWork;
    ENG_TRY(grow(ws->wave_prof, ws->wave_prof_cap, 3 * static_cast<size_t>(a.grid) * kAnyWaves));
  }
  // the launch sequence: flags and counters cleared by kernels (no memset
  // node, so the sequence can be captured), then the persistent kernel
  auto enqueue = [&](hipStream_t q) {
    zero_i32_grid<<<grid_for(nq), kBlock, 0, q>>>(ws->qflag, nq);
    set_u64<<<1, 1, 0, q>>>(C + kCtrFlagged, 0);
    if (prof) zero_u64<<<1, 64, 0, q>>>(prof, 9);
    anyhit_launch(a, ws->tune, L, ws->all_stk, prof, ws->wave_prof, q, ws->qflag,
                  C + kCtrFlagged);
  };
  size_t t = 0;
  ENG_TRY(clock_begin(ws, s, t));
  if (ws->tune.fast_graph && !prof) {
    // replayed from a graph while every argument of the sequence is the same
    const CollideLaunch* Lp = &L;
    std::vector<unsigned long long> key(sizeof(CollideLaunch) / 8 + 12);
    std::memcpy(key.data(), Lp, sizeof(CollideLaunch) / 8 * 8);
    size_t k = sizeof(CollideLaunch) / 8;
    key[k++] = reinterpret_cast<uintptr_t>(ws->all_stk);
    key[k++] = reinterpret_cast<uintptr_t>(ws->qflag);
    key[k++] = static_cast<unsigned long long>(a.grid);
    key[k++] = static_cast<unsigned long long>(a.K);
    key[k++] = static_cast<unsigned long long>(a.wps);
    key[k++] = static_cast<unsigned long long>(a.width);
    key[k++] = static_cast<unsigned long long>(a.slot_cap);
    key[k++] = static_cast<unsigned long long>(a.wide_cap);
    key[k++] = static_cast<unsigned long long>(ws->tune.bool_leaf_batch);
    key[k++] = static_cast<unsigned long long>(ws->tune.bool_refill);
    key[k++] = static_cast<unsigned long long>(nq);
    key[k++] = reinterpret_cast<uintptr_t>(C);
    if (!ws->all_graph || key != ws->all_key) {
      if (!ws->cap_stream)
        ENG_TRY(hipStreamCreateWithFlags(&ws->cap_stream, hipStreamNonBlocking));
      if (ws->all_graph) {
        ENG_TRY(hipGraphExecDestroy(ws->all_graph));
        ws->all_graph = nullptr;
      }
      hipGraph_t g = nullptr;
      ENG_TRY(hipStreamBeginCapture(ws->cap_stream, hipStreamCaptureModeThreadLocal));
      enqueue(ws->cap_stream);
      const hipError_t ec = hipStreamEndCapture(ws->cap_stream, &g);
      if (ec != hipSuccess) return ec;
      const hipError_t ei = hipGraphInstantiate(&ws->all_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        ws->all_graph = nullptr;
        return ei;
      }
      ws->all_key = key;
    }
    ENG_TRY(hipGraphLaunch(ws->all_graph, s));
  } else {
    enqueue(s);
  }
  ENG_TRY(hipGetLastError());
  clock_end(ws, s, t, kPhaseDfs);
  unsigned long long nflag = 0;
  ENG_TRY(read_counter(ws, kCtrFlagged, nflag, s));
  if (nflag > 0) {
    ENG_TRY(clock_begin(ws, s, t));
    if (L.bv_tests || L.leaf_tests || stack_need > 64)
      ENG_TRY(fixup_flagged(ws, L, stack_need, nflag, s));
    else
      ENG_TRY(fixup_keys(ws, L, nflag, s));
    clock_end(ws, s, t, kPhaseFixup);
  }
  rep.flagged = static_cast<int>(nflag);
  rep.front_size = nq;
  rep.dfs_waves = static_cast<int>(a.grid) * kAnyWaves;
  rep.dfs_rounds = 1;
  if (prof) {
    unsigned long long h[9];
    ENG_TRY(hipMemcpyAsync(h, prof, sizeof(h), hipMemcpyDeviceToHost, s));
    ENG_TRY(hipStreamSynchronize(s));
    rep.dfs_bv_tests = static_cast<long long>(h[3]);    // FP32 BV lanes
    rep.dfs_leaf_tests = static_cast<long long>(h[4]);  // triangle tests
    rep.fp64_fallbacks = static_cast<long long>(h[5]);  // FP64 BV tests
    for (int k = 0; k < 3; ++k) {
      rep.bool_steps[k] = static_cast<long long>(h[k]);
      rep.bool_lanes[k] = static_cast<long long>(h[3 + k]);
    }
  }
  ENG_TRY(clock_collect(ws));
  return hipSuccess;
}

hipError_t engine_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                          hipStream_t s) {
  if (!ws) return hipErrorInvalidValue;
  const int nq = L.num_queries;
  if (nq == 0) return hipSuccess;
  const int cus = device_cus();
  if (L.cap <= kOrderedCapMax) {
    const int slot = ws->ord_next;
    ws->ord_next = (slot + 1) % EngineWorkspace::kOrdSlots;
    if (ws->ord_used[slot]) ENG_TRY(hipStreamWaitEvent(s, ws->ord_ev[slot], 0));
    unsigned long long* ctr = ws->ord_ctr + 16 * slot;
    set_u64<<<1, 1, 0, s>>>(ctr, 0);
    const unsigned grid = static_cast<unsigned>(
        std::min<long long>(grid_for(nq), static_cast<long long>(cus) * 8));
    if (stack_need <= 64)
      ordered_collide<64><<<grid, kBlock, 0, s>>>(L, ctr);
    else if (stack_need <= 128)
      ordered_collide<128><<<grid, kBlock, 0, s>>>(L, ctr);
    else
      ordered_collide<256><<<grid, kBlock, 0, s>>>(L, ctr);
    ENG_TRY(hipGetLastError());
    ENG_TRY(hipEventRecord(ws->ord_ev[slot], s));
    ws->ord_used[slot] = true;
    return hipSuccess;
  }
  if (nq >= (1 << 22)) return hipErrorInvalidValue;  // task packing limit
  if (ws->tune.fast_persistent) return engine_persistent_collide(ws, L, stack_need, s);
  const EngineTuning& tune = ws->tune;
  ENG_TRY(grow(ws->qx32, ws->qx32_cap, nq));
  ENG_TRY(grow(ws->qptr, ws->qptr_cap, nq));
  ENG_TRY(grow(ws->qx64, ws->qx64_cap, nq));
  ENG_TRY(grow(ws->qflag, ws->qflag_cap, nq));
  ENG_TRY(grow(ws->front[0], ws->front_cap[0], std::max<size_t>(nq, 1024)));
  unsigned long long* C = ws->counters;

  EngineArgs A{};
  A.L = L;
  A.qx32 = ws->qx32;
  A.qptr = ws->qptr;
  A.qx64 = ws->qx64;
  A.qflag = ws->qflag;
  A.nflag = C + kCtrFlagged;
  A.work = C + kCtrWork;
  zero_u64<<<1, kNumCounters, 0, s>>>(C, kNumCounters);
  EngineReport& rep = ws->report;
  rep = EngineReport{};
  ws->busy_sum = ws->busy_den = 0;
  ws->spans.clear();
  ws->ev_used = 0;

  // 1. setup -> front[0]
  size_t t = 0;
  ENG_TRY(clock_begin(ws, s, t));
  A.front_out = ws->front[0];
  A.n_out = C + kCtrSetupOut;
  engine_setup<<<grid_for(nq), kBlock, 0, s>>>(A);
  ENG_TRY(hipGetLastError());
  clock_end(ws, s, t, kPhaseSetup);
  unsigned long long nfront = 0;
  ENG_TRY(read_counter(ws, kCtrSetupOut, nfront, s));
  int cur = 0;

  // Breadth-first expansion of front[cur] while it is small; leaf pairs of
  // these levels are tested right after.
  auto expand = [&]() -> hipError_t {
    unsigned long long nleafc = 0;
    ENG_TRY(clock_begin(ws, s, t));
    for (int level = 0; level < tune.max_levels && nfront > 0 &&
                        static_cast<long long>(nfront) < tune.front_target;
         ++level) {
      const int nxt = cur ^ 1;
      ENG_TRY(grow(ws->front[nxt], ws->front_cap[nxt], 2 * nfront));
      ENG_TRY(grow_keep(ws->leafc, ws->leafc_cap, nleafc + nfront, nleafc, s));
      EngineArgs X = A;
      X.front_in = ws->front[cur];
      X.n_in = static_cast<long long>(nfront);
      X.front_out = ws->front[nxt];
      X.n_out = C + kCtrExpandOut;
      X.leafc = ws->leafc;
      X.n_leafc = C + kCtrLeafc;
      set_u64<<<1, 1, 0, s>>>(C + kCtrExpandOut, 0);
      engine_expand<<<grid_for(nfront), kBlock, 0, s>>>(X);
      ENG_TRY(hipGetLastError());
      ENG_TRY(read_counter(ws, kCtrExpandOut, nfront, s));
      ENG_TRY(read_counter(ws, kCtrLeafc, nleafc, s));
      cur = nxt;
      ++rep.levels;
    }
    clock_end(ws, s, t, kPhaseExpand);
    if (nleafc > 0) {
      ENG_TRY(clock_begin(ws, s, t));
      EngineArgs X = A;
      X.in.count = 1;
      X.in.p[0] = ws->leafc;
      X.in.start[0] = 0;
      X.in.start[1] = static_cast<long long>(nleafc);
      X.leaf_work = C + kCtrWork + kWorkLeafKernel;
      engine_leaves<<<leaf_grid(nleafc), kBlock, 0, s>>>(X);
      ENG_TRY(hipGetLastError());
      clock_end(ws, s, t, kPhaseLeaves);
      set_u64<<<1, 1, 0, s>>>(C + kCtrLeafc, 0);
    }
    return hipSuccess;
  };

  // 2. breadth-first expansion
  ENG_TRY(expand());
  rep.front_size = static_cast<long long>(nfront);

  // 3. depth-first rounds over the remaining front
  // register budget of the depth-first kernel: 4, 5 or 6 wavefronts per SIMD
  // (more waves hide more latency; fewer registers spill more)
  // the inline triangle test needs the 128 registers of 4 wavefronts per SIMD
  // (without per-query statistics, the inline form runs the variant that
  // leaves their counting out)
  auto dfs_kernel = tune.dfs_waves_per_simd >= 6   ? engine_dfs<6, false>
                    : tune.dfs_waves_per_simd == 5 ? engine_dfs<5, false>
                    : tune.dfs_inline_leaves       ? (L.bv_tests ? engine_dfs<4, true>
                                                                 : engine_dfs<4, true, false>)
                                                   : engine_dfs<4, false>;
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, dfs_kernel, kBlock, 0) != hipSuccess ||
      occ < 1)
    occ = 1;
  const unsigned blocks =
      static_cast<unsigned>(cus) * static_cast<unsigned>(std::min(occ, tune.dfs_blocks_per_cu));
  const size_t nwaves = static_cast<size_t>(blocks) * kWavesPerBlock;
  const int stack_cap = 2 * kWave * (stack_need + 2) + 2 * kWave;
  // entries per sub-buffer of each deferred buffer
  const long long dsub[2] = {std::max<long long>(kLeafQ, tune.defer_leaf / kSub),
                             std::max<long long>(kUncQ, tune.defer_uncertain / kSub)};
  while (nfront > 0) {
    if (rep.dfs_rounds >= 1000000) return hipErrorUnknown;  // no progress: never expected
    const int nxt = cur ^ 1;
    ENG_TRY(grow(ws->stacks, ws->stacks_cap, nwaves * stack_cap));
    // next front: stack dumps, unconsumed input, children of uncertain tasks
    ENG_TRY(grow(ws->front[nxt], ws->front_cap[nxt], nwaves * stack_cap + nfront));
    for (int w = 0; w < 2; ++w)
      ENG_TRY(grow(ws->defer[w], ws->defer_cap[w], kSub * dsub[w] + nwaves * kQueueCap[w]));
    if (ws->profile) ENG_TRY(grow(ws->wave_prof, ws->wave_prof_cap, 3 * nwaves));
    zero_u64<<<1, kNumCounters, 0, s>>>(C + kCtrRoundBegin, kCtrRoundEnd - kCtrRoundBegin);
    set_u64<<<1, 1, 0, s>>>(C + kCtrHead, 0);
    EngineArgs D = A;
    D.front_in = ws->front[cur];
    D.n_in = static_cast<long long>(nfront);
    D.front_out = ws->front[nxt];
    D.n_out = C + kCtrNext;
    D.head = C + kCtrHead;
    D.stacks = ws->stacks;
    D.stack_cap = stack_cap;
    D.stack_top = tune.dfs_stack_top;
    D.wave_prof = ws->profile ? ws->wave_prof : nullptr;
    for (int w = 0; w < 2; ++w) {
      D.defer[w] = ws->defer[w];
      D.defer_sub[w] = dsub[w];
    }
    D.ctr = C;
    D.stop = C + kCtrStop;
    D.stop_poll = static_cast<unsigned>(tune.stop_poll);
    D.idle = C + kCtrIdle;
    // rebalance only when the front gives every wavefront a full step
    D.idle_limit = static_cast<long long>(nfront) >= static_cast<long long>(kWave * nwaves)
                       ? std::max<unsigned long long>(
                             1, static_cast<unsigned long long>(tune.idle_stop * nwaves))
                       : 0;
    ENG_TRY(clock_begin(ws, s, t));
    dfs_kernel<<<blocks, kBlock, 0, s>>>(D);
    ENG_TRY(hipGetLastError());
    clock_end(ws, s, t, kPhaseDfs);
    ++rep.dfs_rounds;
    ENG_TRY(hipMemcpyAsync(ws->host_counters + kCtrHead, C + kCtrHead, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
    ENG_TRY(hipMemcpyAsync(ws->host_counters + kCtrRoundBegin, C + kCtrRoundBegin,
                           (kCtrRoundEnd - kCtrRoundBegin) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
    ENG_TRY(hipStreamSynchronize(s));
    const unsigned long long* H = ws->host_counters;
    const unsigned long long head = H[kCtrHead];
    unsigned long long nnext = H[kCtrNext];
    // what each deferred buffer holds: the filled part of every sub-buffer
    // (padding included) and the dumps
    TaskRanges in[2];
    for (int w = 0; w < 2; ++w) {
      in[w].count = 0;
      in[w].start[0] = 0;
      auto add = [&](const unsigned long long* p, long long n) {
        if (n <= 0) return;
        in[w].p[in[w].count] = p;
        in[w].start[in[w].count + 1] = in[w].start[in[w].count] + n;
        ++in[w].count;
      };
      for (int x = 0; x < kSub; ++x)
        add(ws->defer[w] + x * dsub[w],
            static_cast<long long>(
                std::min<unsigned long long>(H[kCtrResv + 16 * (w * kSub + x)], dsub[w])));
      add(ws->defer[w] + kSub * dsub[w], static_cast<long long>(H[kCtrDump + 16 * w]));
    }
    const long long n_leaf = in[0].start[in[0].count], n_unc = in[1].start[in[1].count];
    if (ws->profile) ENG_TRY(summarize_waves(ws, nwaves, s));
    // unconsumed input goes to the next front
    const unsigned long long rem = nfront - std::min(head, nfront);
    if (rem > 0) {
      ENG_TRY(hipMemcpyAsync(ws->front[nxt] + nnext, ws->front[cur] + (nfront - rem),
                             sizeof(unsigned long long) * rem, hipMemcpyDeviceToDevice, s));
      nnext += rem;
    }
    // FP64 work deferred by the round
    ENG_TRY(clock_begin(ws, s, t));
    if (n_unc > 0) {
      ENG_TRY(grow_keep(ws->front[nxt], ws->front_cap[nxt], nnext + 2 * n_unc, nnext, s));
      set_u64<<<1, 1, 0, s>>>(C + kCtrNext, nnext);
      EngineArgs X = A;
      X.in = in[1];
      X.front_out = ws->front[nxt];
      X.n_out = C + kCtrNext;
      X.leaf_work = C + kCtrWork + kWorkDfsLeaf;
      engine_uncertain<<<grid_for(n_unc), kBlock, 0, s>>>(X);
      ENG_TRY(hipGetLastError());
    }
    if (n_leaf > 0) {
      EngineArgs X = A;
      X.in = in[0];
      X.leaf_work = C + kCtrWork + kWorkDfsLeaf;
      engine_leaves<<<leaf_grid(n_leaf), kBlock, 0, s>>>(X);
      ENG_TRY(hipGetLastError());
    }
    clock_end(ws, s, t, kPhaseLeaves);
    if (n_unc > 0) ENG_TRY(read_counter(ws, kCtrNext, nnext, s));
    cur = nxt;
    nfront = nnext;
    // widen a front too small to give every wavefront a full step
    if (nfront > 0 && static_cast<long long>(nfront) < static_cast<long long>(kWave * nwaves))
      ENG_TRY(expand());
  }

  // 4. queries that reached the cap
  unsigned long long nflag = 0;
  ENG_TRY(read_counter(ws, kCtrFlagged, nflag, s));
  if (nflag > 0) {
    ENG_TRY(clock_begin(ws, s, t));
    // per-query statistics of a truncated traversal need the depth-first
    // re-run; otherwise the recorded sets are truncated by depth-first keys
    // (which fit 64 bits while the summed tree depths do)
    if (L.bv_tests || L.leaf_tests || stack_need > 64)
      ENG_TRY(fixup_flagged(ws, L, stack_need, nflag, s));
    else
      ENG_TRY(fixup_keys(ws, L, nflag, s));
    clock_end(ws, s, t, kPhaseFixup);
  }
  rep.flagged = static_cast<int>(nflag);
  ENG_TRY(hipMemcpyAsync(ws->host_counters + kCtrWork, C + kCtrWork,
                         7 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  ENG_TRY(hipStreamSynchronize(s));
  const unsigned long long* H = ws->host_counters + kCtrWork;
  rep.expand_bv_tests = static_cast<long long>(H[kWorkExpandBV]);
  rep.expand_leaf_tests = static_cast<long long>(H[kWorkLeafKernel]);
  rep.dfs_bv_tests = static_cast<long long>(H[kWorkDfsBV]);
  rep.dfs_leaf_tests = static_cast<long long>(H[kWorkDfsLeaf] + H[kWorkDfsInline]);
  rep.dfs_inline_leaf_tests = static_cast<long long>(H[kWorkDfsInline]);
  rep.dfs_wave_steps = static_cast<long long>(H[kWorkDfsSteps]);
  rep.fp64_fallbacks = static_cast<long long>(H[kWorkFallback]);
  ENG_TRY(clock_collect(ws));
  return hipSuccess;
}

}  // namespace fclx
