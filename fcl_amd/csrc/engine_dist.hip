// Wavefront-wide mesh distance (fclx_distance_batch with FCLX_ENGINE_FAST).
//
// distanceRecurse (traversal_recurse-inl.h:257-316) visits the BVTT depth
// first, the child of smaller RSS lower bound first, and prunes a pair
// whose bound is not below the best distance so far; the first pair that
// reaches the minimum is reported (DistanceResult::update keeps it on ties,
// distance_result-inl.h:84-105).  Here a wavefront keeps kSlots queries in
// flight, each with a LIFO stack in device memory, and a step computes up
// to 64 bounds across them (an even share of the lanes per query, more for
// deeper stacks), so each query descends nearly depth first while the lanes
// stay busy.  Leaf pairs are queued and their FP64 triangle distances taken
// 64 at a time.
//
// Every task carries its path from the root as a key: at depth k, bit 63-k
// is 0 for the child the reference visits first and 1 for the other, so the
// reference's visit order is the order of the keys.  A query's best is the
// smallest (distance, key) found, and a task is pruned when its bound is
// above the best distance, or equal to it with a key above the best key
// (all pairs of its subtree come later in the reference's order).  Hence
// the minimum and the pair reported are the reference's whatever the order
// of the work; the nearest points are those of that pair, recomputed when
// the query is done.
//
// With FP32 culling (the dist_fp32_cull tunable; off by default, FP64
// bounds throughout) bounds are first taken in FP32 (device/rss32.h: a
// certified lower bound from 64-B float nodes).  A child whose FP32 bound is above the best
// distance is pruned, as the FP64 bound would be.  When one child is left,
// it is pushed with its FP32 bound (a lower bound, so later pruning stays
// safe) and the key bit of its level left 0: the other child's subtree
// holds no pair within the best distance, so no comparison the bit could
// decide is ever made.  When both are left the order between them matters
// for the keys, so the task goes to a queue whose FP64 bounds (the
// reference's rectDistance on the 128-B FP64 nodes) are taken 64 at a time
// and decide the order and the pruning exactly as the FP64 form does; the
// task's two stack entries are reserved when it is queued (in lane order,
// where its children would have gone) and the slot takes no tasks until
// they are filled, so the stack keeps its depth-first shape and bound.
// On this hardware FP32 (unpacked) and FP64 vector instructions issue at
// the same rate, and a certified FP32 bound costs about as many
// instructions as the FP64 one, so the FP32 form does more work in all
// (measured on the 20,480-triangle workload: 28.0 ms against 15.3 ms per
// 10,000 queries) and is kept as an option.
// num_bv_tests counts this engine's bounds (FP32 and FP64 alike),
// num_leaf_tests its triangle distances.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/pose64.h"
#include "device/rss32.h"
#include "kernels.h"

namespace fclx {
namespace {

constexpr int kDWave = 64;
constexpr int kDBlock = 256;
constexpr int kDWaves = kDBlock / kDWave;
constexpr int kDQueue = 128;
constexpr int kDSlotCap = 1024;  // stack entries per query slot (at least)

// task: b1:21 | b2:21 | depth:6; bound; path key
struct DTask {
  unsigned long long t;
  double d;
  unsigned long long key;
};
__device__ __forceinline__ unsigned long long dpack(int b1, int b2, int depth) {
  return (static_cast<unsigned long long>(depth) << 42) |
         (static_cast<unsigned long long>(b1) << 21) | static_cast<unsigned long long>(b2);
}
// leaf queue entry: gen:17 | slot:5 | id1:21 | id2:21 (and the key beside it);
// an FP64-bound queue entry packs node ids b1, b2 the same way
constexpr int kDSlotBits = 5;
constexpr int kDGenMask = (1 << (22 - kDSlotBits)) - 1;
__device__ __forceinline__ unsigned long long lpack(int tag, int i1, int i2) {
  return (static_cast<unsigned long long>(tag) << 42) |
         (static_cast<unsigned long long>(i1) << 21) | static_cast<unsigned long long>(i2);
}

template <int kSlots>
struct DistWave {
  double pose[12][kSlots];  // relative pose R (9), T (3) of each slot, chunk-major
  float pose32[13][kSlots];  // the same in float, then the FP32 bound's margin term
  const RssNode64* rss1[kSlots];
  const RssNode64* rss2[kSlots];
  const RssNode32* rss1f[kSlots];
  const RssNode32* rss2f[kSlots];
  const double* size1[kSlots];
  const double* size2[kSlots];
  int mesh1[kSlots], mesh2[kSlots];
  int sp[kSlots], pend[kSlots], gen[kSlots];
  int upend[kSlots];  // the slot's tasks queued for FP64 bounds (the slot takes no
                      // tasks meanwhile: their children's entries are reserved)
  int offs[kSlots], base[kSlots];
  double best_d[kSlots];
  unsigned long long best_key[kSlots];
  int best1[kSlots], best2[kSlots];
  long long nbv[kSlots], nleaf[kSlots];
  unsigned long long lq[kDQueue];
  unsigned long long lk[kDQueue];
  // tasks whose two children both survived the FP32 bounds: (gen, slot,
  // b1, b2), the task's key and its depth
  unsigned long long uq[kDQueue];
  unsigned long long uk[kDQueue];
  int upos[kDQueue];  // the two stack entries reserved for the children
  unsigned char ud[kDQueue];
  // a leaf step's results, for the slots to pick their minimum
  double cd[kDWave];
  unsigned long long ck[kDWave];
  int ci1[kDWave], ci2[kDWave];
  signed char cslot[kDWave];
  unsigned char lane_slot[kDWave];
};

template <int n>
__device__ __forceinline__ int dscan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);
  if (n > 2) v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);
  if (n > 4) v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);
  if (n > 8) v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);
  return v;
}

__device__ __forceinline__ unsigned long long dbelow() {
  return (1ULL << __lane_id()) - 1ULL;
}

template <int kSlots>
__device__ __forceinline__ void slot_pose(const DistWave<kSlots>& W, int s, M3& R, V3& T) {
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = W.pose[i][s];
  T = v3(W.pose[9][s], W.pose[10][s], W.pose[11][s]);
}

// FP64 distance of triangle id1 of mesh m1 and triangle id2 of mesh m2 moved
// into m1's frame (meshDistanceOrientedNodeLeafTesting, as the plain kernel)
__device__ __forceinline__ double tri_pair(const MeshDev& m1, const MeshDev& m2, const M3& R,
                                           const V3& T, int id1, int id2, V3& P, V3& Q) {
  const int32_t* t1 = m1.tris + 3 * static_cast<size_t>(id1);
  const int32_t* t2 = m2.tris + 3 * static_cast<size_t>(id2);
  const V3 S[3] = {vertex(m1, t1[0]), vertex(m1, t1[1]), vertex(m1, t1[2])};
  const V3 Tt[3] = {mul(R, vertex(m2, t2[0])) + T, mul(R, vertex(m2, t2[1])) + T,
                    mul(R, vertex(m2, t2[2])) + T};
  return tri_distance64(S, Tt, P, Q);
}

template <int kSlots, int kWavesPerSimd, bool kF32>
__global__ void __launch_bounds__(kDBlock, kWavesPerSimd) engine_distance_wave(DistanceLaunch L,
                                                                   unsigned long long* next_q,
                                                                   DTask* stk, int slot_cap,
                                                                   int wide_cap, int width,
                                                                   int leaf_batch) {
  static_assert(kSlots <= 16 && kSlots <= (1 << kDSlotBits), "slots fit a DPP row");
  __shared__ DistWave<kSlots> waves_s[kDWaves];
  const int lane = __lane_id();
  const int wave = threadIdx.x / kDWave;
  DistWave<kSlots>& W = waves_s[wave];
  const long long gw = static_cast<long long>(blockIdx.x) * kDWaves + wave;
  DTask* const stacks = stk + gw * kSlots * static_cast<long long>(slot_cap);
  const bool slot_lane = lane < kSlots;
  DTask* const mine = stacks + static_cast<long long>(slot_lane ? lane : 0) * slot_cap;
  const int nq = L.num_queries;
  int q = -1, spv = 0;
  int nx = 0, ne = 0;
  bool exhausted = false, check = true;
  int nl = 0, nu = 0;
  if (slot_lane) W.gen[lane] = 0;

  for (;;) {
    if (check) {
      check = false;
      // finish queries with nothing left: stack empty, no leaf pair queued
      if (slot_lane && q >= 0) {
        spv = W.sp[lane];
        if (spv == 0 && W.pend[lane] == 0) {
          const MeshDev& m1 = L.meshes[W.mesh1[lane]];
          const MeshDev& m2 = L.meshes[W.mesh2[lane]];
          M3 R;
          V3 T;
          slot_pose(W, lane, R, T);
          V3 P, Q;
          (void)tri_pair(m1, m2, R, T, W.best1[lane], W.best2[lane], P, Q);
          L.distance[q] = W.best_d[lane];
          if (L.ids) {
            L.ids[2 * q] = W.best1[lane];
            L.ids[2 * q + 1] = W.best2[lane];
          }
          if (L.points) {
            M3 R1;
            V3 t1;
            load_pose(L.tf1 + 12 * static_cast<size_t>(q), R1, t1);
            const V3 wp = mul(R1, P) + t1, wq = mul(R1, Q) + t1;
            double* o = L.points + 6 * static_cast<size_t>(q);
            o[0] = wp.x; o[1] = wp.y; o[2] = wp.z;
            o[3] = wq.x; o[4] = wq.y; o[5] = wq.z;
          }
          if (L.bv_tests) L.bv_tests[q] = W.nbv[lane];
          if (L.leaf_tests) L.leaf_tests[q] = W.nleaf[lane];
          q = -1;
          spv = 0;
        }
      }
      const int tasks = __builtin_amdgcn_readlane(dscan<kSlots>(spv), kSlots - 1);
      unsigned long long freem = __ballot(slot_lane && q < 0);
      if (!exhausted && freem != 0 && (__popcll(freem) >= 2 || tasks < kDWave)) {
        while (freem != 0) {
          if (nx == ne) {
            unsigned long long b = 0;
            if (lane == 0) b = atomicAdd(next_q, static_cast<unsigned long long>(kSlots));
            b = __shfl(b, 0);
            if (b >= static_cast<unsigned long long>(nq)) {
              exhausted = true;
              break;
            }
            nx = static_cast<int>(b);
            ne = nq - nx < kSlots ? nq : nx + kSlots;
          }
          const int r = __popcll(freem & dbelow());
          const int n = __popcll(freem) < ne - nx ? __popcll(freem) : ne - nx;
          if (slot_lane && q < 0 && r < n) {
            q = nx + r;
            const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
            if (!query_ok(L.meshes, L.num_meshes, i1, i2, /*need_obbrss=*/false)) {
              L.distance[q] = 1.7976931348623157e308;
              if (L.ids) {
                L.ids[2 * q] = -1;
                L.ids[2 * q + 1] = -1;
              }
              if (L.bv_tests) L.bv_tests[q] = 0;
              if (L.leaf_tests) L.leaf_tests[q] = 0;
              atomicOr(L.status, kStatusBadQuery);
              q = -1;
            } else {
              const MeshDev& m1 = L.meshes[i1];
              const MeshDev& m2 = L.meshes[i2];
              M3 R;
              V3 T;
              inverse_compose(L.tf1 + 12 * static_cast<size_t>(q),
                              L.tf2 + 12 * static_cast<size_t>(q), R, T);
#pragma unroll
              for (int k = 0; k < 9; ++k) W.pose[k][lane] = R.m[k];
              W.pose[9][lane] = T.x;
              W.pose[10][lane] = T.y;
              W.pose[11][lane] = T.z;
#pragma unroll
              for (int k = 0; k < 9; ++k) W.pose32[k][lane] = static_cast<float>(R.m[k]);
              W.pose32[9][lane] = static_cast<float>(T.x);
              W.pose32[10][lane] = static_cast<float>(T.y);
              W.pose32[11][lane] = static_cast<float>(T.z);
              W.pose32[12][lane] = rss_margin32(m1.rcmax, m2.rcmax, T);
              W.rss1[lane] = m1.rss64;
              W.rss2[lane] = m2.rss64;
              W.rss1f[lane] = m1.rss32;
              W.rss2f[lane] = m2.rss32;
              W.size1[lane] = m1.size64;
              W.size2[lane] = m2.size64;
              W.mesh1[lane] = i1;
              W.mesh2[lane] = i2;
              // seed: triangle 0 with triangle 0 (distancePreprocessOrientedNode)
              V3 P, Q;
              W.best_d[lane] = tri_pair(m1, m2, R, T, 0, 0, P, Q);
              W.best_key[lane] = 0;
              W.best1[lane] = 0;
              W.best2[lane] = 0;
              W.nbv[lane] = 0;
              W.nleaf[lane] = 0;
              mine[0] = DTask{dpack(0, 0, 0), -1.0, 0};  // the root, never pruned
              W.sp[lane] = 1;
              W.pend[lane] = 0;
              W.upend[lane] = 0;
              W.gen[lane] = (W.gen[lane] + 1) & kDGenMask;
              spv = 1;
            }
          }
          nx += n;
          freem = __ballot(slot_lane && q < 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      if (exhausted && __ballot(slot_lane && q >= 0) == 0) break;
    }

    // tasks per slot: an even share, then the rest to deeper stacks, up to
    // `width` (all 64 once every query is taken), one at a time near the cap.
    // A slot with tasks waiting for FP64 bounds takes none: their children
    // go on its stack first, so the slot keeps descending depth first and
    // its stack keeps the bound below.
    const int lim = wide_cap - spv > 1 ? wide_cap - spv : 1;
    const int wmax = exhausted ? kDWave : width;
    const int cap_w = wmax < lim ? wmax : lim;
    const int avail = (slot_lane && W.upend[lane] > 0) ? 0 : spv;
    int take = avail < kDWave / kSlots ? avail : kDWave / kSlots;
    take = take < cap_w ? take : cap_w;
    int incl = dscan<kSlots>(take);
    int total = __builtin_amdgcn_readlane(incl, kSlots - 1);
    if (total < kDWave) {
      const int e = (avail < cap_w ? avail : cap_w) - take;
      const int ei = dscan<kSlots>(e);
      const int room = kDWave - total - (ei - e);
      take += room <= 0 ? 0 : (room < e ? room : e);
      incl = dscan<kSlots>(take);
      total = __builtin_amdgcn_readlane(incl, kSlots - 1);
    }
    const int excl = incl - take;

    if (nl >= leaf_batch || (total == 0 && nl > 0)) {
      // FP64 triangle distances of queued leaf pairs
      const int n = nl < kDWave ? nl : kDWave;
      nl -= n;
      int slot = -1;
      if (lane < n) {
        const unsigned long long e = W.lq[nl + lane];
        const unsigned long long key = W.lk[nl + lane];
        const int tag = static_cast<int>(e >> 42);
        const int id1 = static_cast<int>((e >> 21) & 0x1FFFFF), id2 = static_cast<int>(e & 0x1FFFFF);
        const int s = tag & ((1 << kDSlotBits) - 1);
        if ((tag >> kDSlotBits) == W.gen[s]) {
          const MeshDev& m1 = L.meshes[W.mesh1[s]];
          const MeshDev& m2 = L.meshes[W.mesh2[s]];
          M3 R;
          V3 T;
          slot_pose(W, s, R, T);
          V3 P, Q;
          const double d = tri_pair(m1, m2, R, T, id1, id2, P, Q);
          W.cd[lane] = d;
          W.ck[lane] = key;
          W.ci1[lane] = id1;
          W.ci2[lane] = id2;
          slot = s;
          atomicSub(&W.pend[s], 1);
          atomicAdd(reinterpret_cast<unsigned long long*>(&W.nleaf[s]), 1ULL);
        }
      }
      W.cslot[lane] = static_cast<signed char>(slot);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (slot_lane) {
        // the slot's smallest (distance, key), found before or now
        double bd = W.best_d[lane];
        unsigned long long bk = W.best_key[lane];
        int b1 = W.best1[lane], b2 = W.best2[lane];
        for (int j = 0; j < n; ++j) {
          if (W.cslot[j] != lane) continue;
          const double d = W.cd[j];
          const unsigned long long k = W.ck[j];
          if (d < bd || (d == bd && k < bk)) {
            bd = d;
            bk = k;
            b1 = W.ci1[j];
            b2 = W.ci2[j];
          }
        }
        W.best_d[lane] = bd;
        W.best_key[lane] = bk;
        W.best1[lane] = b1;
        W.best2[lane] = b2;
      }
      check = true;
    } else if (kF32 && (nu >= kDWave / 2 || (total == 0 && nu > 0))) {
      // FP64 bounds of both children of up to 32 queued tasks, one bound per
      // lane (lanes 2j and 2j+1: task j's two children); the reference's
      // rectDistance decides their order (the key bit) and their pruning
      const int n = nu < kDWave / 2 ? nu : kDWave / 2;
      nu -= n;
      const int j = lane >> 1;
      const bool second = (lane & 1) != 0;
      int slot = 0, pos = 0;
      bool live = false;
      double dm = 0.0;
      int a1 = 0, a2 = 0, e1 = 0, e2 = 0, depth = 0;
      unsigned long long key = 0;
      if (j < n) {
        const unsigned long long e = W.uq[nu + j];
        key = W.uk[nu + j];
        depth = W.ud[nu + j];
        pos = W.upos[nu + j];
        const int tag = static_cast<int>(e >> 42);
        const int b1 = static_cast<int>((e >> 21) & 0x1FFFFF), b2 = static_cast<int>(e & 0x1FFFFF);
        slot = tag & ((1 << kDSlotBits) - 1);
        live = (tag >> kDSlotBits) == W.gen[slot];
        if (live) {
          const RssNode64* r1 = W.rss1[slot];
          const RssNode64* r2 = W.rss2[slot];
          const int fc1 = r1[b1].first_child, fc2 = r2[b2].first_child;
          if ((fc2 < 0) || (fc1 >= 0 && (W.size1[slot][b1] > W.size2[slot][b2]))) {
            a1 = fc1; a2 = b2; e1 = fc1 + 1; e2 = b2;
          } else {
            a1 = b1; a2 = fc2; e1 = b1; e2 = fc2 + 1;
          }
          M3 R;
          V3 T;
          slot_pose(W, slot, R, T);
          dm = second ? rss_distance64(R, T, r1[e1], r2[e2]) : rss_distance64(R, T, r1[a1], r2[a2]);
        }
      }
      const double dother = __shfl_xor(dm, 1);
      if (live && !second) {
        const double d1 = dm, d2 = dother;
        const double bd = W.best_d[slot];
        const unsigned long long bk = W.best_key[slot];
        const unsigned long long bit = 1ULL << (63 - depth);
        // the reference visits the nearer child first (a on ties)
        DTask near, far;
        if (d2 < d1) {
          near = DTask{dpack(e1, e2, depth + 1), d2, key};
          far = DTask{dpack(a1, a2, depth + 1), d1, key | bit};
        } else {
          near = DTask{dpack(a1, a2, depth + 1), d1, key};
          far = DTask{dpack(e1, e2, depth + 1), d2, key | bit};
        }
        const bool keep_far = !(far.d > bd || (far.d == bd && far.key > bk));
        const bool keep_near = !(near.d > bd || (near.d == bd && near.key > bk));
        // the reserved entries: far below near, so near comes off first; a
        // pruned child leaves an entry that is pruned when it comes off
        const DTask dead{0, INFINITY, ~0ULL};
        if (pos + 2 <= slot_cap) {
          DTask* st = stacks + static_cast<long long>(slot) * slot_cap;
          st[pos] = keep_far ? far : dead;
          st[pos + 1] = keep_near ? near : dead;
        }
        atomicSub(&W.pend[slot], 1);
        atomicSub(&W.upend[slot], 1);
        atomicAdd(reinterpret_cast<unsigned long long*>(&W.nbv[slot]), 2ULL);
      }
      check = true;
    } else if (total > 0) {
      // bounds of the children of the tasks on top of the stacks (FP64, or
      // FP32 lower bounds with FP64 for the tasks whose children both stay)
      if (slot_lane) {
        W.offs[lane] = excl;
        W.base[lane] = spv - take;
        for (int j = 0; j < take; ++j) W.lane_slot[excl + j] = static_cast<unsigned char>(lane);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      int slot = 0, first = 0, at = 0;
      bool expand = false, leafc = false, unsure = false;
      DTask c0{0, 0, 0}, c1{0, 0, 0};
      int npush = 0, ub1 = 0, ub2 = 0, udepth = 0;
      unsigned long long ukey = 0;
      if (lane < total) {
        slot = W.lane_slot[lane];
        first = W.offs[slot];
        at = W.base[slot];
        const DTask t = stacks[static_cast<long long>(slot) * slot_cap + at + lane - first];
        const double bd = W.best_d[slot];
        const unsigned long long bk = W.best_key[slot];
        const bool pruned = t.d > bd || (t.d == bd && t.key > bk);
        if (!pruned) {
          const int b1 = static_cast<int>((t.t >> 21) & 0x1FFFFF), b2 = static_cast<int>(t.t & 0x1FFFFF);
          const int depth = static_cast<int>(t.t >> 42);
          if constexpr (kF32) {
            const RssNode32* f1 = W.rss1f[slot];
            const RssNode32* f2 = W.rss2f[slot];
            const int fc1 = f1[b1].first_child, fc2 = f2[b2].first_child;
            if (fc1 < 0 && fc2 < 0) {
              leafc = true;
              c0 = DTask{lpack((W.gen[slot] << kDSlotBits) | slot, -(fc1 + 1), -(fc2 + 1)), 0.0,
                         t.key};
            } else {
              int a1, a2, e1, e2;
              if ((fc2 < 0) || (fc1 >= 0 && (W.size1[slot][b1] > W.size2[slot][b2]))) {
                a1 = fc1; a2 = b2; e1 = fc1 + 1; e2 = b2;
              } else {
                a1 = b1; a2 = fc2; e1 = b1; e2 = fc2 + 1;
              }
              float R0[9], T0[3];
#pragma unroll
              for (int k = 0; k < 9; ++k) R0[k] = W.pose32[k][slot];
              T0[0] = W.pose32[9][slot];
              T0[1] = W.pose32[10][slot];
              T0[2] = W.pose32[11][slot];
              const float ms = W.pose32[12][slot];
              const float l1 = rss_lb32(R0, T0, ms, f1[a1], f2[a2]);
              const float l2 = rss_lb32(R0, T0, ms, f1[e1], f2[e2]);
              expand = true;
              // a child is pruned when its lower bound is above the best
              // distance (its FP64 bound is then too)
              const bool k1 = !(static_cast<double>(l1) > bd);
              const bool k2 = !(static_cast<double>(l2) > bd);
              if (k1 && k2) {
                // their order needs the FP64 bounds: two stack entries are
                // reserved for them in lane order, so the stack stays as the
                // pushes of the FP64 bounds would have left it
                unsure = true;
                npush = 2;
                ub1 = b1;
                ub2 = b2;
                udepth = depth;
                ukey = t.key;
              } else if (k1) {
                c0 = DTask{dpack(a1, a2, depth + 1), static_cast<double>(l1), t.key};
                npush = 1;
              } else if (k2) {
                c0 = DTask{dpack(e1, e2, depth + 1), static_cast<double>(l2), t.key};
                npush = 1;
              }
            }
          } else {
            const RssNode64* r1 = W.rss1[slot];
            const RssNode64* r2 = W.rss2[slot];
            const int fc1 = r1[b1].first_child, fc2 = r2[b2].first_child;
            if (fc1 < 0 && fc2 < 0) {
              leafc = true;
              c0 = DTask{lpack((W.gen[slot] << kDSlotBits) | slot, -(fc1 + 1), -(fc2 + 1)), 0.0,
                         t.key};
            } else {
              int a1, a2, e1, e2;
              if ((fc2 < 0) || (fc1 >= 0 && (W.size1[slot][b1] > W.size2[slot][b2]))) {
                a1 = fc1; a2 = b2; e1 = fc1 + 1; e2 = b2;
              } else {
                a1 = b1; a2 = fc2; e1 = b1; e2 = fc2 + 1;
              }
              M3 R;
              V3 T;
              slot_pose(W, slot, R, T);
              const double d1 = rss_distance64(R, T, r1[a1], r2[a2]);
              const double d2 = rss_distance64(R, T, r1[e1], r2[e2]);
              expand = true;
              const unsigned long long bit = 1ULL << (63 - depth);
              // the reference visits the nearer child first (a on ties)
              DTask near, far;
              if (d2 < d1) {
                near = DTask{dpack(e1, e2, depth + 1), d2, t.key};
                far = DTask{dpack(a1, a2, depth + 1), d1, t.key | bit};
              } else {
                near = DTask{dpack(a1, a2, depth + 1), d1, t.key};
                far = DTask{dpack(e1, e2, depth + 1), d2, t.key | bit};
              }
              const bool keep_far = !(far.d > bd || (far.d == bd && far.key > bk));
              const bool keep_near = !(near.d > bd || (near.d == bd && near.key > bk));
              // pushed far below near, so near comes off first
              if (keep_far) c0 = far;
              if (keep_near) (keep_far ? c1 : c0) = near;
              npush = (keep_far ? 1 : 0) + (keep_near ? 1 : 0);
            }
          }
        }
      }
      // children to the slots' stacks, in lane order
      const unsigned long long m1b = __ballot(npush >= 1), m2b = __ballot(npush == 2);
      int o_stack = 0;
      if (npush > 0) {
        const unsigned long long same = ~((1ULL << first) - 1ULL) & dbelow();
        o_stack = at + __popcll(m1b & same) + __popcll(m2b & same);
        DTask* s = stacks + static_cast<long long>(slot) * slot_cap;
        if (!unsure && o_stack + npush <= slot_cap) {
          s[o_stack] = c0;
          if (npush == 2) s[o_stack + 1] = c1;
        }
      }
      const unsigned long long lm = __ballot(leafc);
      if (leafc) {
        const int o = nl + __popcll(lm & dbelow());
        W.lq[o] = c0.t;
        W.lk[o] = c0.key;
      }
      nl += __popcll(lm);
      const unsigned long long um = __ballot(unsure);
      if (unsure) {
        const int o = nu + __popcll(um & dbelow());
        W.uq[o] = lpack((W.gen[slot] << kDSlotBits) | slot, ub1, ub2);
        W.uk[o] = ukey;
        W.upos[o] = o_stack;
        W.ud[o] = static_cast<unsigned char>(udepth);
      }
      nu += __popcll(um);
      const unsigned long long em = __ballot(expand);
      bool emptied = false;
      if (slot_lane && take > 0) {
        const unsigned long long own =
            (take >= kDWave ? ~0ULL : ((1ULL << take) - 1ULL)) << excl;
        const int nsp = spv - take + __popcll(m1b & own) + __popcll(m2b & own);
        if (nsp > slot_cap) {
          atomicOr(L.status, kStatusStackOverflow);
          spv = 0;  // the query's answer is incomplete; the status says so
        } else {
          spv = nsp;
        }
        W.sp[lane] = spv;
        W.pend[lane] += __popcll((lm | um) & own);
        W.upend[lane] += __popcll(um & own);
        W.nbv[lane] += 2 * __popcll(em & own);
        emptied = spv == 0;
      }
      if (__ballot(emptied) != 0) check = true;
    } else {
      check = true;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

__global__ void dist_zero(unsigned long long* p) { *p = 0; }

}  // namespace

hipError_t launch_distance_wave(const DistanceLaunch& L, int stack_need, const DistanceTuning& tu,
                                hipStream_t s) {
  if (L.num_queries == 0) return hipSuccess;
  // path keys hold 63 levels of descent
  if (stack_need > 62) return launch_distance_plain(L, stack_need, s);
  int dev = 0, cus = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess ||
      (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) !=
          hipSuccess)
    return e;
  const int K = tu.queries_per_wave;
  const int wps = tu.waves_per_simd;
  const int width = std::max(tu.width, kDWave / K);
  const int leaf_batch = tu.leaf_batch;
  const long long nw = std::max<long long>(
      1, std::min<long long>((L.num_queries + K - 1) / K, static_cast<long long>(cus) * 4 * wps));
  const unsigned grid = static_cast<unsigned>((nw + kDWaves - 1) / kDWaves);
  // as the BOOL engine: sorted by level, at most 2 entries per level above
  // the point where steps go one task at a time
  const int slot_cap = std::max(kDSlotCap, 4 * (stack_need + 2) + 2 * kDQueue);
  const int wide_cap = slot_cap - (stack_need + 2) - 2 * kDQueue;
  DTask* stk = nullptr;
  unsigned long long* ctr = nullptr;
  if ((e = hipMallocAsync(reinterpret_cast<void**>(&stk),
                          sizeof(DTask) * static_cast<size_t>(grid) * kDWaves * K * slot_cap,
                          s)) != hipSuccess)
    return e;
  if ((e = hipMallocAsync(reinterpret_cast<void**>(&ctr), sizeof(unsigned long long), s)) !=
      hipSuccess) {
    (void)hipFreeAsync(stk, s);
    return e;
  }
  dist_zero<<<1, 1, 0, s>>>(ctr);
#define FCLX_DIST(k, w, f)                                                               \
  engine_distance_wave<k, w, f><<<grid, kDBlock, 0, s>>>(L, ctr, stk, slot_cap, wide_cap, width, \
                                                         leaf_batch)
  if (tu.fp32_cull) {
    if (K == 2)
      wps >= 3 ? FCLX_DIST(2, 3, true) : FCLX_DIST(2, 2, true);
    else if (K == 8)
      wps >= 3 ? FCLX_DIST(8, 3, true) : FCLX_DIST(8, 2, true);
    else
      wps >= 3 ? FCLX_DIST(4, 3, true) : FCLX_DIST(4, 2, true);
  } else {
    if (K == 2)
      wps >= 3 ? FCLX_DIST(2, 3, false) : FCLX_DIST(2, 2, false);
    else if (K == 8)
      wps >= 3 ? FCLX_DIST(8, 3, false) : FCLX_DIST(8, 2, false);
    else
      wps >= 3 ? FCLX_DIST(4, 3, false) : FCLX_DIST(4, 2, false);
  }
#undef FCLX_DIST
  e = hipGetLastError();
  (void)hipFreeAsync(ctr, s);
  (void)hipFreeAsync(stk, s);
  return e;
}

}  // namespace fclx
