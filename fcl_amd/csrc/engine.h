// Wavefront-cooperative traversal engine (FCLX_ENGINE_FAST).
#pragma once
#include "kernels.h"

namespace fclx {

struct EngineWorkspace;

EngineWorkspace* engine_workspace_create();
void engine_workspace_destroy(EngineWorkspace* ws);

// Collide batch with the fast engine.  num_max_contacts <= kOrderedCapMax
// runs the per-lane ordered traversal with FP32 culling; larger caps run the
// breadth-first front expansion + persistent wavefront DFS, and re-run any
// query whose contact count reaches the cap with the plain kernel so that
// its truncated list is the reference's depth-first prefix.
// Booleans only: counts[q] = 1 when the query's pair collides, else 0; no
// records or statistics.  Batches on different streams may overlap.
hipError_t engine_anyhit_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                                 hipStream_t s);
hipError_t engine_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                          hipStream_t s);

// Tunables of the exhaustive path, per context (defaults chosen by measurement).
struct EngineTuning {
  DistanceTuning dist;          // FAST distance engine
  int meshshape_waves_per_simd = 2;  // register budget of the mesh-vs-primitive kernel
  int meshshape_leaf_share = 0;      // percent of live lanes at leaves that starts a leaf round
                                     // (0: 50 for booleans, 75 with contact geometry)
  int max_levels = 24;          // breadth-first levels before a depth-first round
  long long front_target = 1 << 18;  // stop expanding once the front is this large (a full
                                     // step for every resident wavefront)
  int dfs_blocks_per_cu = 8;    // persistent grid cap: blocks of 4 wavefronts per CU
  long long defer_leaf = 1LL << 27;       // deferred leaf-pair buffer (entries)
  long long defer_uncertain = 1LL << 24;  // deferred uncertain-task buffer (entries)
  double idle_stop = 0.5;       // end a round once this share of wavefronts is idle
  int stop_poll = 16;           // depth-first steps between polls of the round's stop flag
  int dfs_waves_per_simd = 4;   // register budget of the depth-first kernel (4..6)
  int dfs_inline_leaves = 1;    // 1: the depth-first kernel runs its triangle tests itself
                                // (at 4 wavefronts per SIMD only)
  int dfs_stack_top = 512;      // with inline triangle tests: stack entries per wavefront
                                // kept in LDS (128, 256 or 512)
  int bool_queries_per_wave = 8;  // queries a boolean-engine wavefront keeps in flight (4, 8, 16,
                                  // 32; 0: 4 below 65536 queries, else 8)
  int bool_waves_per_simd = 4;    // register budget of the boolean kernel (2, 3, 4; 32 slots: 2, 3)
  int bool_leaf_batch = 64;       // leaf pairs queued before the boolean kernel tests them
  int bool_width = 16;            // most tasks of one query a boolean-kernel step tests
  int bool_refill = 4;            // free slots that make a boolean wavefront take new queries
  int fast_persistent = 0;        // 1: exhaustive FAST batches as one persistent launch (the
                                  // boolean kernel's exhaustive form) instead of rounds
  int fast_graph = 1;             // 1: that launch sequence replayed from a captured HIP graph
                                  // while its arguments stay the same
};
EngineTuning& engine_tuning(EngineWorkspace* ws);  // per context

// The boolean kernel over a work area the caller holds (area_words 8-byte
// words, 8-byte aligned): no allocation, no events; the grid is cut to the
// workgroups whose stacks fit.  Calls that overlap on one area race.
// hipErrorInvalidValue when not even one workgroup fits
// (engine_anyhit_min_words gives that size).
hipError_t engine_anyhit_collide_area(const EngineTuning& tu, const CollideLaunch& L,
                                      int stack_need, unsigned long long* area,
                                      size_t area_words, hipStream_t s);
size_t engine_anyhit_min_words(const EngineTuning& tu, long long num_queries, int stack_need);

constexpr int kOrderedCapMax = 64;

struct EngineReport {
  double setup_ms, expand_ms, leaves_ms, dfs_ms, fixup_ms;
  long long expand_bv_tests, expand_leaf_tests, dfs_bv_tests, dfs_leaf_tests, fp64_fallbacks;
  long long front_size;
  int levels, flagged;
  int dfs_waves, dfs_rounds;
  double dfs_wave_busy, dfs_end_p50_ms, dfs_end_p90_ms, dfs_end_max_ms;
  // BOOL engine (profiling): wavefront steps and the lanes they used, per
  // kind (FP32 BV, FP64 triangle, FP64 BV); tasks dropped because their
  // query was answered; query rounds
  long long bool_steps[3], bool_lanes[3], bool_skipped, bool_rounds, bool_stack_max;
  double bool_wave_ms;  // mean wavefront lifetime
  double bool_end_p50_ms, bool_end_p90_ms, bool_span_ms;  // wavefront ends after the first start
  long long dfs_inline_leaf_tests;  // of dfs_leaf_tests: run inside the depth-first kernel
  long long dfs_wave_steps;         // BV-test steps of the depth-first wavefronts
};
hipError_t engine_set_profiling(EngineWorkspace* ws, bool on);
const EngineReport& engine_report(const EngineWorkspace* ws);

}  // namespace fclx
