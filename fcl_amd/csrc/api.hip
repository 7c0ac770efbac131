// C ABI of the engine: contexts, mesh upload (host BVH build -> HBM), and
// the batched / single-query collide and distance entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fclx.h"
#include "engine.h"
#include "host/bvh_build.h"
#include "kernels.h"

using namespace fclx;

namespace {

struct MeshSlot {
  bool live = false;
  HostBVH bvh;
  ObbNode64* obb64 = nullptr;
  RssNode64* rss64 = nullptr;
  double* size64 = nullptr;
  ObbNode32* obb32 = nullptr;        // obb32_alloc + 1 (see fclx_mesh_create)
  ObbNode32* obb32_alloc = nullptr;
  RssNode32* rss32 = nullptr;
  double* verts = nullptr;
  int32_t* tris = nullptr;
  double* tri_verts = nullptr;
  int32_t* span = nullptr;
  int32_t* prim_pos = nullptr;
  int64_t bytes = 0;
  MeshDev dev{};
  // device refit: scratch, the nodes level by level (bottom-up refit) and
  // whether the device BVs differ from the host build (export reads them);
  // refit_cmax holds the two cull-error scales (OBB centres, RSS origins)
  int32_t* refit_order = nullptr;
  double* refit_cmax = nullptr;
  int32_t* level_nodes = nullptr;
  std::vector<int32_t> level_offsets;
  bool refitted = false;
  // the BVs were last merged from their children (bottom-up refit): the
  // reference's OBB and RSS merges need not contain every child, so the
  // bounds are no longer lower bounds of every pair below them
  bool merged = false;
};

template <typename T>
hipError_t dmalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(n, 1));
}

__global__ void clear_fronts(int32_t* sizes, int32_t* meshes, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sizes[i] = 0;
  meshes[2 * i] = -1;
  meshes[2 * i + 1] = -1;
}

__global__ void zero_i64(long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

struct fclx_context {
  int device = 0;
  std::mutex mu;          // meshes, table, single-query staging
  std::string err;
  std::vector<MeshSlot> meshes;
  MeshDev* d_table = nullptr;
  int table_cap = 0;
  int32_t* d_status = nullptr;
  hipStream_t stream = nullptr;  // single-query calls
  // single-query staging
  double* s_tf = nullptr;        // 24
  int32_t* s_qm = nullptr;       // 2
  int32_t* s_count = nullptr;    // 1
  long long* s_num = nullptr;    // 1
  long long* s_stats = nullptr;  // 2
  int32_t* s_pairs = nullptr;
  double* s_geo = nullptr;       // contact geometry: 7 doubles per staged record
  long long s_pairs_cap = 0;
  double* s_dist = nullptr;      // 1 + 6 points
  int32_t* s_ids = nullptr;      // 2
  EngineWorkspace* ws = nullptr;
  std::mutex ws_mu;       // the FAST engine's work buffers (one batch at a time)
  ShapeWorkspace* sws = nullptr;
  std::mutex sws_mu;      // the shape path's work buffers
  char* s_shape = nullptr;        // single shape-pair staging (under mu)
  size_t s_shape_bytes = 0;
  int epa_blocks_per_cu = 0;
  int shape_waves = 2;
  std::atomic<int> serve_fail_launches{0};  // failure injection for the serving loop
};

namespace {

int fail(fclx_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(fclx_context* ctx, hipError_t e, const char* what) {
  return fail(ctx, e == hipErrorOutOfMemory ? FCLX_ERR_OUT_OF_MEMORY : FCLX_ERR_DEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                              \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

void free_slot(MeshSlot& s) {
  (void)hipFree(s.obb64);
  (void)hipFree(s.rss64);
  (void)hipFree(s.size64);
  (void)hipFree(s.obb32_alloc);
  (void)hipFree(s.rss32);
  (void)hipFree(s.verts);
  (void)hipFree(s.tri_verts);
  (void)hipFree(s.tris);
  (void)hipFree(s.span);
  (void)hipFree(s.prim_pos);
  (void)hipFree(s.refit_order);
  (void)hipFree(s.refit_cmax);
  (void)hipFree(s.level_nodes);
  s = MeshSlot{};
}

int table_sync(fclx_context* ctx) {
  const int n = static_cast<int>(ctx->meshes.size());
  if (n > ctx->table_cap) {
    int cap = std::max(16, ctx->table_cap * 2);
    while (cap < n) cap *= 2;
    MeshDev* t = nullptr;
    HIP_TRY(ctx, dmalloc(&t, cap));
    (void)hipFree(ctx->d_table);
    ctx->d_table = t;
    ctx->table_cap = cap;
  }
  std::vector<MeshDev> host(n);
  for (int i = 0; i < n; ++i) host[i] = ctx->meshes[i].live ? ctx->meshes[i].dev : MeshDev{};
  HIP_TRY(ctx, hipMemcpy(ctx->d_table, host.data(), sizeof(MeshDev) * n, hipMemcpyHostToDevice));
  return FCLX_OK;
}

bool valid_mesh(const fclx_context* ctx, int32_t id) {
  return id >= 0 && id < static_cast<int32_t>(ctx->meshes.size()) && ctx->meshes[id].live;
}

int max_depth(const fclx_context* ctx) {
  int d = 0;
  for (const auto& m : ctx->meshes)
    if (m.live) d = std::max(d, m.bvh.depth);
  return d;
}

// Some live mesh holds BVs merged by a bottom-up refit.
bool any_merged(const fclx_context* ctx) {
  for (const auto& m : ctx->meshes)
    if (m.live && m.merged) return true;
  return false;
}

size_t max_nodes(const fclx_context* ctx) {
  size_t n = 0;
  for (const auto& m : ctx->meshes)
    if (m.live) n = std::max(n, m.bvh.nodes.size());
  return n;
}

// The FAST engine packs (query, node, node) into 64 bits: 21 bits per node.
constexpr size_t kFastMaxNodes = size_t(1) << 21;
// ... and 22 bits per query.
constexpr int32_t kFastMaxQueries = int32_t(1) << 22;

// Work area of fclx_collide_batch_noalloc: the boolean kernel's range
// counters and per-query stacks.  Each GPU the library runs on holds its
// own copy for the life of the process.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): fclx_init resolves the device address of each GPU's copy with
// hipGetSymbolAddress, with that GPU current, and every use goes through
// that address.
constexpr size_t kScratchWords = FCLX_NOALLOC_SCRATCH_BYTES / sizeof(unsigned long long);
__device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Resolves the current device's copy of `scratch` once (device `device` is current).
void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = static_cast<unsigned long long*>(p);
}

// Largest L1 norm of an OBB centre over the mesh's nodes; the FP32 cull's
// translation error term scales with it.
double centre_bound(const std::vector<HostNode>& nodes) {
  double mx = 0;
  for (const auto& n : nodes)
    mx = std::max(mx, std::fabs(n.obb_to[0]) + std::fabs(n.obb_to[1]) + std::fabs(n.obb_to[2]));
  return mx;
}

// Largest L1 norm of an RSS origin over the mesh's nodes; the FP32 distance
// bound's translation error term scales with it.
double origin_bound(const std::vector<HostNode>& nodes) {
  double mx = 0;
  for (const auto& n : nodes)
    mx = std::max(mx, std::fabs(n.rss_to[0]) + std::fabs(n.rss_to[1]) + std::fabs(n.rss_to[2]));
  return mx;
}

}  // namespace

namespace fclx {
int context_device(const fclx_context* ctx) { return ctx ? ctx->device : 0; }
}  // namespace fclx

extern "C" {

int fclx_init(int32_t device, fclx_context** out) {
  if (!out) return FCLX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FCLX_ERR_DEVICE;
  if (device < 0 || device >= ndev) return FCLX_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return FCLX_ERR_DEVICE;
  auto* ctx = new (std::nothrow) fclx_context();
  if (!ctx) return FCLX_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  bool ok = dmalloc(&ctx->d_status, 1) == hipSuccess &&
            hipMemset(ctx->d_status, 0, sizeof(int32_t)) == hipSuccess &&
            hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
            dmalloc(&ctx->s_tf, 24) == hipSuccess && dmalloc(&ctx->s_qm, 2) == hipSuccess &&
            dmalloc(&ctx->s_count, 1) == hipSuccess && dmalloc(&ctx->s_num, 1) == hipSuccess &&
            dmalloc(&ctx->s_stats, 2) == hipSuccess && dmalloc(&ctx->s_dist, 7) == hipSuccess &&
            dmalloc(&ctx->s_ids, 2) == hipSuccess;
  if (!ok) {
    fclx_shutdown(ctx);
    return FCLX_ERR_DEVICE;
  }
  ctx->ws = engine_workspace_create();
  resolve_scratch(device);
  *out = ctx;
  return FCLX_OK;
}

void fclx_shutdown(fclx_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto& m : ctx->meshes)
    if (m.live) free_slot(m);
  engine_workspace_destroy(ctx->ws);
  shape_workspace_destroy(ctx->sws);
  (void)hipFree(ctx->d_table);
  (void)hipFree(ctx->d_status);
  (void)hipFree(ctx->s_tf);
  (void)hipFree(ctx->s_qm);
  (void)hipFree(ctx->s_count);
  (void)hipFree(ctx->s_num);
  (void)hipFree(ctx->s_stats);
  (void)hipFree(ctx->s_pairs);
  (void)hipFree(ctx->s_geo);
  (void)hipFree(ctx->s_dist);
  (void)hipFree(ctx->s_ids);
  (void)hipFree(ctx->s_shape);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* fclx_last_error(const fclx_context* ctx) { return ctx ? ctx->err.c_str() : ""; }

int fclx_mesh_create(fclx_context* ctx, int32_t bv_type, const double* vertices,
                     int32_t num_vertices, const int32_t* tris, int32_t num_tris,
                     int32_t* mesh_id) {
  if (!ctx || !mesh_id) return FCLX_ERR_INVALID_ARGUMENT;
  if (bv_type != FCLX_BV_OBBRSS && bv_type != FCLX_BV_RSS)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "only OBBRSS and RSS hierarchies are supported");
  if (num_tris <= 0 || num_vertices <= 0 || !vertices || !tris)
    return fail(ctx, FCLX_ERR_EMPTY_MODEL, "model has no triangles or vertices");
  for (int64_t i = 0; i < 3LL * num_tris; ++i)
    if (tris[i] < 0 || tris[i] >= num_vertices)
      return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "triangle index out of range");
  for (int64_t i = 0; i < 3LL * num_vertices; ++i)
    if (!std::isfinite(vertices[i]))
      return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "non-finite vertex coordinate");
  MeshSlot s;
  if (build_bvh(s.bvh, bv_type, vertices, num_vertices, tris, num_tris) != 0)
    return fail(ctx, FCLX_ERR_EMPTY_MODEL, "empty model");
  const auto& nodes = s.bvh.nodes;
  const int nb = static_cast<int>(nodes.size());
  std::vector<ObbNode64> o64(nb);
  std::vector<RssNode64> r64(nb);
  std::vector<ObbNode32> o32(nb);
  std::vector<RssNode32> r32(nb);
  std::vector<double> sz(nb);
  for (int i = 0; i < nb; ++i) {
    const HostNode& h = nodes[i];
    ObbNode64& o = o64[i];
    std::memcpy(o.axis, h.obb_axis, sizeof(o.axis));
    std::memcpy(o.to, h.obb_to, sizeof(o.to));
    std::memcpy(o.extent, h.obb_extent, sizeof(o.extent));
    o.first_child = h.first_child;
    o.pad = 0;
    RssNode64& r = r64[i];
    std::memcpy(r.axis, h.rss_axis, sizeof(r.axis));
    std::memcpy(r.to, h.rss_to, sizeof(r.to));
    r.l[0] = h.rss_l[0];
    r.l[1] = h.rss_l[1];
    r.r = h.rss_r;
    r.first_child = h.first_child;
    r.pad = 0;
    ObbNode32& f = o32[i];
    for (int k = 0; k < 9; ++k) f.axis[k] = static_cast<float>(h.obb_axis[k]);
    for (int k = 0; k < 3; ++k) {
      f.to[k] = static_cast<float>(h.obb_to[k]);
      f.extent[k] = static_cast<float>(h.obb_extent[k]);
    }
    f.first_child = h.first_child;
    RssNode32& g = r32[i];
    for (int k = 0; k < 9; ++k) g.axis[k] = static_cast<float>(h.rss_axis[k]);
    for (int k = 0; k < 3; ++k) g.to[k] = static_cast<float>(h.rss_to[k]);
    g.l[0] = static_cast<float>(h.rss_l[0]);
    g.l[1] = static_cast<float>(h.rss_l[1]);
    g.r = static_cast<float>(h.rss_r);
    g.first_child = h.first_child;
    sz[i] = s.bvh.node_size(i);
  }
  // primitive ranges of the nodes and each triangle's position in the build
  // order: they locate a triangle's leaf from the root (depth-first keys)
  std::vector<int32_t> span(2 * static_cast<size_t>(nb));
  std::vector<int32_t> pos(num_tris);
  for (int i = 0; i < nb; ++i) {
    span[2 * i] = nodes[i].first_primitive;
    span[2 * i + 1] = nodes[i].num_primitives;
  }
  for (int i = 0; i < num_tris; ++i) pos[s.bvh.primitive_indices[i]] = i;
  // triangle-major copy of the vertices: one 128-B line per triangle
  std::vector<double> tv(static_cast<size_t>(kTriVertsStride) * num_tris, 0.0);
  for (int i = 0; i < num_tris; ++i)
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c)
        tv[kTriVertsStride * static_cast<size_t>(i) + 3 * k + c] =
            vertices[3 * static_cast<size_t>(tris[3 * i + k]) + c];
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  // The FP32 nodes start one 64-B record into their allocation (256-B
  // aligned): node i sits at 64 (i + 1) bytes, so each pair of children
  // (first_child odd) shares one 128-B line, which the two child tasks of a
  // descent read in the same step.
  bool ok = dmalloc(&s.obb64, nb) == hipSuccess && dmalloc(&s.rss64, nb) == hipSuccess &&
            dmalloc(&s.size64, nb) == hipSuccess && dmalloc(&s.obb32_alloc, nb + 1) == hipSuccess &&
            dmalloc(&s.rss32, nb) == hipSuccess &&
            dmalloc(&s.verts, 3 * static_cast<size_t>(num_vertices)) == hipSuccess &&
            dmalloc(&s.tris, 3 * static_cast<size_t>(num_tris)) == hipSuccess &&
            dmalloc(&s.tri_verts, tv.size()) == hipSuccess &&
            dmalloc(&s.span, 2 * static_cast<size_t>(nb)) == hipSuccess &&
            dmalloc(&s.prim_pos, static_cast<size_t>(num_tris)) == hipSuccess;
  if (ok) s.obb32 = s.obb32_alloc + 1;
  if (ok)
    ok = hipMemcpy(s.span, span.data(), sizeof(int32_t) * span.size(), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.prim_pos, pos.data(), sizeof(int32_t) * pos.size(),
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(s.obb64, o64.data(), sizeof(ObbNode64) * nb, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.rss64, r64.data(), sizeof(RssNode64) * nb, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.size64, sz.data(), sizeof(double) * nb, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.obb32, o32.data(), sizeof(ObbNode32) * nb, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.rss32, r32.data(), sizeof(RssNode32) * nb, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.verts, vertices, sizeof(double) * 3 * num_vertices,
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(s.tris, tris, sizeof(int32_t) * 3 * num_tris, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(s.tri_verts, tv.data(), sizeof(double) * tv.size(), hipMemcpyHostToDevice) ==
             hipSuccess;
  if (!ok) {
    free_slot(s);
    return fail(ctx, FCLX_ERR_OUT_OF_MEMORY, "device allocation or upload failed");
  }
  s.bytes = sizeof(ObbNode32) + static_cast<int64_t>(nb) * (sizeof(ObbNode64) + sizeof(RssNode64) + sizeof(double) +
                                        sizeof(ObbNode32) + sizeof(RssNode32)) +
            24LL * num_vertices + 16LL * num_tris + 8LL * nb +
            8LL * kTriVertsStride * num_tris;
  s.live = true;
  MeshDev& d = s.dev;
  d.span = s.span;
  d.prim_pos = s.prim_pos;
  d.obb64 = s.obb64;
  d.rss64 = s.rss64;
  d.size64 = s.size64;
  d.obb32 = s.obb32;
  d.rss32 = s.rss32;
  d.verts = s.verts;
  d.tris = s.tris;
  d.tri_verts = s.tri_verts;
  d.num_bvs = nb;
  d.num_tris = num_tris;
  d.num_verts = num_vertices;
  d.depth = s.bvh.depth;
  d.bv_type = bv_type;
  d.pad = 0;
  d.cmax = centre_bound(nodes);
  d.rcmax = origin_bound(nodes);
  int id = -1;
  for (size_t i = 0; i < ctx->meshes.size(); ++i)
    if (!ctx->meshes[i].live) {
      id = static_cast<int>(i);
      break;
    }
  if (id < 0) {
    id = static_cast<int>(ctx->meshes.size());
    ctx->meshes.emplace_back();
  }
  ctx->meshes[id] = std::move(s);
  int rc = table_sync(ctx);
  if (rc != FCLX_OK) {
    free_slot(ctx->meshes[id]);
    return rc;
  }
  *mesh_id = id;
  return FCLX_OK;
}

int fclx_mesh_destroy(fclx_context* ctx, int32_t mesh_id) {
  if (!ctx) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh_id)) return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();  // no kernel may still read the mesh
  free_slot(ctx->meshes[mesh_id]);
  return table_sync(ctx);
}

int fclx_mesh_get_info(fclx_context* ctx, int32_t mesh_id, fclx_mesh_info* info) {
  if (!ctx || !info) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh_id)) return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  const MeshSlot& s = ctx->meshes[mesh_id];
  info->num_bvs = static_cast<int32_t>(s.bvh.nodes.size());
  info->num_tris = static_cast<int32_t>(s.bvh.tris.size() / 3);
  info->num_vertices = static_cast<int32_t>(s.bvh.vertices.size() / 3);
  info->depth = s.bvh.depth;
  info->bv_type = s.bvh.bv_type;
  info->reserved = 0;
  info->device_bytes = s.bytes;
  return FCLX_OK;
}

}  // extern "C"

namespace {
void export_nodes(const HostBVH& b, double* bv_data, int32_t* links, double* sizes) {
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const HostNode& n = b.nodes[i];
    if (bv_data) {
      double* o = bv_data + 31 * i;
      std::memcpy(o, n.obb_axis, 9 * sizeof(double));
      std::memcpy(o + 9, n.obb_to, 3 * sizeof(double));
      std::memcpy(o + 12, n.obb_extent, 3 * sizeof(double));
      std::memcpy(o + 15, n.rss_axis, 9 * sizeof(double));
      std::memcpy(o + 24, n.rss_to, 3 * sizeof(double));
      o[27] = n.rss_l[0];
      o[28] = n.rss_l[1];
      o[29] = n.rss_r;
      o[30] = 0;
    }
    if (links) {
      links[3 * i] = n.first_child;
      links[3 * i + 1] = n.first_primitive;
      links[3 * i + 2] = n.num_primitives;
    }
    if (sizes) sizes[i] = b.node_size(static_cast<int>(i));
  }
}
}  // namespace

extern "C" {

int fclx_mesh_export_nodes(fclx_context* ctx, int32_t mesh_id, double* bv_data, int32_t* links,
                           double* sizes) {
  if (!ctx) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh_id)) return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  const MeshSlot& m = ctx->meshes[mesh_id];
  export_nodes(m.bvh, bv_data, links, sizes);
  if (!m.refitted || (!bv_data && !sizes)) return FCLX_OK;
  // refit on the device: the BVs and sizes come from device memory
  (void)hipSetDevice(ctx->device);
  const size_t nb = m.bvh.nodes.size();
  std::vector<ObbNode64> o(nb);
  std::vector<RssNode64> r(nb);
  HIP_TRY(ctx, hipMemcpy(o.data(), m.obb64, nb * sizeof(ObbNode64), hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(r.data(), m.rss64, nb * sizeof(RssNode64), hipMemcpyDeviceToHost));
  if (sizes)
    HIP_TRY(ctx, hipMemcpy(sizes, m.size64, nb * sizeof(double), hipMemcpyDeviceToHost));
  if (bv_data)
    for (size_t i = 0; i < nb; ++i) {
      double* d = bv_data + 31 * i;
      std::memcpy(d, o[i].axis, 9 * sizeof(double));
      std::memcpy(d + 9, o[i].to, 3 * sizeof(double));
      std::memcpy(d + 12, o[i].extent, 3 * sizeof(double));
      std::memcpy(d + 15, r[i].axis, 9 * sizeof(double));
      std::memcpy(d + 24, r[i].to, 3 * sizeof(double));
      d[27] = r[i].l[0];
      d[28] = r[i].l[1];
      d[29] = r[i].r;
      d[30] = 0;
    }
  return FCLX_OK;
}

int fclx_mesh_refit(fclx_context* ctx, int32_t mesh_id, const double* vertices,
                    int32_t bottomup, void* stream) {
  if (!ctx || !vertices) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh_id)) return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  (void)hipSetDevice(ctx->device);
  MeshSlot& m = ctx->meshes[mesh_id];
  const auto& nodes = m.bvh.nodes;
  const int nb = static_cast<int>(nodes.size());
  const int nt = static_cast<int>(m.bvh.tris.size() / 3);
  const int nv = static_cast<int>(m.bvh.vertices.size() / 3);
  if (!m.refit_order) {
    HIP_TRY(ctx, dmalloc(&m.refit_order, static_cast<size_t>(nt)));
    HIP_TRY(ctx, dmalloc(&m.refit_cmax, 2));
  }
  if (bottomup && !m.level_nodes) {
    // the nodes level by level from the root
    std::vector<int32_t> order{0}, offsets{0, 1};
    for (size_t a = 0; a < order.size();) {
      const size_t b = order.size();
      for (size_t i = a; i < b; ++i) {
        const int fc = nodes[order[i]].first_child;
        if (fc >= 0) {
          order.push_back(fc);
          order.push_back(fc + 1);
        }
      }
      if (order.size() > b) offsets.push_back(static_cast<int32_t>(order.size()));
      a = b;
    }
    HIP_TRY(ctx, dmalloc(&m.level_nodes, order.size()));
    HIP_TRY(ctx, hipMemcpy(m.level_nodes, order.data(), order.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));
    m.level_offsets = offsets;
  }
  RefitLaunch L{};
  L.new_verts = vertices;
  L.verts = m.verts;
  L.tri_verts = m.tri_verts;
  L.tris = m.tris;
  L.span = m.span;
  L.prim_pos = m.prim_pos;
  L.order = m.refit_order;
  L.obb64 = m.obb64;
  L.rss64 = m.rss64;
  L.size64 = m.size64;
  L.obb32 = m.obb32;
  L.rss32 = m.rss32;
  L.cmax = m.refit_cmax;
  L.rcmax = m.refit_cmax + 1;
  L.num_verts = nv;
  L.num_tris = nt;
  L.num_bvs = nb;
  L.bv_type = m.bvh.bv_type;
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(ctx, launch_refit(L, bottomup != 0, m.level_nodes, m.level_offsets.data(),
                            static_cast<int>(m.level_offsets.size()) - 1, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  double cmax[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpy(cmax, m.refit_cmax, 2 * sizeof(double), hipMemcpyDeviceToHost));
  m.dev.cmax = cmax[0];
  m.dev.rcmax = cmax[1];
  m.refitted = true;
  m.merged = bottomup != 0;
  return table_sync(ctx);
}

int fclx_bvh_build(int32_t bv_type, const double* vertices, int32_t num_vertices,
                   const int32_t* tris, int32_t num_tris, double* bv_data, int32_t* links,
                   double* sizes, int32_t* depth) {
  if (bv_type != FCLX_BV_OBBRSS && bv_type != FCLX_BV_RSS) return FCLX_ERR_UNSUPPORTED;
  if (num_tris <= 0 || num_vertices <= 0 || !vertices || !tris) return FCLX_ERR_EMPTY_MODEL;
  for (int64_t i = 0; i < 3LL * num_tris; ++i)
    if (tris[i] < 0 || tris[i] >= num_vertices) return FCLX_ERR_INVALID_ARGUMENT;
  HostBVH b;
  if (build_bvh(b, bv_type, vertices, num_vertices, tris, num_tris) != 0)
    return FCLX_ERR_EMPTY_MODEL;
  export_nodes(b, bv_data, links, sizes);
  if (depth) *depth = b.depth;
  return FCLX_OK;
}

}  // extern "C"

namespace fclx {

// fclx_collide_batch with the device status word given by the caller (the
// serving loop keeps one per stream slot, so that a batch's flags are its
// own).
int collide_batch_status(fclx_context* ctx, const fclx_collide_batch_args* a, void* stream,
                         int32_t* d_status) {
  if (!ctx || !a || !d_status) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->num_queries < 0 || a->num_max_contacts < 1 || !a->counts || !a->num_pairs ||
      (a->num_queries > 0 && (!a->query_meshes || !a->tf1 || !a->tf2)) ||
      (a->pair_capacity > 0 && !a->pairs) ||
      (a->enable_contact && a->pairs &&
       (!a->contact_pos || !a->contact_normal || !a->contact_depth)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid collide batch arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int depth, nmesh;
  size_t nodes;
  const MeshDev* table;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    depth = max_depth(ctx);
    nodes = max_nodes(ctx);
    nmesh = static_cast<int>(ctx->meshes.size());
    table = ctx->d_table;
  }
  if (2 * depth > kMaxStack)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "hierarchy deeper than the traversal stack bound");
  (void)hipSetDevice(ctx->device);
  zero_i64<<<1, 64, 0, s>>>(reinterpret_cast<long long*>(a->num_pairs), 1);
  CollideLaunch L{};
  L.meshes = table;
  L.num_meshes = nmesh;
  L.num_queries = a->num_queries;
  L.query_meshes = a->query_meshes;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.cap = a->num_max_contacts;
  L.counts = a->counts;
  L.pairs = a->pairs;
  L.pair_capacity = a->pairs ? a->pair_capacity : 0;
  L.num_pairs = reinterpret_cast<long long*>(a->num_pairs);
  L.bv_tests = reinterpret_cast<long long*>(a->bv_tests);
  L.leaf_tests = reinterpret_cast<long long*>(a->leaf_tests);
  L.status = d_status;
  hipError_t e;
  if (a->enable_contact) {
    // contact geometry follows the reference's order and per-contact cap
    if (a->engine != FCLX_ENGINE_PLAIN && a->engine != FCLX_ENGINE_FAST)
      return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "unknown engine");
    L.contact = 1;
    L.contact_pos = a->contact_pos;
    L.contact_normal = a->contact_normal;
    L.contact_depth = a->contact_depth;
    e = launch_collide_plain(L, 2 * depth, s);
  } else if (a->engine == FCLX_ENGINE_BOOL) {
    if (nodes >= kFastMaxNodes || a->num_queries >= kFastMaxQueries) {
      L.cap = 1;  // min(count, 1)
      e = launch_collide_plain(L, 2 * depth, s);
    } else {
      std::lock_guard<std::mutex> lock(ctx->ws_mu);
      e = engine_anyhit_collide(ctx->ws, L, 2 * depth, s);
    }
  } else if (a->engine == FCLX_ENGINE_PLAIN ||
             (a->engine == FCLX_ENGINE_FAST && a->num_max_contacts > kOrderedCapMax &&
              (nodes >= kFastMaxNodes || a->num_queries >= kFastMaxQueries))) {
    // meshes and batches beyond the FAST engine's task encoding (node and
    // query index widths) take the plain kernel
    e = launch_collide_plain(L, 2 * depth, s);
  } else if (a->engine == FCLX_ENGINE_FAST) {
    std::lock_guard<std::mutex> lock(ctx->ws_mu);
    e = engine_collide(ctx->ws, L, 2 * depth, s);
  } else {
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "unknown engine");
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "collide launch");
  return FCLX_OK;
}

// The FCLX_* code of a set of device status flags (0 when none is set).
int status_code(int32_t flags) {
  if (flags & (kStatusStackOverflow | kStatusQueueOverflow)) return FCLX_ERR_STACK_OVERFLOW;
  if (flags & kStatusBadQuery) return FCLX_ERR_INVALID_ARGUMENT;
  if (flags & (kStatusPairOverflow | kStatusFrontOverflow)) return FCLX_ERR_CAPACITY;
  return FCLX_OK;
}

// Server batch launches still to fail on purpose (see "serve_fail_launches").
bool take_injected_failure(fclx_context* ctx) {
  int n = ctx->serve_fail_launches.load();
  while (n > 0)
    if (ctx->serve_fail_launches.compare_exchange_weak(n, n - 1)) return true;
  return false;
}

}  // namespace fclx

extern "C" {

int fclx_collide_batch(fclx_context* ctx, const fclx_collide_batch_args* a, void* stream) {
  if (!ctx || !a) return FCLX_ERR_INVALID_ARGUMENT;
  return fclx::collide_batch_status(ctx, a, stream, ctx->d_status);
}

int fclx_collide_batch_noalloc(fclx_context* ctx, const fclx_collide_batch_args* a,
                               void* stream) {
  if (!ctx || !a) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->engine != FCLX_ENGINE_BOOL)
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "the no-allocation collide runs FCLX_ENGINE_BOOL");
  if (a->enable_contact)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "the no-allocation collide has no contact geometry");
  if (a->num_queries < 0 || a->num_max_contacts < 1 || !a->counts || !a->num_pairs ||
      (a->num_queries > 0 && (!a->query_meshes || !a->tf1 || !a->tf2)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid collide batch arguments");
  if (a->num_queries >= kFastMaxQueries)
    return fail(ctx, FCLX_ERR_UNSUPPORTED,
                "the no-allocation collide takes fewer than 2^22 queries per call");
  unsigned long long* area = nullptr;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    area = g_scratch_ptr[ctx->device];
    if (!area)
      return fail(ctx, FCLX_ERR_DEVICE,
                  std::string("no device address for the work area: ") +
                      hipGetErrorString(g_scratch_err[ctx->device]));
  } else {
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "device ordinal beyond the work-area table");
  }
  if (!ctx->ws) return fail(ctx, FCLX_ERR_DEVICE, "engine workspace missing");
  int depth, nmesh;
  size_t nodes;
  const MeshDev* table;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    depth = max_depth(ctx);
    nodes = max_nodes(ctx);
    nmesh = static_cast<int>(ctx->meshes.size());
    table = ctx->d_table;
  }
  if (2 * depth > kMaxStack)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "hierarchy deeper than the traversal stack bound");
  if (nodes >= kFastMaxNodes)
    return fail(ctx, FCLX_ERR_UNSUPPORTED,
                "the no-allocation collide takes meshes of fewer than 2^20 triangles");
  EngineTuning tu;
  {
    std::lock_guard<std::mutex> lock(ctx->ws_mu);
    tu = engine_tuning(ctx->ws);
  }
  if (engine_anyhit_min_words(tu, a->num_queries, 2 * depth) > kScratchWords)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "the work area holds too few query stacks");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  zero_i64<<<1, 64, 0, s>>>(reinterpret_cast<long long*>(a->num_pairs), 1);
  CollideLaunch L{};
  L.meshes = table;
  L.num_meshes = nmesh;
  L.num_queries = a->num_queries;
  L.query_meshes = a->query_meshes;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.cap = a->num_max_contacts;
  L.counts = a->counts;
  L.num_pairs = reinterpret_cast<long long*>(a->num_pairs);
  L.bv_tests = reinterpret_cast<long long*>(a->bv_tests);
  L.leaf_tests = reinterpret_cast<long long*>(a->leaf_tests);
  L.status = ctx->d_status;
  const hipError_t e = engine_anyhit_collide_area(tu, L, 2 * depth, area, kScratchWords, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "no-allocation collide launch");
  return FCLX_OK;
}

int fclx_distance_batch(fclx_context* ctx, const fclx_distance_batch_args* a, void* stream) {
  if (!ctx || !a) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->num_queries < 0 || !a->distance ||
      (a->num_queries > 0 && (!a->query_meshes || !a->tf1 || !a->tf2)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid distance batch arguments");
  if (a->engine != FCLX_ENGINE_PLAIN && a->engine != FCLX_ENGINE_FAST)
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "unknown engine");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int depth, nmesh;
  bool merged;
  const MeshDev* table;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    depth = max_depth(ctx);
    nmesh = static_cast<int>(ctx->meshes.size());
    table = ctx->d_table;
    merged = any_merged(ctx);
  }
  if (2 * depth > kMaxStack)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "hierarchy deeper than the traversal stack bound");
  (void)hipSetDevice(ctx->device);
  DistanceLaunch L{};
  L.meshes = table;
  L.num_meshes = nmesh;
  L.num_queries = a->num_queries;
  L.query_meshes = a->query_meshes;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.distance = a->distance;
  L.points = a->enable_nearest_points ? a->points : nullptr;
  L.ids = a->ids;
  L.bv_tests = reinterpret_cast<long long*>(a->bv_tests);
  L.leaf_tests = reinterpret_cast<long long*>(a->leaf_tests);
  L.status = ctx->d_status;
  DistanceTuning tu;
  {
    std::lock_guard<std::mutex> lock(ctx->ws_mu);
    tu = engine_tuning(ctx->ws).dist;
  }
  // The FAST engine finds the reference's answer in any visit order only
  // when every bound is a lower bound of the pairs below it; merged BVs
  // need not be, and then only the reference's own order gives its answer.
  hipError_t e = a->engine == FCLX_ENGINE_FAST && !merged
                     ? launch_distance_wave(L, 2 * depth, tu, s)
                     : launch_distance_plain(L, 2 * depth, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "distance launch");
  return FCLX_OK;
}

int fclx_mesh_shape_collide_batch(fclx_context* ctx, const fclx_mesh_shape_collide_batch_args* a,
                                  void* stream) {
  if (!ctx || !a) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->num_queries < 0 || a->num_shapes < 0 || a->num_vertices < 0 ||
      a->num_max_contacts < 1 || !(a->gjk_tolerance > 0) ||
      (a->num_queries > 0 && (!a->query_meshes || !a->query_shapes || !a->shapes || !a->tf1 ||
                              !a->tf2 || !a->counts)) ||
      (a->num_queries > 0 && a->enable_contact &&
       (!a->contact_pos || !a->contact_normal || !a->contact_depth)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid mesh-shape batch arguments");
  MeshShapeLaunch L{};
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    L.meshes = ctx->d_table;
    L.num_meshes = static_cast<int32_t>(ctx->meshes.size());
    L.max_depth = max_depth(ctx);
  }
  (void)hipSetDevice(ctx->device);
  L.num_queries = a->num_queries;
  L.query_meshes = a->query_meshes;
  L.query_shapes = a->query_shapes;
  L.num_shapes = a->num_shapes;
  L.shapes = a->shapes;
  L.vertices = a->vertices;
  L.num_vertices = a->vertices ? a->num_vertices : 0;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.cap = a->num_max_contacts;
  L.contact = a->enable_contact ? 1 : 0;
  L.tol = a->gjk_tolerance;
  L.counts = a->counts;
  L.prims = a->primitives;
  L.contact_pos = a->contact_pos;
  L.contact_normal = a->contact_normal;
  L.contact_depth = a->contact_depth;
  L.stats = reinterpret_cast<long long*>(a->stats);
  L.status = ctx->d_status;
  int waves;
  {
    std::lock_guard<std::mutex> lock(ctx->ws_mu);
    waves = engine_tuning(ctx->ws).meshshape_waves_per_simd;
    L.leaf_share = engine_tuning(ctx->ws).meshshape_leaf_share;
  }
  hipError_t e = launch_mesh_shape_collide(L, waves, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(ctx, e, "mesh-shape collide launch");
  return FCLX_OK;
}

int fclx_shape_collide_batch(fclx_context* ctx, const fclx_shape_collide_batch_args* a,
                             void* stream) {
  if (!ctx || !a) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->num_pairs < 0 || a->num_shapes < 0 || a->num_vertices < 0 ||
      a->num_max_contacts < 1 || !(a->gjk_tolerance > 0) ||
      (a->num_pairs > 0 &&
       (!a->shapes || !a->pair_shapes || !a->tf1 || !a->tf2 || !a->counts)) ||
      (a->num_pairs > 0 && a->enable_contact &&
       (!a->contact_pos || !a->contact_normal || !a->contact_depth)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid shape batch arguments");
  (void)hipSetDevice(ctx->device);
  std::lock_guard<std::mutex> lock(ctx->sws_mu);
  if (!ctx->sws) ctx->sws = shape_workspace_create(ctx->device);
  shape_workspace_set_epa_blocks_per_cu(ctx->sws, ctx->epa_blocks_per_cu);
  shape_workspace_set_waves(ctx->sws, ctx->shape_waves);
  ShapeLaunch L{};
  L.num_pairs = a->num_pairs;
  L.num_shapes = a->num_shapes;
  L.shapes = a->shapes;
  L.vertices = a->vertices;
  L.num_vertices = a->vertices ? a->num_vertices : 0;
  L.pair_shapes = a->pair_shapes;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.cap = a->num_max_contacts;
  L.contact = a->enable_contact ? 1 : 0;
  L.tol = a->gjk_tolerance;
  L.counts = a->counts;
  L.contact_pos = a->contact_pos;
  L.contact_normal = a->contact_normal;
  L.contact_depth = a->contact_depth;
  L.passes = a->passes;
  L.status = ctx->d_status;
  hipError_t e = shape_collide_run(ctx->sws, L, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(ctx, e, "shape collide launch");
  return FCLX_OK;
}

int fclx_shape_collide(fclx_context* ctx, const fclx_shape* s1, const double* verts1,
                       const double tf1[12], const fclx_shape* s2, const double* verts2,
                       const double tf2[12], int32_t cap, int32_t enable_contact,
                       double gjk_tolerance, double* pos, double* normal, double* depth,
                       int32_t* num_contacts, int32_t* stats) {
  if (!ctx || !s1 || !s2 || !tf1 || !tf2 || !num_contacts || !(gjk_tolerance > 0))
    return FCLX_ERR_INVALID_ARGUMENT;
  *num_contacts = 0;
  if (cap <= 0) return FCLX_OK;
  if (enable_contact && (!pos || !normal || !depth))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "contact outputs are required");
  const fclx_shape* in[2] = {s1, s2};
  const double* vin[2] = {verts1, verts2};
  fclx_shape rec[2];
  int64_t nv = 0;
  for (int k = 0; k < 2; ++k) {
    rec[k] = *in[k];
    rec[k].vertex_offset = 0;
    if (rec[k].type == FCLX_SHAPE_CONVEX) {
      if (!vin[k] || rec[k].num_vertices < 1)
        return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "a convex shape needs its vertices");
      rec[k].vertex_offset = nv;
      nv += rec[k].num_vertices;
    }
  }
  // staging: records 80 B, pair ids 8, poses 192, count 4, passes 8, contacts
  // 4 x 56 B, then the vertices
  const size_t off_ids = 80, off_tf = 88, off_cnt = 280, off_pass = 288, off_geo = 296,
               off_v = 296 + 4 * 56;
  const size_t bytes = off_v + 24 * static_cast<size_t>(nv);
  std::vector<char> host(bytes, 0);
  std::memcpy(host.data(), rec, 80);
  const int32_t ids[2] = {0, 1};
  std::memcpy(host.data() + off_ids, ids, 8);
  std::memcpy(host.data() + off_tf, tf1, 96);
  std::memcpy(host.data() + off_tf + 96, tf2, 96);
  for (int k = 0; k < 2; ++k)
    if (rec[k].type == FCLX_SHAPE_CONVEX)
      std::memcpy(host.data() + off_v + 24 * rec[k].vertex_offset, vin[k],
                  24 * static_cast<size_t>(rec[k].num_vertices));
  const int32_t kcap = std::min<int32_t>(cap, 4);
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (bytes > ctx->s_shape_bytes) {
    (void)hipFree(ctx->s_shape);
    ctx->s_shape = nullptr;
    ctx->s_shape_bytes = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->s_shape, bytes));
    ctx->s_shape_bytes = bytes;
  }
  char* d = ctx->s_shape;
  HIP_TRY(ctx, hipMemcpyAsync(d, host.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
  ShapeLaunch L{};
  L.num_pairs = 1;
  L.num_shapes = 2;
  L.shapes = reinterpret_cast<const fclx_shape*>(d);
  L.vertices = nv ? reinterpret_cast<const double*>(d + off_v) : nullptr;
  L.num_vertices = nv;
  L.pair_shapes = reinterpret_cast<const int32_t*>(d + off_ids);
  L.tf1 = reinterpret_cast<const double*>(d + off_tf);
  L.tf2 = reinterpret_cast<const double*>(d + off_tf + 96);
  L.cap = kcap;
  L.contact = enable_contact ? 1 : 0;
  L.tol = gjk_tolerance;
  L.counts = reinterpret_cast<int32_t*>(d + off_cnt);
  L.contact_pos = reinterpret_cast<double*>(d + off_geo);
  L.contact_normal = reinterpret_cast<double*>(d + off_geo + 96);
  L.contact_depth = reinterpret_cast<double*>(d + off_geo + 192);
  L.passes = reinterpret_cast<int32_t*>(d + off_pass);
  L.status = ctx->d_status;
  {
    std::lock_guard<std::mutex> sl(ctx->sws_mu);
    if (!ctx->sws) ctx->sws = shape_workspace_create(ctx->device);
    shape_workspace_set_epa_blocks_per_cu(ctx->sws, ctx->epa_blocks_per_cu);
    shape_workspace_set_waves(ctx->sws, ctx->shape_waves);
    hipError_t e = shape_collide_run(ctx->sws, L, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "shape collide launch");
  }
  HIP_TRY(ctx, hipMemcpyAsync(host.data(), d, off_v, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  int32_t st = 0;
  HIP_TRY(ctx, hipMemcpy(&st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost));
  if (st & kStatusBadQuery) {
    HIP_TRY(ctx, hipMemset(ctx->d_status, 0, sizeof(int32_t)));
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "malformed shape");
  }
  int32_t n = 0;
  std::memcpy(&n, host.data() + off_cnt, 4);
  *num_contacts = n;
  if (stats) std::memcpy(stats, host.data() + off_pass, 8);
  if (enable_contact) {
    std::memcpy(pos, host.data() + off_geo, 24 * static_cast<size_t>(n));
    std::memcpy(normal, host.data() + off_geo + 96, 24 * static_cast<size_t>(n));
    std::memcpy(depth, host.data() + off_geo + 192, 8 * static_cast<size_t>(n));
  }
  return FCLX_OK;
}

int fclx_set_profiling(fclx_context* ctx, int32_t enable) {
  if (!ctx) return FCLX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  hipError_t e = engine_set_profiling(ctx->ws, enable != 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "engine_set_profiling");
  return FCLX_OK;
}

int fclx_get_engine_param(fclx_context* ctx, const char* name, double* value) {
  if (!ctx || !name || !value || !ctx->ws) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->ws_mu);
  const EngineTuning& t = engine_tuning(ctx->ws);
  const std::string n(name);
  if (n == "max_levels") *value = t.max_levels;
  else if (n == "front_target") *value = static_cast<double>(t.front_target);
  else if (n == "dfs_blocks_per_cu") *value = t.dfs_blocks_per_cu;
  else if (n == "defer_leaf") *value = static_cast<double>(t.defer_leaf);
  else if (n == "defer_uncertain") *value = static_cast<double>(t.defer_uncertain);
  else if (n == "idle_stop") *value = t.idle_stop;
  else if (n == "stop_poll") *value = t.stop_poll;
  else if (n == "dfs_waves_per_simd") *value = t.dfs_waves_per_simd;
  else if (n == "dist_queries_per_wave") *value = t.dist.queries_per_wave;
  else if (n == "dist_waves_per_simd") *value = t.dist.waves_per_simd;
  else if (n == "meshshape_waves_per_simd") *value = t.meshshape_waves_per_simd;
  else if (n == "meshshape_leaf_share") *value = t.meshshape_leaf_share;
  else if (n == "dist_width") *value = t.dist.width;
  else if (n == "dist_leaf_batch") *value = t.dist.leaf_batch;
  else if (n == "dist_fp32_cull") *value = t.dist.fp32_cull;
  else if (n == "fast_persistent") *value = t.fast_persistent;
  else if (n == "dfs_inline_leaves") *value = t.dfs_inline_leaves;
  else if (n == "dfs_stack_top") *value = t.dfs_stack_top;
  else if (n == "fast_graph") *value = t.fast_graph;
  else if (n == "bool_queries_per_wave") *value = t.bool_queries_per_wave;
  else if (n == "bool_waves_per_simd") *value = t.bool_waves_per_simd;
  else if (n == "bool_width") *value = t.bool_width;
  else if (n == "bool_refill") *value = t.bool_refill;
  else if (n == "bool_leaf_batch") *value = t.bool_leaf_batch;
  else if (n == "epa_blocks_per_cu") {
    std::lock_guard<std::mutex> sl(ctx->sws_mu);
    *value = ctx->epa_blocks_per_cu;
  } else if (n == "shape_waves_per_simd") {
    std::lock_guard<std::mutex> sl(ctx->sws_mu);
    *value = ctx->shape_waves;
  } else if (n == "serve_fail_launches") {
    *value = ctx->serve_fail_launches.load();
  } else {
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "unknown engine parameter");
  }
  return FCLX_OK;
}

int fclx_set_engine_param(fclx_context* ctx, const char* name, double value) {
  if (!ctx || !name || !ctx->ws) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->ws_mu);
  EngineTuning& t = engine_tuning(ctx->ws);
  const std::string n(name);
  auto in = [&](double lo, double hi) { return value >= lo && value <= hi; };
  if (n == "max_levels" && in(1, 64)) {
    t.max_levels = static_cast<int>(value);
  } else if (n == "front_target" && in(64, 1LL << 40)) {
    t.front_target = static_cast<long long>(value);
  } else if (n == "dfs_blocks_per_cu" && in(1, 32)) {
    t.dfs_blocks_per_cu = static_cast<int>(value);
  } else if (n == "defer_leaf" && in(64, 1LL << 34)) {
    t.defer_leaf = static_cast<long long>(value);
  } else if (n == "defer_uncertain" && in(64, 1LL << 34)) {
    t.defer_uncertain = static_cast<long long>(value);
  } else if (n == "idle_stop" && in(0, 1)) {
    t.idle_stop = value;
  } else if (n == "stop_poll" && in(1, 1024)) {
    t.stop_poll = static_cast<int>(value);
  } else if (n == "dfs_waves_per_simd" && in(4, 6)) {
    t.dfs_waves_per_simd = static_cast<int>(value);
  } else if (n == "dist_queries_per_wave" && (value == 2 || value == 4 || value == 8)) {
    t.dist.queries_per_wave = static_cast<int>(value);
  } else if (n == "dist_waves_per_simd" && in(2, 3)) {
    t.dist.waves_per_simd = static_cast<int>(value);
  } else if (n == "meshshape_waves_per_simd" && in(1, 3)) {
    t.meshshape_waves_per_simd = static_cast<int>(value);
  } else if (n == "meshshape_leaf_share" && in(0, 100)) {
    t.meshshape_leaf_share = static_cast<int>(value);
  } else if (n == "dist_width" && in(1, 64)) {
    t.dist.width = static_cast<int>(value);
  } else if (n == "dist_leaf_batch" && in(1, 64)) {
    t.dist.leaf_batch = static_cast<int>(value);
  } else if (n == "dist_fp32_cull" && in(0, 1)) {
    t.dist.fp32_cull = static_cast<int>(value);
  } else if (n == "dfs_inline_leaves" && in(0, 1)) {
    t.dfs_inline_leaves = static_cast<int>(value);
  } else if (n == "dfs_stack_top" && (value == 128 || value == 256 || value == 512)) {
    t.dfs_stack_top = static_cast<int>(value);
  } else if (n == "fast_persistent" && in(0, 1)) {
    t.fast_persistent = static_cast<int>(value);
  } else if (n == "fast_graph" && in(0, 1)) {
    t.fast_graph = static_cast<int>(value);
  } else if (n == "bool_queries_per_wave" &&
             (value == 0 || value == 4 || value == 8 || value == 16 || value == 32)) {
    t.bool_queries_per_wave = static_cast<int>(value);
  } else if (n == "bool_waves_per_simd" && in(2, 4)) {
    t.bool_waves_per_simd = static_cast<int>(value);
  } else if (n == "bool_width" && in(1, 64)) {
    t.bool_width = static_cast<int>(value);
  } else if (n == "bool_refill" && in(1, 32)) {
    t.bool_refill = static_cast<int>(value);
  } else if (n == "bool_leaf_batch" && in(1, 64)) {
    t.bool_leaf_batch = static_cast<int>(value);
  } else if (n == "epa_blocks_per_cu" && in(0, 32)) {
    std::lock_guard<std::mutex> sl(ctx->sws_mu);
    ctx->epa_blocks_per_cu = static_cast<int>(value);
  } else if (n == "shape_waves_per_simd" && in(2, 4)) {
    std::lock_guard<std::mutex> sl(ctx->sws_mu);
    ctx->shape_waves = static_cast<int>(value);
  } else if (n == "serve_fail_launches" && in(0, 1 << 20)) {
    ctx->serve_fail_launches.store(static_cast<int>(value));
  } else {
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "unknown engine parameter or value out of range");
  }
  return FCLX_OK;
}

int fclx_engine_last_report(fclx_context* ctx, fclx_engine_report* out) {
  if (!ctx || !out || !ctx->ws) return FCLX_ERR_INVALID_ARGUMENT;
  const EngineReport& r = engine_report(ctx->ws);
  out->setup_ms = r.setup_ms;
  out->expand_ms = r.expand_ms;
  out->leaves_ms = r.leaves_ms;
  out->dfs_ms = r.dfs_ms;
  out->fixup_ms = r.fixup_ms;
  out->expand_bv_tests = r.expand_bv_tests;
  out->expand_leaf_tests = r.expand_leaf_tests;
  out->dfs_bv_tests = r.dfs_bv_tests;
  out->dfs_leaf_tests = r.dfs_leaf_tests;
  out->fp64_fallbacks = r.fp64_fallbacks;
  out->front_size = r.front_size;
  out->levels = r.levels;
  out->flagged = r.flagged;
  out->dfs_waves = r.dfs_waves;
  out->dfs_rounds = r.dfs_rounds;
  out->dfs_wave_busy = r.dfs_wave_busy;
  out->dfs_end_p50_ms = r.dfs_end_p50_ms;
  out->dfs_end_p90_ms = r.dfs_end_p90_ms;
  out->dfs_end_max_ms = r.dfs_end_max_ms;
  for (int k = 0; k < 3; ++k) {
    out->bool_steps[k] = r.bool_steps[k];
    out->bool_lanes[k] = r.bool_lanes[k];
  }
  out->bool_skipped = r.bool_skipped;
  out->bool_rounds = r.bool_rounds;
  out->bool_stack_max = r.bool_stack_max;
  out->bool_wave_ms = r.bool_wave_ms;
  out->bool_end_p50_ms = r.bool_end_p50_ms;
  out->bool_end_p90_ms = r.bool_end_p90_ms;
  out->bool_span_ms = r.bool_span_ms;
  out->dfs_inline_leaf_tests = r.dfs_inline_leaf_tests;
  out->dfs_wave_steps = r.dfs_wave_steps;
  return FCLX_OK;
}

struct fclx_front_set {
  fclx_context* ctx = nullptr;
  int32_t num_sequences = 0;
  int32_t capacity = 0;
  int2* fronts = nullptr;
  int2* second = nullptr;
  int32_t* sizes = nullptr;
  int32_t* meshes = nullptr;
};

namespace {
void free_front_set(fclx_front_set* f) {
  (void)hipFree(f->fronts);
  (void)hipFree(f->second);
  (void)hipFree(f->sizes);
  (void)hipFree(f->meshes);
  delete f;
}
}  // namespace

int fclx_front_set_create(fclx_context* ctx, int32_t num_sequences, int32_t capacity,
                          fclx_front_set** out) {
  if (!ctx || !out || num_sequences < 1 || capacity < 1) return FCLX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  (void)hipSetDevice(ctx->device);
  auto* f = new (std::nothrow) fclx_front_set();
  if (!f) return FCLX_ERR_OUT_OF_MEMORY;
  f->ctx = ctx;
  f->num_sequences = num_sequences;
  f->capacity = capacity;
  const size_t entries = static_cast<size_t>(num_sequences) * capacity;
  hipError_t e;
  if ((e = hipMalloc(&f->fronts, entries * sizeof(int2))) != hipSuccess ||
      (e = hipMalloc(&f->second, entries * sizeof(int2))) != hipSuccess ||
      (e = hipMalloc(&f->sizes, num_sequences * sizeof(int32_t))) != hipSuccess ||
      (e = hipMalloc(&f->meshes, 2 * num_sequences * sizeof(int32_t))) != hipSuccess) {
    free_front_set(f);
    return hip_fail(ctx, e, "front set allocation");
  }
  clear_fronts<<<(num_sequences + 255) / 256, 256, 0, ctx->stream>>>(f->sizes, f->meshes,
                                                                    num_sequences);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
    free_front_set(f);
    return hip_fail(ctx, e, "front set clear");
  }
  *out = f;
  return FCLX_OK;
}

int fclx_front_set_destroy(fclx_front_set* f) {
  if (!f) return FCLX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(f->ctx->device);
  free_front_set(f);
  return FCLX_OK;
}

int fclx_front_set_clear(fclx_front_set* f, void* stream) {
  if (!f) return FCLX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(f->ctx->device);
  clear_fronts<<<(f->num_sequences + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      f->sizes, f->meshes, f->num_sequences);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? FCLX_OK : hip_fail(f->ctx, e, "front set clear");
}

int fclx_front_set_read(fclx_front_set* f, int32_t sequence, int32_t* pairs, int32_t room,
                        int32_t* n) {
  if (!f || !n || sequence < 0 || sequence >= f->num_sequences || room < 0 ||
      (room > 0 && !pairs))
    return FCLX_ERR_INVALID_ARGUMENT;
  fclx_context* ctx = f->ctx;
  (void)hipSetDevice(ctx->device);
  int32_t size = 0;
  HIP_TRY(ctx, hipMemcpy(&size, f->sizes + sequence, sizeof(size), hipMemcpyDeviceToHost));
  *n = size;
  const int32_t k = std::min(size, room);
  if (k > 0)
    HIP_TRY(ctx, hipMemcpy(pairs, f->fronts + static_cast<size_t>(sequence) * f->capacity,
                           static_cast<size_t>(k) * sizeof(int2), hipMemcpyDeviceToHost));
  return FCLX_OK;
}

int fclx_collide_front_batch(fclx_context* ctx, const fclx_collide_front_batch_args* a,
                             fclx_front_set* f, void* stream) {
  if (!ctx || !a || !f || f->ctx != ctx) return FCLX_ERR_INVALID_ARGUMENT;
  if (a->num_sequences < 0 || a->num_sequences > f->num_sequences || a->num_max_contacts < 1 ||
      !a->counts || !a->num_pairs || (a->pair_capacity > 0 && !a->pairs) ||
      (a->num_sequences > 0 &&
       (!a->seq_offsets || !a->query_meshes || !a->tf1 || !a->tf2)))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid front batch arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FrontLaunch L{};
  int depth;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    depth = max_depth(ctx);
    L.meshes = ctx->d_table;
    L.num_meshes = static_cast<int32_t>(ctx->meshes.size());
  }
  if (2 * depth > kMaxStack)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "hierarchy deeper than the traversal stack bound");
  (void)hipSetDevice(ctx->device);
  zero_i64<<<1, 64, 0, s>>>(reinterpret_cast<long long*>(a->num_pairs), 1);
  L.num_sequences = a->num_sequences;
  L.seq_offsets = a->seq_offsets;
  L.query_meshes = a->query_meshes;
  L.tf1 = a->tf1;
  L.tf2 = a->tf2;
  L.cap = a->num_max_contacts;
  L.counts = a->counts;
  L.pairs = a->pairs;
  L.pair_capacity = a->pairs ? a->pair_capacity : 0;
  L.num_pairs = reinterpret_cast<long long*>(a->num_pairs);
  L.bv_tests = reinterpret_cast<long long*>(a->bv_tests);
  L.leaf_tests = reinterpret_cast<long long*>(a->leaf_tests);
  L.fronts = f->fronts;
  L.second = f->second;
  L.front_sizes = f->sizes;
  L.front_meshes = f->meshes;
  L.front_cap = f->capacity;
  L.status = ctx->d_status;
  const hipError_t e = launch_collide_front(L, 2 * depth, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "front collide launch");
  return FCLX_OK;
}

int fclx_broadphase_pairs(fclx_context* ctx, int32_t num_objects, const double* aabbs,
                          const int32_t* groups, int32_t* pairs, int64_t pair_capacity,
                          int64_t* num_pairs, void* stream) {
  if (!ctx || num_objects < 0 || !num_pairs || pair_capacity < 0 ||
      (num_objects > 0 && !aabbs) || (pair_capacity > 0 && !pairs) ||
      pair_capacity > (int64_t(1) << 31) - 1)
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "invalid broadphase arguments");
  (void)hipSetDevice(ctx->device);
  const hipError_t e = launch_broadphase(num_objects, aabbs, groups, pairs, pair_capacity,
                                         reinterpret_cast<long long*>(num_pairs),
                                         static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(ctx, e, "broadphase launch");
  return FCLX_OK;
}

int fclx_sync(fclx_context* ctx, void* stream) {
  if (!ctx) return FCLX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  HIP_TRY(ctx, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  int32_t st = 0;
  HIP_TRY(ctx, hipMemcpy(&st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost));
  if (st != 0) {
    HIP_TRY(ctx, hipMemset(ctx->d_status, 0, sizeof(int32_t)));
    if (st & (kStatusStackOverflow | kStatusQueueOverflow))
      return fail(ctx, FCLX_ERR_STACK_OVERFLOW, "a traversal exceeded its stack bound");
    if (st & kStatusBadQuery)
      return fail(ctx, FCLX_ERR_INVALID_ARGUMENT,
                  "a query named a missing mesh or shape, or an unsupported BV pair");
    if (st & kStatusPairOverflow)
      return fail(ctx, FCLX_ERR_CAPACITY, "pair output exceeded pair_capacity");
    if (st & kStatusFrontOverflow)
      return fail(ctx, FCLX_ERR_CAPACITY, "a BVTT front outgrew its front-set capacity");
  }
  return FCLX_OK;
}

namespace {

// One collide() query through the reference-order kernel with host staging;
// contact mode also returns the geometry of each contact.
int collide_one(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                const double tf2[12], int32_t cap, int32_t* pairs, double* pos, double* normal,
                double* depth, int32_t capacity, int32_t* num_contacts, int64_t* stats,
                bool contact) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh1) || !valid_mesh(ctx, mesh2))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  if (ctx->meshes[mesh1].bvh.bv_type != FCLX_BV_OBBRSS ||
      ctx->meshes[mesh2].bvh.bv_type != FCLX_BV_OBBRSS)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "mesh collide needs OBBRSS hierarchies");
  (void)hipSetDevice(ctx->device);
  double tf[24];
  std::memcpy(tf, tf1, 12 * sizeof(double));
  std::memcpy(tf + 12, tf2, 12 * sizeof(double));
  const int32_t qm[2] = {mesh1, mesh2};
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->s_tf, tf, sizeof(tf), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->s_qm, qm, sizeof(qm), hipMemcpyHostToDevice, s));
  const int depth_sum = ctx->meshes[mesh1].bvh.depth + ctx->meshes[mesh2].bvh.depth;
  if (depth_sum > kMaxStack)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "hierarchy deeper than the traversal stack bound");
  // device staging for the records starts small and grows to the count found
  long long want = std::min<long long>(cap, 4096);
  int32_t count = 0;
  long long st[2] = {0, 0};
  long long produced = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (want > ctx->s_pairs_cap) {
      (void)hipFree(ctx->s_pairs);
      (void)hipFree(ctx->s_geo);
      ctx->s_pairs = nullptr;
      ctx->s_geo = nullptr;
      ctx->s_pairs_cap = 0;
      HIP_TRY(ctx, dmalloc(&ctx->s_pairs, 3 * static_cast<size_t>(want)));
      HIP_TRY(ctx, dmalloc(&ctx->s_geo, 7 * static_cast<size_t>(want)));
      ctx->s_pairs_cap = want;
    }
    zero_i64<<<1, 64, 0, s>>>(ctx->s_num, 1);
    CollideLaunch L{};
    L.meshes = ctx->d_table;
    L.num_meshes = static_cast<int32_t>(ctx->meshes.size());
    L.num_queries = 1;
    L.query_meshes = ctx->s_qm;
    L.tf1 = ctx->s_tf;
    L.tf2 = ctx->s_tf + 12;
    L.cap = cap;
    L.counts = ctx->s_count;
    L.pairs = ctx->s_pairs;
    L.pair_capacity = ctx->s_pairs_cap;
    L.num_pairs = ctx->s_num;
    L.bv_tests = ctx->s_stats;
    L.leaf_tests = ctx->s_stats + 1;
    L.status = ctx->d_status;
    if (contact) {
      L.contact = 1;
      L.contact_pos = ctx->s_geo;
      L.contact_normal = ctx->s_geo + 3 * ctx->s_pairs_cap;
      L.contact_depth = ctx->s_geo + 6 * ctx->s_pairs_cap;
    }
    HIP_TRY(ctx, launch_collide_plain(L, depth_sum, s));
    HIP_TRY(ctx, hipMemcpyAsync(&count, ctx->s_count, sizeof(count), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(st, ctx->s_stats, sizeof(st), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(&produced, ctx->s_num, sizeof(produced), hipMemcpyDeviceToHost,
                                s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (produced <= ctx->s_pairs_cap) break;
    want = produced;  // every contact of this query fits on the second run
  }
  int32_t flag = 0;
  HIP_TRY(ctx, hipMemcpy(&flag, ctx->d_status, sizeof(flag), hipMemcpyDeviceToHost));
  if (flag) HIP_TRY(ctx, hipMemset(ctx->d_status, 0, sizeof(int32_t)));
  if (flag & kStatusStackOverflow)
    return fail(ctx, FCLX_ERR_STACK_OVERFLOW, "traversal stack overflow");
  *num_contacts = count;
  if (stats) {
    stats[0] = st[0];
    stats[1] = st[1];
  }
  const int32_t ncopy = std::min(count, capacity);
  if (ncopy > 0) {
    std::vector<int32_t> rec(3 * static_cast<size_t>(ncopy));
    HIP_TRY(ctx, hipMemcpy(rec.data(), ctx->s_pairs, sizeof(int32_t) * rec.size(),
                           hipMemcpyDeviceToHost));
    for (int i = 0; i < ncopy; ++i) {
      pairs[2 * i] = rec[3 * i + 1];
      pairs[2 * i + 1] = rec[3 * i + 2];
    }
    if (contact) {
      const size_t cap_rec = static_cast<size_t>(ctx->s_pairs_cap);
      HIP_TRY(ctx, hipMemcpy(pos, ctx->s_geo, sizeof(double) * 3 * ncopy, hipMemcpyDeviceToHost));
      HIP_TRY(ctx, hipMemcpy(normal, ctx->s_geo + 3 * cap_rec, sizeof(double) * 3 * ncopy,
                             hipMemcpyDeviceToHost));
      HIP_TRY(ctx, hipMemcpy(depth, ctx->s_geo + 6 * cap_rec, sizeof(double) * ncopy,
                             hipMemcpyDeviceToHost));
    }
  }
  if (count > capacity)
    return fail(ctx, FCLX_ERR_CAPACITY, "pairs buffer smaller than the contact count");
  return FCLX_OK;
}

}  // namespace

int fclx_collide(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                 const double tf2[12], int32_t cap, int32_t* pairs, int32_t pairs_capacity,
                 int32_t* num_contacts, int64_t* stats) {
  if (!ctx || !tf1 || !tf2 || !num_contacts || cap < 1 || pairs_capacity < 0 ||
      (pairs_capacity > 0 && !pairs))
    return FCLX_ERR_INVALID_ARGUMENT;
  return collide_one(ctx, mesh1, tf1, mesh2, tf2, cap, pairs, nullptr, nullptr, nullptr,
                     pairs_capacity, num_contacts, stats, false);
}

int fclx_collide_contact(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                         const double tf2[12], int32_t cap, int32_t* pairs, double* pos,
                         double* normal, double* depth, int32_t capacity,
                         int32_t* num_contacts, int64_t* stats) {
  if (!ctx || !tf1 || !tf2 || !num_contacts || cap < 1 || capacity < 0 ||
      (capacity > 0 && (!pairs || !pos || !normal || !depth)))
    return FCLX_ERR_INVALID_ARGUMENT;
  return collide_one(ctx, mesh1, tf1, mesh2, tf2, cap, pairs, pos, normal, depth, capacity,
                     num_contacts, stats, true);
}

int fclx_distance(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                  const double tf2[12], int32_t enable_nearest_points, double* min_distance,
                  double* points, int32_t* ids, int64_t* stats) {
  if (!ctx || !tf1 || !tf2 || !min_distance) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!valid_mesh(ctx, mesh1) || !valid_mesh(ctx, mesh2))
    return fail(ctx, FCLX_ERR_INVALID_ARGUMENT, "bad mesh id");
  if (ctx->meshes[mesh1].bvh.bv_type != ctx->meshes[mesh2].bvh.bv_type)
    return fail(ctx, FCLX_ERR_UNSUPPORTED, "both meshes must use the same BV type");
  (void)hipSetDevice(ctx->device);
  double tf[24];
  std::memcpy(tf, tf1, 12 * sizeof(double));
  std::memcpy(tf + 12, tf2, 12 * sizeof(double));
  const int32_t qm[2] = {mesh1, mesh2};
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->s_tf, tf, sizeof(tf), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->s_qm, qm, sizeof(qm), hipMemcpyHostToDevice, s));
  DistanceLaunch L{};
  L.meshes = ctx->d_table;
  L.num_meshes = static_cast<int32_t>(ctx->meshes.size());
  L.num_queries = 1;
  L.query_meshes = ctx->s_qm;
  L.tf1 = ctx->s_tf;
  L.tf2 = ctx->s_tf + 12;
  L.distance = ctx->s_dist;
  L.points = enable_nearest_points ? ctx->s_dist + 1 : nullptr;
  L.ids = ctx->s_ids;
  L.bv_tests = ctx->s_stats;
  L.leaf_tests = ctx->s_stats + 1;
  L.status = ctx->d_status;
  const int depth = ctx->meshes[mesh1].bvh.depth + ctx->meshes[mesh2].bvh.depth;
  HIP_TRY(ctx, launch_distance_plain(L, depth, s));
  double out[7];
  int32_t idv[2];
  long long st[2];
  HIP_TRY(ctx, hipMemcpyAsync(out, ctx->s_dist, sizeof(out), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(idv, ctx->s_ids, sizeof(idv), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(st, ctx->s_stats, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  *min_distance = out[0];
  if (points && enable_nearest_points) std::memcpy(points, out + 1, 6 * sizeof(double));
  if (ids) {
    ids[0] = idv[0];
    ids[1] = idv[1];
  }
  if (stats) {
    stats[0] = st[0];
    stats[1] = st[1];
  }
  int32_t flag = 0;
  HIP_TRY(ctx, hipMemcpy(&flag, ctx->d_status, sizeof(flag), hipMemcpyDeviceToHost));
  if (flag & kStatusStackOverflow) {
    HIP_TRY(ctx, hipMemset(ctx->d_status, 0, sizeof(int32_t)));
    return fail(ctx, FCLX_ERR_STACK_OVERFLOW, "traversal stack overflow");
  }
  return FCLX_OK;
}

}  // extern "C"
