// Serving: a request loop that renders frames of deforming scenes.
//
// Scenes stay resident in one device arena whose size is the HBM budget;
// a request is admitted when a slot (a stream with its own workspace and
// frame buffers) is free and its scene is resident or can be made resident
// by evicting scenes no request is using (least recently used first).
// Admitted requests run concurrently, one per slot: the GPU overlaps their
// kernels while the host keeps admitting. Per request, on its slot's stream:
// upload the scene if it was not resident, deform its vertices for the
// request's time (the same launch clears a refit's split counters), rebuild
// or refit its tree, trace the camera frame, and checksum it; the checksum
// launch's last workgroup writes the checksum and then the request's ticket
// to pinned host memory, which is how the host sees the request finish (four
// launches per refit request). Requests of one scene
// are ordered, since they share its buffers: a request whose scene's latest
// request still runs on another slot waits for an event recorded there.
//
// Nothing here allocates per request: the arena, the slot workspaces and the
// pinned result words are allocated at creation or when a scene is added.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "rk_device.h"
#include "rk_scenes.h"

// rk_fast.hip
rk_status rk_frame_checksum_to_host(hipStream_t st, const uint32_t* faces_id, const rk_float2* coords,
                                    const float* depths, uint32_t count, const uint32_t* indices,
                                    uint32_t face_count, uint64_t* partials, uint32_t* counter, uint64_t* out_host,
                                    uint64_t done_value, const uint32_t* build_flags);
rk_status rk_deform_and_clear(hipStream_t st, const rk_float4* base, rk_float4* out, uint32_t count, float amplitude,
                              float omega, float time, uint64_t seed, int along_z, uint32_t* clear,
                              uint32_t clear_n);
uint32_t* rk_fast_refit_counters(void* ws, size_t ws_bytes, uint32_t n);
rk_status rk_fast_refit_precleared(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh,
                                   const void* sorted_keys, void* ws, size_t ws_bytes);

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kSumParts = 2 * 256;  // two words per checksum workgroup, at most 256 of them


struct Scene {
  rk_float4* h_verts = nullptr;  // pinned host copies, the source of uploads
  rk_uint4* h_faces = nullptr;
  uint32_t nv = 0, nf = 0;
  int along_z = 0;
  rk_aabox hitbox{};  // base hitbox grown by the deformation amplitude
  float amplitude = 0.0f;
  size_t bytes = 0;
  bool resident = false;
  size_t off = 0;
  rk_float4 *base = nullptr, *verts = nullptr;
  rk_uint4* faces = nullptr;
  void *nodes = nullptr, *tris = nullptr, *keys = nullptr;
  uint32_t *indices = nullptr, *info = nullptr;
  bool uploaded = false;  // base vertices and faces are on the device
  bool built = false;     // the tree matches the faces (keys kept for refits)
  uint32_t refits = 0;
  int in_use = 0;
  int last_slot = -1;
  uint64_t last_use = 0;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  uint32_t* fid = nullptr;
  rk_float2* uv = nullptr;
  float* dep = nullptr;
  uint64_t* sum_parts = nullptr;   // per-workgroup shares of the checksum, then a zeroed counter
  uint64_t* sum_host = nullptr;    // pinned, coherent: checksum, hits, ticket + 1 once the frame is done, build flags
  uint64_t* sum_host_dev = nullptr;  // the device's address of sum_host
  rk_float4 *rays_o = nullptr, *rays_d = nullptr;  // plain path: primary rays
  uint64_t* sum_dev = nullptr;                      // plain path: checksum in device memory
  int64_t ticket = -1;
  bool recorded = false;  // 'done' marks the end of the running request
  uint32_t polls = 0;
};

struct Pending {
  rk_request req;
  uint64_t ticket;
};

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

size_t scene_bytes(uint32_t nv, uint32_t nf) {
  return align_up(16ull * nv) * 2 + align_up(16ull * nf) + align_up(64ull * nf) + align_up(48ull * nf) +
         align_up(4ull * nf) * 2 + align_up(4ull * RK_INFO_WORDS);
}

}  // namespace

struct rk_server_st {
  rk_server_config cfg{};
  char* arena = nullptr;
  size_t arena_bytes = 0;
  std::map<size_t, size_t> free_list;  // offset -> size, coalesced
  size_t resident_bytes = 0, peak_resident = 0;
  std::vector<Scene> scenes;
  std::vector<Slot> slots;
  std::deque<Pending> queue;
  std::vector<rk_result> results;
  std::vector<uint8_t> finished;
  uint64_t tick = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  uint64_t rebuilds = 0, refits = 0, uploads = 0, evictions = 0, failed_requests = 0;
  rk_status failed = RK_SUCCESS;  // the first launch failure; the server admits nothing after it

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_list.begin(); it != free_list.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, at = it->first + bytes;
        free_list.erase(it);
        if (rest) free_list[at] = rest;
        return true;
      }
    }
    return false;
  }

  void release(size_t off, size_t bytes) {
    auto it = free_list.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_list.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_list.erase(next);
    }
    if (it != free_list.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_list.erase(it);
      }
    }
  }

  void place(Scene& s) {
    char* p = arena + s.off;
    auto take = [&](size_t b) { char* q = p; p += align_up(b); return q; };
    s.base = reinterpret_cast<rk_float4*>(take(16ull * s.nv));
    s.verts = reinterpret_cast<rk_float4*>(take(16ull * s.nv));
    s.faces = reinterpret_cast<rk_uint4*>(take(16ull * s.nf));
    s.nodes = take(64ull * s.nf);
    s.tris = take(48ull * s.nf);
    s.indices = reinterpret_cast<uint32_t*>(take(4ull * s.nf));
    s.keys = take(4ull * s.nf);
    s.info = reinterpret_cast<uint32_t*>(take(4ull * RK_INFO_WORDS));
  }

  void evict(Scene& s) {
    release(s.off, s.bytes);
    s.resident = false;
    s.uploaded = false;
    s.built = false;
    resident_bytes -= s.bytes;
    ++evictions;
  }

  // Makes scene 'id' resident, evicting idle scenes (least recently used
  // first). Returns false when the scene cannot fit until running requests end.
  bool make_resident(uint32_t id, uint32_t* evicted) {
    Scene& s = scenes[id];
    if (s.resident) return true;
    size_t off = 0;
    while (!alloc(s.bytes, &off)) {
      Scene* lru = nullptr;
      for (auto& c : scenes)
        if (c.resident && c.in_use == 0 && (!lru || c.last_use < lru->last_use)) lru = &c;
      if (!lru) return false;
      evict(*lru);
      ++*evicted;
    }
    s.off = off;
    s.resident = true;
    s.uploaded = false;
    s.built = false;
    resident_bytes += s.bytes;
    peak_resident = std::max(peak_resident, resident_bytes);
    place(s);
    return true;
  }

  void mark_running(Slot& sl, Scene& s, int slot_index, uint64_t ticket) {
    sl.ticket = (int64_t)ticket;
    sl.recorded = false;
    sl.polls = 0;
    s.in_use++;
    s.last_slot = slot_index;
    s.last_use = ++tick;
  }

  // The plain per-request path (rk_server_config.plain): what a caller of the
  // plain entry points does per frame, in the same order of results.
  rk_status launch_plain(Slot& sl, Scene& s, const rk_request& r, uint64_t ticket, rk_result& res) {
    rk_stream st = reinterpret_cast<rk_stream>(sl.st);
    rk_status e;
    rk_fast_bvh bvh = {s.nodes, s.tris, s.indices, nullptr, s.nf, 0};
    if ((e = rk_deform_vertices(st, s.base, s.verts, s.nv, s.amplitude, cfg.omega, r.time, 0x5EED0000ull + r.scene,
                                s.along_z)) != RK_SUCCESS)
      return e;
    if ((e = rk_fast_build(st, s.faces, s.verts, s.nf, &s.hitbox, RK_MORTON_30, cfg.leaf_max, &bvh, s.info, sl.ws,
                           sl.ws_bytes)) != RK_SUCCESS)
      return e;
    s.built = false;  // no keys kept: a later request of the request loop rebuilds
    ++rebuilds;
    res.rebuilt = 1;
    const uint32_t n = r.width * r.height;
    uint32_t* fid = r.faces_id ? r.faces_id : sl.fid;
    rk_float2* uv = r.coords ? r.coords : sl.uv;
    float* dep = r.depths ? r.depths : sl.dep;
    // the frame as the camera kernel leaves misses: face ~0, coords 0, depth FLT_MAX
    if ((e = rk_fill_u32(st, fid, n, 0xFFFFFFFFu)) != RK_SUCCESS || (e = rk_fill_f2(st, uv, n, {0.0f, 0.0f})) !=
        RK_SUCCESS || (e = rk_fill_f32(st, dep, n, RK_HUGE)) != RK_SUCCESS)
      return e;
    if ((e = rk_calc_primary_rays(st, (int)r.width, (int)r.height, r.camera.origin, r.camera.direction,
                                  r.camera.step_x, r.camera.step_y, sl.rays_o, sl.rays_d)) != RK_SUCCESS)
      return e;
    rk_trace_options opt = {cfg.det_epsilon, 0u, 0u, 0u};
    if ((e = rk_fast_trace(st, &bvh, sl.rays_o, sl.rays_d, n, fid, uv, dep, &opt)) != RK_SUCCESS) return e;
    if ((e = rk_frame_checksum(st, fid, uv, dep, n, s.indices, s.nf, sl.sum_dev)) != RK_SUCCESS) return e;
    RK_HIP(hipMemcpyAsync(sl.sum_host, sl.sum_dev, 16, hipMemcpyDeviceToHost, sl.st));
    uint32_t flags = 0;
    RK_HIP(hipMemcpyAsync(&flags, s.info + RK_INFO_FLAGS, 4, hipMemcpyDeviceToHost, sl.st));
    RK_HIP(hipStreamSynchronize(sl.st));  // one request at a time, finished before the next
    sl.sum_host[3] = flags;
    sl.sum_host[2] = ticket + 1;
    return RK_SUCCESS;
  }

  rk_status launch(Slot& sl, int slot_index, const rk_request& r, uint64_t ticket, uint32_t evicted) {
    Scene& s = scenes[r.scene];
    rk_status e = enqueue(sl, slot_index, s, r, ticket, evicted);
    if (e != RK_SUCCESS) {
      // work for this scene may be partly enqueued: its device copy is no
      // longer trusted, so a later request uploads and rebuilds it
      s.uploaded = false;
      s.built = false;
      failed = e;
      return e;
    }
    mark_running(sl, s, slot_index, ticket);
    return RK_SUCCESS;
  }

  rk_status enqueue(Slot& sl, int slot_index, Scene& s, const rk_request& r, uint64_t ticket, uint32_t evicted) {
    rk_result& res = results[ticket];
    res.start = now();
    res.evictions = evicted;
    res.slot = (uint32_t)slot_index;
    if (s.in_use > 0 && s.last_slot >= 0 && s.last_slot != slot_index) {
      // the scene's latest request still runs on another slot: mark its end
      // there (that slot holds nothing else) and wait for it
      Slot& prev = slots[s.last_slot];
      if (!prev.recorded) {
        RK_HIP(hipEventRecord(prev.done, prev.st));
        prev.recorded = true;
      }
      RK_HIP(hipStreamWaitEvent(sl.st, prev.done, 0));
    }
    if (!s.uploaded) {  // a scene made resident for this request starts from its host copy
      RK_HIP(hipMemcpyAsync(s.base, s.h_verts, 16ull * s.nv, hipMemcpyHostToDevice, sl.st));
      RK_HIP(hipMemcpyAsync(s.faces, s.h_faces, 16ull * s.nf, hipMemcpyHostToDevice, sl.st));
      s.uploaded = true;
      s.built = false;
      res.uploaded = 1;
      ++uploads;
    }
    return cfg.plain ? launch_plain(sl, s, r, ticket, res) : launch_loop(sl, s, r, ticket, res);
  }

  // The request loop's work for one request: deform, refit or rebuild, trace
  // the camera frame, checksum (with the build's flags) to host memory.
  rk_status launch_loop(Slot& sl, Scene& s, const rk_request& r, uint64_t ticket, rk_result& res) {
    rk_stream st = reinterpret_cast<rk_stream>(sl.st);
    rk_status e;
    rk_fast_bvh bvh = {s.nodes, s.tris, s.indices, nullptr, s.nf, 0};
    const bool force_rebuild = (r.flags & RK_REQUEST_FORCE_REBUILD) != 0;
    const bool force_refit = (r.flags & RK_REQUEST_FORCE_REFIT) != 0;
    const bool rebuild = !s.built || force_rebuild ||
                         (!force_refit && (cfg.rebuild_every == 0 || s.refits >= cfg.rebuild_every));
    // A refit's split counters are cleared by the deform launch; a rebuild
    // clears its own in its key kernel.
    uint32_t* counters = rebuild ? nullptr : rk_fast_refit_counters(sl.ws, sl.ws_bytes, s.nf);
    if (!rebuild && !counters) return RK_ERROR_WORKSPACE;
    if ((e = rk_deform_and_clear(sl.st, s.base, s.verts, s.nv, s.amplitude, cfg.omega, r.time,
                                 0x5EED0000ull + r.scene, s.along_z, counters, rebuild ? 0u : s.nf)) != RK_SUCCESS)
      return e;
    if (rebuild) {
      s.built = false;
      e = rk_fast_build_keyed(st, s.faces, s.verts, s.nf, &s.hitbox, RK_MORTON_30, cfg.leaf_max, &bvh, s.info, s.keys,
                              sl.ws, sl.ws_bytes);
      if (e != RK_SUCCESS) return e;
      s.built = true;  // until the frame's checksum reports the build's flags
      s.refits = 0;
      ++rebuilds;
    } else {
      e = rk_fast_refit_precleared(st, s.faces, s.verts, s.nf, RK_MORTON_30, cfg.leaf_max, &bvh, s.keys, sl.ws,
                                   sl.ws_bytes);
      ++s.refits;
      ++refits;
    }
    if (e != RK_SUCCESS) return e;
    res.rebuilt = rebuild ? 1 : 0;
    uint32_t* fid = r.faces_id ? r.faces_id : sl.fid;
    rk_float2* uv = r.coords ? r.coords : sl.uv;
    float* dep = r.depths ? r.depths : sl.dep;
    rk_trace_options opt = {cfg.det_epsilon, 0u, cfg.trace_config, 0u};
    if ((e = rk_fast_trace_camera(st, &bvh, &r.camera, (int)r.width, (int)r.height, fid, uv, dep, &opt)) !=
        RK_SUCCESS)
      return e;
    if ((e = rk_frame_checksum_to_host(sl.st, fid, uv, dep, r.width * r.height, s.indices, s.nf, sl.sum_parts,
                                       reinterpret_cast<uint32_t*>(sl.sum_parts + kSumParts), sl.sum_host_dev,
                                       ticket + 1, s.info + RK_INFO_FLAGS)) != RK_SUCCESS)
      return e;
    return RK_SUCCESS;
  }

  rk_status progress(uint32_t* completed) {
    uint32_t done_now = 0;
    for (auto& sl : slots) {
      if (sl.ticket < 0) continue;
      // the checksum kernel, the request's last, writes ticket + 1 last
      if (__atomic_load_n(&sl.sum_host[2], __ATOMIC_ACQUIRE) != (uint64_t)sl.ticket + 1) {
        if ((++sl.polls & 1023u) == 0) {  // now and then, ask the runtime whether the stream failed
          hipError_t q = hipStreamQuery(sl.st);
          if (q != hipSuccess && q != hipErrorNotReady) return RK_ERROR_DEVICE;
        }
        continue;
      }
      rk_result& res = results[(size_t)sl.ticket];
      res.done = now();
      res.checksum = sl.sum_host[0];
      res.hits = sl.sum_host[1];
      res.build_flags = (uint32_t)sl.sum_host[3];
      finished[(size_t)sl.ticket] = 1;
      Scene& s = scenes[res.scene];
      if (res.build_flags & RK_FLAG_SPIN_LIMIT) {
        // a look-back wait of the build's sort gave up: the tree and this
        // frame are invalid; the scene's next request rebuilds
        res.status = RK_ERROR_DEVICE;
        s.built = false;
        ++failed_requests;
      }
      s.in_use--;
      sl.ticket = -1;
      ++done_now;
    }
    // admit in arrival order while a slot is free and the head fits; nothing
    // more after a launch failed
    while (!queue.empty() && failed == RK_SUCCESS) {
      int free_slot = -1;
      for (size_t i = 0; i < slots.size(); ++i)
        if (slots[i].ticket < 0) { free_slot = (int)i; break; }
      if (free_slot < 0) break;
      Pending& p = queue.front();
      uint32_t evicted = 0;
      if (!make_resident(p.req.scene, &evicted)) break;  // wait for running requests to free memory
      rk_status e = launch(slots[(size_t)free_slot], free_slot, p.req, p.ticket, evicted);
      if (e != RK_SUCCESS) return e;
      queue.pop_front();
    }
    if (completed) *completed = done_now;
    return failed;
  }

  bool idle() const {
    if (!queue.empty()) return false;
    for (auto& sl : slots)
      if (sl.ticket >= 0) return false;
    return true;
  }
};

extern "C" rk_server_config rk_server_config_default(void) {
  rk_server_config c = {};
  c.slots = 16;
  c.leaf_max = 1;
  c.rebuild_every = 8;
  c.max_faces = 2u << 20;
  c.max_pixels = 1024u * 1024u;
  c.det_epsilon = 0.0f;
  c.amplitude_frac = 0.01f;
  c.omega = 6.2831853f;
  c.hbm_budget = 200ull << 30;
  return c;
}

extern "C" rk_status rk_server_create(const rk_server_config* cfg, uint64_t arena_bytes, rk_server* out) {
  if (!cfg || !out || cfg->slots == 0 || cfg->slots > 64 || cfg->leaf_max < 1 || cfg->leaf_max > 16 ||
      cfg->max_faces == 0 || cfg->max_faces >= (1u << 27) || cfg->max_pixels == 0 || arena_bytes == 0 ||
      arena_bytes > cfg->hbm_budget)
    return RK_ERROR_INVALID_VALUE;
  rk_server_st* s = new (std::nothrow) rk_server_st;
  if (!s) return RK_ERROR_OUT_OF_MEMORY;
  s->cfg = *cfg;
  auto fail = [&](rk_status e) { rk_server_destroy(s); return e; };
  if (hipMalloc(reinterpret_cast<void**>(&s->arena), arena_bytes) != hipSuccess) {
    s->arena = nullptr;
    return fail(RK_ERROR_OUT_OF_MEMORY);
  }
  s->arena_bytes = arena_bytes;
  s->free_list[0] = arena_bytes;
  s->slots.resize(cfg->slots);
  const size_t ws = rk_fast_build_workspace_size(cfg->max_faces, RK_MORTON_30);
  for (auto& sl : s->slots) {
    if (hipStreamCreate(&sl.st) != hipSuccess) return fail(RK_ERROR_DEVICE);
    if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) return fail(RK_ERROR_DEVICE);
    sl.ws_bytes = ws;
    if (hipMalloc(&sl.ws, ws) != hipSuccess) return fail(RK_ERROR_OUT_OF_MEMORY);
    if (hipMalloc(reinterpret_cast<void**>(&sl.fid), 4ull * cfg->max_pixels) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&sl.uv), 8ull * cfg->max_pixels) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&sl.dep), 4ull * cfg->max_pixels) != hipSuccess)
      return fail(RK_ERROR_OUT_OF_MEMORY);
    const size_t parts_bytes = 8 * (kSumParts + 1);
    if (hipMalloc(reinterpret_cast<void**>(&sl.sum_parts), parts_bytes) != hipSuccess)
      return fail(RK_ERROR_OUT_OF_MEMORY);
    if (hipMemset(sl.sum_parts, 0, parts_bytes) != hipSuccess) return fail(RK_ERROR_DEVICE);
    if (hipHostMalloc(reinterpret_cast<void**>(&sl.sum_host), 32, hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess)
      return fail(RK_ERROR_OUT_OF_MEMORY);
    std::memset(sl.sum_host, 0, 32);
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.sum_host_dev), sl.sum_host, 0) != hipSuccess)
      return fail(RK_ERROR_DEVICE);
    if (cfg->plain && (hipMalloc(reinterpret_cast<void**>(&sl.rays_o), 16ull * cfg->max_pixels) != hipSuccess ||
                       hipMalloc(reinterpret_cast<void**>(&sl.rays_d), 16ull * cfg->max_pixels) != hipSuccess ||
                       hipMalloc(reinterpret_cast<void**>(&sl.sum_dev), 32) != hipSuccess))
      return fail(RK_ERROR_OUT_OF_MEMORY);
  }
  *out = s;
  return RK_SUCCESS;
}

extern "C" rk_status rk_server_destroy(rk_server s) {
  if (!s) return RK_SUCCESS;
  for (auto& sl : s->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    (void)hipFree(sl.ws);
    (void)hipFree(sl.fid);
    (void)hipFree(sl.uv);
    (void)hipFree(sl.dep);
    (void)hipFree(sl.sum_parts);
    (void)hipFree(sl.rays_o);
    (void)hipFree(sl.rays_d);
    (void)hipFree(sl.sum_dev);
    if (sl.sum_host) (void)hipHostFree(sl.sum_host);
  }
  for (auto& sc : s->scenes) {
    if (sc.h_verts) (void)hipHostFree(sc.h_verts);
    if (sc.h_faces) (void)hipHostFree(sc.h_faces);
  }
  if (s->arena) (void)hipFree(s->arena);
  delete s;
  return RK_SUCCESS;
}

extern "C" rk_status rk_server_add_scene(rk_server s, const rk_float4* vertices, uint32_t vertex_count,
                                         const rk_uint4* faces, uint32_t face_count, int deform_along_z,
                                         uint32_t* scene_id) {
  if (!s || !vertices || !faces || vertex_count == 0 || face_count == 0 || !scene_id) return RK_ERROR_INVALID_VALUE;
  if (face_count > s->cfg.max_faces) return RK_ERROR_UNSUPPORTED;
  for (uint32_t f = 0; f < face_count; ++f)
    if (faces[f].x >= vertex_count || faces[f].y >= vertex_count || faces[f].z >= vertex_count)
      return RK_ERROR_INVALID_VALUE;
  Scene sc;
  sc.nv = vertex_count;
  sc.nf = face_count;
  sc.along_z = deform_along_z;
  sc.bytes = scene_bytes(vertex_count, face_count);
  if (sc.bytes > s->arena_bytes) return RK_ERROR_OUT_OF_MEMORY;
  if (hipHostMalloc(reinterpret_cast<void**>(&sc.h_verts), 16ull * vertex_count, hipHostMallocDefault) !=
      hipSuccess)
    return RK_ERROR_OUT_OF_MEMORY;
  if (hipHostMalloc(reinterpret_cast<void**>(&sc.h_faces), 16ull * face_count, hipHostMallocDefault) != hipSuccess) {
    (void)hipHostFree(sc.h_verts);
    return RK_ERROR_OUT_OF_MEMORY;
  }
  std::memcpy(sc.h_verts, vertices, 16ull * vertex_count);
  std::memcpy(sc.h_faces, faces, 16ull * face_count);
  rk_aabox hb;
  rk_status e = rk_scene_hitbox(vertices, vertex_count, &hb);
  if (e != RK_SUCCESS) {
    (void)hipHostFree(sc.h_verts);
    (void)hipHostFree(sc.h_faces);
    return e;
  }
  const float ext = std::max(hb.max.x - hb.min.x, std::max(hb.max.y - hb.min.y, hb.max.z - hb.min.z));
  sc.amplitude = s->cfg.amplitude_frac * ext;
  const float a = sc.amplitude;
  sc.hitbox = {{hb.min.x - a, hb.min.y - a, hb.min.z - a, 0.0f}, {hb.max.x + a, hb.max.y + a, hb.max.z + a, 0.0f}};
  *scene_id = (uint32_t)s->scenes.size();
  s->scenes.push_back(sc);
  return RK_SUCCESS;
}

extern "C" size_t rk_server_scene_bytes(uint32_t vertex_count, uint32_t face_count) {
  return scene_bytes(vertex_count, face_count);
}

extern "C" double rk_server_now(rk_server s) { return s ? s->now() : 0.0; }

extern "C" rk_status rk_server_submit(rk_server s, const rk_request* r, uint64_t* ticket) {
  if (!s || !r || !ticket || r->scene >= s->scenes.size() || r->width == 0 || r->height == 0)
    return RK_ERROR_INVALID_VALUE;
  const uint64_t px = (uint64_t)r->width * r->height;
  const bool own = r->faces_id || r->coords || r->depths;
  if (own && !(r->faces_id && r->coords && r->depths)) return RK_ERROR_INVALID_VALUE;
  if (!own && px > s->cfg.max_pixels) return RK_ERROR_UNSUPPORTED;
  if (px > 0xFFFFFFFFull) return RK_ERROR_UNSUPPORTED;
  const uint64_t t = s->results.size();
  rk_result res = {};
  res.arrival = r->arrival >= 0.0 ? r->arrival : s->now();
  res.scene = r->scene;
  res.pixels = (uint32_t)px;
  s->results.push_back(res);
  s->finished.push_back(0);
  s->queue.push_back(Pending{*r, t});
  *ticket = t;
  return RK_SUCCESS;
}

extern "C" rk_status rk_server_progress(rk_server s, uint32_t* completed) {
  if (!s) return RK_ERROR_INVALID_VALUE;
  return s->progress(completed);
}

extern "C" rk_status rk_server_wait_all(rk_server s) {
  if (!s) return RK_ERROR_INVALID_VALUE;
  while (!s->idle()) {
    rk_status e = s->progress(nullptr);
    if (e != RK_SUCCESS) return e;
    if (!s->queue.empty() || !s->idle()) std::this_thread::yield();
  }
  return RK_SUCCESS;
}

extern "C" rk_status rk_server_result(rk_server s, uint64_t ticket, rk_result* out) {
  if (!s || !out || ticket >= s->results.size()) return RK_ERROR_INVALID_VALUE;
  if (!s->finished[ticket]) return RK_ERROR_UNSUPPORTED;
  *out = s->results[ticket];
  return RK_SUCCESS;
}

extern "C" rk_status rk_server_stats_get(rk_server s, uint64_t first_ticket, uint64_t count, rk_server_stats* st) {
  if (!s || !st || first_ticket + count > s->results.size()) return RK_ERROR_INVALID_VALUE;
  rk_server_stats o = {};
  std::vector<double> lat;
  lat.reserve(count);
  double first = 1e300, last = 0.0;
  for (uint64_t t = first_ticket; t < first_ticket + count; ++t) {
    if (!s->finished[t]) continue;
    const rk_result& r = s->results[t];
    lat.push_back((r.done - r.arrival) * 1e3);
    first = std::min(first, r.arrival);
    last = std::max(last, r.done);
    o.rays += r.pixels;
    o.rebuilds += r.rebuilt;
    o.refits += r.rebuilt ? 0 : 1;
    o.uploads += r.uploaded;
    o.evictions += r.evictions;
    o.failed += r.status != RK_SUCCESS ? 1 : 0;
  }
  o.completed = lat.size();
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) { return lat[std::min(lat.size() - 1, (size_t)std::floor(p * (lat.size() - 1) + 0.5))]; };
    o.p50_ms = pct(0.50);
    o.p99_ms = pct(0.99);
    o.max_ms = lat.back();
    double sum = 0.0;
    for (double v : lat) sum += v;
    o.mean_ms = sum / lat.size();
    o.first_arrival = first;
    o.last_done = last;
    o.mrays_per_s = last > first ? (double)o.rays / (last - first) / 1e6 : 0.0;
  }
  o.peak_resident_bytes = s->peak_resident;
  *st = o;
  return RK_SUCCESS;
}

// Open-loop driver: Poisson arrivals (exponential gaps, SplitMix64) at 'rate'
// per second, or all at once when rate is 0. Scene uniform over the added
// scenes, time = arrival, square frames with log-uniform side, camera at 30
// degrees elevation on an orbit around the scene's hitbox (60 degree fov).
extern "C" rk_status rk_server_run_poisson(rk_server s, double rate, uint32_t count, uint32_t side_lo,
                                           uint32_t side_hi, uint64_t seed, rk_server_stats* stats,
                                           uint64_t* first_ticket) {
  if (!s || s->scenes.empty() || count == 0 || side_lo == 0 || side_hi < side_lo || rate < 0.0)
    return RK_ERROR_INVALID_VALUE;
  if ((uint64_t)side_hi * side_hi > s->cfg.max_pixels) return RK_ERROR_UNSUPPORTED;
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  auto u01 = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
  std::vector<rk_request> reqs(count);
  std::vector<double> at(count);
  double t = 0.0;
  for (uint32_t i = 0; i < count; ++i) {
    if (rate > 0.0) t += -std::log(1.0 - u01()) / rate;
    at[i] = t;
    rk_request& r = reqs[i];
    std::memset(&r, 0, sizeof r);
    r.scene = (uint32_t)(next() % s->scenes.size());
    const uint32_t side =
        std::min(side_hi, (uint32_t)std::floor(side_lo * std::pow((double)side_hi / side_lo, u01())));
    r.width = r.height = side;
    r.time = (float)t;
    const Scene& sc = s->scenes[r.scene];
    const float cx = 0.5f * (sc.hitbox.min.x + sc.hitbox.max.x), cy = 0.5f * (sc.hitbox.min.y + sc.hitbox.max.y),
                cz = 0.5f * (sc.hitbox.min.z + sc.hitbox.max.z);
    const float dx = sc.hitbox.max.x - sc.hitbox.min.x, dy = sc.hitbox.max.y - sc.hitbox.min.y,
                dz = sc.hitbox.max.z - sc.hitbox.min.z;
    const float dist = 1.2f * std::sqrt(dx * dx + dy * dy + dz * dz);
    const float az = (float)(6.283185307179586 * u01()), el = 0.5235988f;
    const rk_float4 eye = {cx + dist * std::cos(el) * std::cos(az), cy + dist * std::cos(el) * std::sin(az),
                           cz + dist * std::sin(el), 0.0f};
    rk_float4 cam[4];
    rk_status e = rk_camera_lookat_up(eye, rk_float4{cx, cy, cz, 0.0f}, rk_float4{0, 0, 1, 0}, 1.0471976f, 1.0f,
                                      (int)side, (int)side, cam);
    if (e != RK_SUCCESS) return e;
    r.camera = {cam[0], cam[1], cam[2], cam[3]};
  }
  const uint64_t first = s->results.size();
  const double start = s->now();
  uint32_t submitted = 0;
  while (submitted < count || !s->idle()) {
    const double now = s->now() - start;
    while (submitted < count && at[submitted] <= now) {
      reqs[submitted].arrival = start + at[submitted];
      uint64_t ticket;
      rk_status e = rk_server_submit(s, &reqs[submitted], &ticket);
      if (e != RK_SUCCESS) return e;
      ++submitted;
    }
    rk_status e = s->progress(nullptr);
    if (e != RK_SUCCESS) return e;
    std::this_thread::yield();
  }
  if (first_ticket) *first_ticket = first;
  if (stats) return rk_server_stats_get(s, first, count, stats);
  return RK_SUCCESS;
}
