/*
 * xdg_hip.hpp -- C++ facade over the xdg_hip C-ABI with the query semantics
 * of xdg::RayTracer (include/xdg/ray_tracing_interface.h:16-147 in the
 * reference): single-ray calls with the reference's signatures, argument
 * meanings and error behaviour, plus batched calls on device arrays.
 *
 * Header only; needs nothing but include/xdg_hip.h and libxdg_hip.so.
 *
 * Error behaviour, as the reference's Embree backend:
 *   - a query on an unknown surface tree throws std::out_of_range
 *     (std::map::at in src/embree/ray_tracer.cpp:264,297,331,352);
 *   - find_element on a tree without elements warns and returns ID_NONE
 *     (ray_tracer.cpp:236-239);
 *   - ray_fire appends the hit face to the exclusion list
 *     (ray_tracer.cpp:324);
 *   - any other failure throws xdg_hip::Error with the library's message.
 * Single-ray calls share one staging buffer and are serialised by a mutex,
 * so concurrent callers (OpenMP loops in the reference's tools) are safe.
 */
#ifndef XDG_HIP_HPP
#define XDG_HIP_HPP

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "xdg_hip.h"

namespace xdg_hip {

using MeshID = int32_t;
using TreeID = int32_t;
constexpr MeshID ID_NONE = -1;     // constants.h ID_NONE
constexpr TreeID TREE_NONE = -1;   // constants.h TREE_NONE
constexpr double INFTY = DBL_MAX;  // constants.h INFTY

// HitOrientation (include/xdg/shared_enums.h:11-15)
enum class HitOrientation : int { ANY = XDG_HIP_ANY, EXITING = XDG_HIP_EXITING, ENTERING = XDG_HIP_ENTERING };
enum class Precision : int { PLAIN = XDG_HIP_PLAIN, MIXED = XDG_HIP_MIXED };

struct Position {
  double x, y, z;
};
using Direction = Position;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class HipRayTracer {
 public:
  explicit HipRayTracer(int device = 0, Precision precision = Precision::MIXED) : precision_(precision)
  {
    const int rc = xdg_hip_create(&ctx_, device);
    if (rc != XDG_HIP_OK) throw Error(rc, "xdg_hip_create failed (no usable GPU?)");
    try {
      check(xdg_hip_malloc(ctx_, kStageBytes, &stage_));
    } catch (...) {
      xdg_hip_destroy(ctx_);
      throw;
    }
  }
  ~HipRayTracer()
  {
    if (stage_) xdg_hip_free(ctx_, stage_);
    xdg_hip_destroy(ctx_);
  }
  HipRayTracer(const HipRayTracer&) = delete;
  HipRayTracer& operator=(const HipRayTracer&) = delete;

  // RayTracer::init (ray_tracing_interface.h:22): trees are ready when built
  void init() {}
  // RTLibrary: this backend is the HIP one (the enum value a host adds)
  const char* library() const { return "HIP"; }
  xdg_hip_ctx* handle() const { return ctx_; }
  Precision precision() const { return precision_; }
  void set_precision(Precision p) { precision_ = p; }

  // ---- trees -------------------------------------------------------------
  // RayTracer::create_surface_tree (ray_tracing_interface.h:51): the
  // triangles of one volume; flip[i] = 1 where the volume is the reverse-
  // sense parent of the triangle's surface; bump = bounding_box_bump.
  TreeID create_surface_tree(const std::vector<double>& verts, const std::vector<int32_t>& tris,
                             const std::vector<int32_t>& face_ids, const std::vector<int32_t>& surf_ids,
                             const std::vector<uint8_t>& flip, double bump)
  {
    int32_t t = TREE_NONE;
    const int64_t nt = (int64_t)face_ids.size();
    if ((int64_t)tris.size() != 3 * nt || (int64_t)surf_ids.size() != nt || (int64_t)flip.size() != nt)
      throw std::invalid_argument("create_surface_tree: array sizes disagree");
    check(xdg_hip_build_surface_tree(ctx_, (int64_t)verts.size() / 3, verts.data(), nt, tris.data(),
                                     face_ids.data(), surf_ids.data(), flip.data(), bump, &t));
    return t;
  }
  // RayTracer::create_element_tree (ray_tracing_interface.h:63)
  TreeID create_element_tree(const std::vector<double>& verts, const std::vector<int32_t>& tets,
                             const std::vector<int32_t>& adj, const std::array<int32_t, 12>& face_order)
  {
    int32_t t = TREE_NONE;
    check(xdg_hip_build_element_tree(ctx_, (int64_t)verts.size() / 3, verts.data(), (int64_t)tets.size() / 4,
                                     tets.data(), adj.data(), face_order.data(), &t));
    return t;
  }
  int num_registered_surface_trees() const { return counts().first; }
  int num_registered_element_trees() const { return counts().second; }
  int num_registered_trees() const
  {
    auto c = counts();
    return c.first + c.second;
  }

  // ---- single-ray queries (RayTracer semantics) -------------------------
  // RayTracer::point_in_volume (ray_tracing_interface.h:76-79)
  bool point_in_volume(TreeID tree, const Position& point, const Direction* direction = nullptr,
                       const std::vector<MeshID>* exclude = nullptr) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* p = s.put_d3(point);
    double* d = direction ? s.put_d3(*direction) : nullptr;
    const int32_t* ex = nullptr;
    int k = 0;
    if (exclude && !exclude->empty()) {
      ex = s.put_ids(*exclude);
      k = (int)exclude->size();
    }
    uint8_t* out = (uint8_t*)s.take(8);
    tree_check(xdg_hip_point_in_volume(ctx_, tree, 1, p, d, ex, k, (int)precision_, out, nullptr), tree);
    uint8_t r = 0;
    s.get(&r, out, 1);
    return r != 0;
  }

  // RayTracer::ray_fire (ray_tracing_interface.h:81-86); the hit face is
  // appended to *exclude (src/embree/ray_tracer.cpp:324)
  std::pair<double, MeshID> ray_fire(TreeID tree, const Position& origin, const Direction& direction,
                                     double dist_limit = INFTY,
                                     HitOrientation orientation = HitOrientation::EXITING,
                                     std::vector<MeshID>* exclude = nullptr)
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* o = s.put_d3(origin);
    double* d = s.put_d3(direction);
    double* tm = s.put_d(dist_limit);
    const int32_t* ex = nullptr;
    int k = 0;
    if (exclude && !exclude->empty()) {
      ex = s.put_ids(*exclude);
      k = (int)exclude->size();
    }
    double* t = (double*)s.take(8);
    int32_t* sf = (int32_t*)s.take(8);
    int32_t* fc = (int32_t*)s.take(8);
    tree_check(xdg_hip_ray_fire(ctx_, tree, nullptr, 1, o, d, tm, (int)orientation, ex, k, (int)precision_, t, sf,
                                fc, nullptr),
               tree);
    double th = INFTY;
    int32_t sh = ID_NONE, fh = ID_NONE;
    s.get(&th, t, 8);
    s.get(&sh, sf, 4);
    s.get(&fh, fc, 4);
    if (sh == ID_NONE) return {INFTY, ID_NONE};
    if (exclude) exclude->push_back(fh);
    return {th, sh};
  }

  // RayTracer::find_element(TreeID, const Position&) (ray_tracing_interface.h:107)
  MeshID find_element(TreeID tree, const Position& point) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* p = s.put_d3(point);
    int32_t* e = (int32_t*)s.take(8);
    const int rc = xdg_hip_find_element(ctx_, tree, 1, p, e, nullptr);
    if (rc == XDG_HIP_ERR_NOT_FOUND) {
      std::fprintf(stderr, "WARNING: Tree %d does not have a point location tree\n", (int)tree);
      return ID_NONE;
    }
    check(rc);
    int32_t r = ID_NONE;
    s.get(&r, e, 4);
    return r;
  }

  // RayTracer::closest (ray_tracing_interface.h:109-110): {distance, face}
  std::pair<double, MeshID> closest(TreeID tree, const Position& point)
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* p = s.put_d3(point);
    double* dist = (double*)s.take(8);
    int32_t* face = (int32_t*)s.take(8);
    tree_check(xdg_hip_closest(ctx_, tree, 1, p, dist, face, nullptr), tree);
    double dh = INFTY;
    int32_t fh = ID_NONE;
    s.get(&dh, dist, 8);
    s.get(&fh, face, 4);
    if (fh == ID_NONE) return {INFTY, ID_NONE};
    return {dh, fh};
  }

  // RayTracer::occluded (ray_tracing_interface.h:112-115): dist is
  // -INFTY on a hit, INFTY on a miss
  bool occluded(TreeID tree, const Position& origin, const Direction& direction, double& dist) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* o = s.put_d3(origin);
    double* d = s.put_d3(direction);
    uint8_t* hit = (uint8_t*)s.take(8);
    double* dd = (double*)s.take(8);
    tree_check(xdg_hip_occluded(ctx_, tree, 1, o, d, (int)precision_, hit, dd, nullptr), tree);
    s.get(&dist, dd, 8);
    return dist != INFTY;
  }

  // XDG::find_volume (src/xdg.cpp:113-128): first volume (in the given
  // order, skipping ipc) whose point_in_volume is true, else ipc
  MeshID find_volume(const std::vector<TreeID>& vol_trees, const std::vector<MeshID>& vol_ids, MeshID ipc,
                     const Position& point, const Direction& direction) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* p = s.put_d3(point);
    double* d = s.put_d3(direction);
    int32_t* v = (int32_t*)s.take(8);
    check(xdg_hip_find_volume(ctx_, (int)vol_trees.size(), vol_trees.data(), vol_ids.data(), ipc, 1, p, d,
                              (int)precision_, v, nullptr));
    int32_t r = ipc;
    s.get(&r, v, 4);
    return r;
  }

  // Boundary faces of an element tree for the implicit-complement re-entry
  // of segments (src/xdg.cpp:164-185): the complement's surface tree and,
  // indexed by its face ids, the element index behind each face (-1 none).
  void attach_boundary(TreeID element_tree, TreeID surface_tree, const std::vector<int32_t>& bface_elem)
  {
    tree_check(xdg_hip_attach_boundary(ctx_, element_tree, surface_tree, (int64_t)bface_elem.size(),
                                       bface_elem.data()),
               element_tree);
  }

  // MeshManager::next_element (mesh_manager_interface.cpp:150-200) on
  // element indices of the tree: {next element or ID_NONE, exit distance}.
  std::pair<MeshID, double> next_element(TreeID element_tree, MeshID element, const Position& r,
                                         const Direction& u) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    int32_t* e = (int32_t*)s.take(8);
    check(xdg_hip_memcpy_htod(ctx_, e, &element, 4));
    double* rp = s.put_d3(r);
    double* up = s.put_d3(u);
    int32_t* nx = (int32_t*)s.take(8);
    double* d = (double*)s.take(8);
    tree_check(xdg_hip_next_element(ctx_, element_tree, 1, e, rp, up, nx, d, nullptr), element_tree);
    int32_t nh = ID_NONE;
    double dh = INFTY;
    s.get(&nh, nx, 4);
    s.get(&dh, d, 8);
    return {nh, dh};
  }

  // XDG::segments(start, end) (src/xdg.cpp:142-198) over an element tree:
  // the (element index, length) pieces in walk order.  At most `cap` pieces
  // are returned (the walk's own count is in *count when given).
  std::vector<std::pair<MeshID, double>> segments(TreeID element_tree, const Position& start, const Position& end,
                                                  int cap = 2048, int32_t* count = nullptr) const
  {
    std::lock_guard<std::mutex> lk(mu_);
    Stage s(*this);
    double* p0 = s.put_d3(start);
    double* p1 = s.put_d3(end);
    int32_t* cnt = (int32_t*)s.take(8);
    uint64_t* hsh = (uint64_t*)s.take(8);
    const int64_t off[2] = {0, cap};
    int64_t* offd = (int64_t*)s.take(sizeof off);
    check(xdg_hip_memcpy_htod(ctx_, offd, off, sizeof off));
    int32_t* se = (int32_t*)s.take((size_t)cap * 4);
    double* sl = (double*)s.take((size_t)cap * 8);
    tree_check(xdg_hip_tally_segments(ctx_, element_tree, 1, p0, p1, nullptr, cnt, hsh, offd, se, sl, nullptr),
               element_tree);
    int32_t n = 0;
    s.get(&n, cnt, 4);
    if (count) *count = n;
    const int m = std::min<int>(n, cap);
    std::vector<int32_t> eh((size_t)m);
    std::vector<double> lh((size_t)m);
    if (m > 0) {
      s.get(eh.data(), se, (size_t)m * 4);
      s.get(lh.data(), sl, (size_t)m * 8);
    }
    std::vector<std::pair<MeshID, double>> out;
    out.reserve((size_t)m);
    for (int i = 0; i < m; ++i) out.emplace_back(eh[(size_t)i], lh[(size_t)i]);
    return out;
  }

  // ---- batched queries on device arrays (the performance path) ----------
  void ray_fire(TreeID tree, const int32_t* tree_per_ray, int64_t n, const double* org, const double* dir,
                const double* tmax, HitOrientation orientation, const int32_t* excl, int k, double* t_out,
                int32_t* surf_out, int32_t* face_out, void* stream = nullptr)
  {
    tree_check(xdg_hip_ray_fire(ctx_, tree, tree_per_ray, n, org, dir, tmax, (int)orientation, excl, k,
                                (int)precision_, t_out, surf_out, face_out, stream),
               tree);
  }
  void synchronize(void* stream = nullptr) const { check(xdg_hip_synchronize(ctx_, stream)); }

 private:
  static constexpr size_t kStageBytes = 1 << 16;

  // bump allocator over the staging buffer for one call
  struct Stage {
    const HipRayTracer& rt;
    size_t off = 0;
    explicit Stage(const HipRayTracer& r) : rt(r) {}
    void* take(size_t bytes)
    {
      bytes = (bytes + 15) & ~(size_t)15;
      if (off + bytes > kStageBytes) throw std::length_error("xdg_hip: single-ray arguments exceed staging buffer");
      void* p = (char*)rt.stage_ + off;
      off += bytes;
      return p;
    }
    double* put_d3(const Position& v)
    {
      const double h[3] = {v.x, v.y, v.z};
      double* p = (double*)take(sizeof h);
      rt.check(xdg_hip_memcpy_htod(rt.ctx_, p, h, sizeof h));
      return p;
    }
    double* put_d(double v)
    {
      double* p = (double*)take(8);
      rt.check(xdg_hip_memcpy_htod(rt.ctx_, p, &v, 8));
      return p;
    }
    const int32_t* put_ids(const std::vector<MeshID>& ids)
    {
      if (ids.size() > XDG_HIP_MAX_EXCLUDE) throw std::length_error("xdg_hip: exclusion list longer than 1024");
      int32_t* p = (int32_t*)take(ids.size() * 4);
      rt.check(xdg_hip_memcpy_htod(rt.ctx_, p, ids.data(), ids.size() * 4));
      return p;
    }
    // the query ran on the context's (non-blocking) stream: wait for it
    // before the copy back
    void get(void* dst, const void* src, size_t bytes)
    {
      if (!synced) {
        rt.check(xdg_hip_synchronize(rt.ctx_, nullptr));
        synced = true;
      }
      rt.check(xdg_hip_memcpy_dtoh(rt.ctx_, dst, src, bytes));
    }
    bool synced = false;
  };

  void check(int rc) const
  {
    if (rc != XDG_HIP_OK) throw Error(rc, xdg_hip_last_error(ctx_));
  }
  void tree_check(int rc, TreeID tree) const
  {
    if (rc == XDG_HIP_ERR_NOT_FOUND) throw std::out_of_range("xdg_hip: unknown tree " + std::to_string(tree));
    check(rc);
  }
  std::pair<int, int> counts() const
  {
    int32_t s = 0, e = 0;
    check(xdg_hip_num_trees(ctx_, &s, &e));
    return {s, e};
  }

  xdg_hip_ctx* ctx_ = nullptr;
  void* stage_ = nullptr;
  Precision precision_;
  mutable std::mutex mu_;
};

}  // namespace xdg_hip

#endif /* XDG_HIP_HPP */
