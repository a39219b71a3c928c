/*
 * xdg_oracle.c -- CPU FP64 oracle for the XDG geometry-query hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product never calls into this file.  Build with -O2 -ffp-contract=off and
 * no -march / -ffast-math: every expression below keeps the reference's
 * operation order so that results are bit-identical to it.
 *
 * Vector helpers follow include/xdg/vec3da.h:
 *   dot(a,b)   = a.x*b.x + a.y*b.y + a.z*b.z        (vec3da.h:80-83)
 *   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *                                                    (vec3da.h:85-90)
 *   length     = sqrt(x*x + y*y + z*z)               (vec3da.h:58-59)
 */
#include "xdg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define INFTY DBL_MAX
/* constants.h:18-23 (20*DBL_EPSILON), constants.h:30 (TINY_BIT) */
#define PLUCKER_ZERO_TOL (20.0 * DBL_EPSILON)
#define TINY_BIT 1e-10
/* constants.h:80: float min_rcp_input = FLT_MIN * 1E5 (double product,
 * stored as float, compared after promotion to double) */
static const float MIN_RCP_INPUT = FLT_MIN * 1E5;

#define MAX_WALK_STEPS (1 << 16)
#define MAX_SEG_LOOPS 64

/* planted mistake for tests (xo_set_fault); XO_FAULT_NONE in normal use */
static int g_fault = XO_FAULT_NONE;

void xo_set_fault(int code) { g_fault = code; }

/* ------------------------------------------------------------------------ */
/* vector helpers                                                            */
/* ------------------------------------------------------------------------ */

static inline void v_sub(const double* a, const double* b, double* r)
{
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}

static inline double v_dot(const double* a, const double* b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline void v_cross(const double* a, const double* b, double* r)
{
  double x = a[1] * b[2] - a[2] * b[1];
  double y = a[2] * b[0] - a[0] * b[2];
  double z = a[0] * b[1] - a[1] * b[0];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}

/* Vec3da::normalize (vec3da.h:61-66) */
static inline void v_normalize(double* a)
{
  double len = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  len = len < (double)MIN_RCP_INPUT ? (double)MIN_RCP_INPUT : len;
  a[0] /= len;
  a[1] /= len;
  a[2] /= len;
}

/* ------------------------------------------------------------------------ */
/* Pluecker ray-triangle test (include/xdg/geometry/plucker.h)               */
/* ------------------------------------------------------------------------ */

/* first() plucker.h:35-43: lexicographic (x, y, z) order */
static inline int first(const double* a, const double* b)
{
  if (a[0] < b[0]) return 1;
  if (a[0] > b[0]) return 0;
  if (a[1] < b[1]) return 1;
  if (a[1] > b[1]) return 0;
  return a[2] < b[2];
}

/* plucker_edge_test plucker.h:45-62 */
double xo_plucker_edge_test(const double va[3], const double vb[3],
                            const double ray[3], const double ray_normal[3])
{
  double pip;
  double edge[3], edge_normal[3];
  if (first(va, vb)) {
    v_sub(vb, va, edge);
    v_cross(edge, va, edge_normal);
    pip = v_dot(ray, edge_normal) + v_dot(ray_normal, edge);
  } else {
    v_sub(va, vb, edge);
    v_cross(edge, vb, edge_normal);
    pip = v_dot(ray, edge_normal) + v_dot(ray_normal, edge);
    pip = -pip;
  }
  if (PLUCKER_ZERO_TOL > fabs(pip)) pip = 0.0;
  return pip;
}

/* plucker_ray_tri_intersect plucker.h:64-161 */
int xo_plucker_ray_tri_intersect(const double v[9], const double o[3],
                                 const double d[3], double tmax, double tmin,
                                 int use_orientation, int orientation,
                                 double* t)
{
  const double* v0 = v;
  const double* v1 = v + 3;
  const double* v2 = v + 6;
  double rayb[3];
  v_cross(d, o, rayb); /* plucker.h:75 */
  *t = 0.0;
  double ori = (double)orientation;

  double c0 = xo_plucker_edge_test(v0, v1, d, rayb);
  if (use_orientation && ori * c0 > 0) return 0;

  double c1 = xo_plucker_edge_test(v1, v2, d, rayb);
  if (use_orientation) {
    if (ori * c1 > 0) return 0;
  } else if ((0.0 < c0 && 0.0 > c1) || (0.0 > c0 && 0.0 < c1)) {
    return 0;
  }

  double c2 = xo_plucker_edge_test(v2, v0, d, rayb);
  if (use_orientation) {
    if (ori * c2 > 0) return 0;
  } else if ((0.0 < c1 && 0.0 > c2) || (0.0 > c1 && 0.0 < c2) ||
             (0.0 < c0 && 0.0 > c2) || (0.0 > c0 && 0.0 < c2)) {
    return 0;
  }

  if (0.0 == c0 && 0.0 == c1 && 0.0 == c2) return 0; /* coplanar */

  const double inverse_sum = 1.0 / (c0 + c1 + c2);

  /* largest-|dir| axis, first wins ties (plucker.h:137-144) */
  int idx = 0;
  double max_abs_dir = 0;
  for (int i = 0; i < 3; ++i) {
    if (fabs(d[i]) > max_abs_dir) {
      idx = i;
      max_abs_dir = fabs(d[i]);
    }
  }
  /* intersection = (c0*inv)*v2 + (c1*inv)*v1... in the reference order
   * (plucker.h:132-134): c0 pairs with v2, c1 with v0, c2 with v1 */
  double w0 = c0 * inverse_sum;
  double w1 = c1 * inverse_sum;
  double w2 = c2 * inverse_sum;
  double ix = w0 * v2[idx] + w1 * v0[idx] + w2 * v1[idx];
  double dist = (ix - o[idx]) / d[idx];

  /* barycentric guard plucker.h:149-155 */
  double bu = w2;
  double bv = w0;
  if (bu < 0.0 || bv < 0.0 || (bu + bv) > 1.0) dist = -1.0;

  if (dist < tmin || dist > tmax) return 0;
  *t = dist;
  return 1;
}

/* MeshManager::face_normal mesh_manager_interface.cpp:216-220 */
void xo_face_normal(const double v[9], double n[3])
{
  double a[3], b[3];
  v_sub(v + 3, v, a);
  v_sub(v + 6, v, b);
  v_cross(a, b, n);
  v_normalize(n);
}

/* ------------------------------------------------------------------------ */
/* closest point on triangle (src/geometry/closest.cpp:6-183)                */
/* ------------------------------------------------------------------------ */

static int region_of(double s, double t, double det)
{
  if (s + t < det) {
    if (s < 0) return t < 0 ? 4 : 3;
    if (t < 0) return 5;
    return 0;
  }
  if (s < 0) return 2;
  if (t < 0) return 6;
  return 1;
}

typedef struct {
  double sv[3], tv[3], pv[3];
  double ss, st, tt, sp, tp, det, s, t;
} closest_terms;

static void closest_setup(const double v[9], const double p[3],
                          closest_terms* c)
{
  v_sub(v + 3, v, c->sv);
  v_sub(v + 6, v, c->tv);
  v_sub(v, p, c->pv);
  c->ss = v_dot(c->sv, c->sv);
  c->st = v_dot(c->sv, c->tv);
  c->tt = v_dot(c->tv, c->tv);
  c->sp = v_dot(c->sv, c->pv);
  c->tp = v_dot(c->tv, c->pv);
  c->det = c->ss * c->tt - c->st * c->st;
  c->s = c->st * c->tp - c->tt * c->sp;
  c->t = c->st * c->sp - c->ss * c->tp;
}

int xo_determine_region(const double v[9], const double p[3])
{
  closest_terms c;
  closest_setup(v, p, &c);
  return region_of(c.s, c.t, c.det);
}

/* out = a + f*b  (Vec3da a + double*Vec3da) */
static inline void axpy(const double* a, double f, const double* b, double* o)
{
  o[0] = a[0] + f * b[0];
  o[1] = a[1] + f * b[1];
  o[2] = a[2] + f * b[2];
}
/* out = a - f*b */
static inline void axmy(const double* a, double f, const double* b, double* o)
{
  o[0] = a[0] - f * b[0];
  o[1] = a[1] - f * b[1];
  o[2] = a[2] - f * b[2];
}
/* out = f*a + g*b */
static inline void lin2(double f, const double* a, double g, const double* b,
                        double* o)
{
  o[0] = f * a[0] + g * b[0];
  o[1] = f * a[1] + g * b[1];
  o[2] = f * a[2] + g * b[2];
}
static inline void v_copy(const double* a, double* o)
{
  o[0] = a[0];
  o[1] = a[1];
  o[2] = a[2];
}

void xo_closest_location_on_triangle(const double v[9], const double p[3],
                                     double out[3])
{
  const double* v0 = v;
  const double* v1 = v + 3;
  const double* v2 = v + 6;
  closest_terms c;
  closest_setup(v, p, &c);
  double s = c.s, t = c.t;
  const double ss = c.ss, st = c.st, tt = c.tt, sp = c.sp, tp = c.tp;
  switch (region_of(s, t, c.det)) {
  case 0: {
    const double inv_det = 1.0 / c.det;
    s *= inv_det;
    t *= inv_det;
    /* vertices[0] + s*sv + t*tv: (v0 + s*sv) + t*tv */
    double a[3];
    axpy(v0, s, c.sv, a);
    axpy(a, t, c.tv, out);
    return;
  }
  case 1: {
    const double num = tt + tp - st - sp;
    if (num <= 0) {
      v_copy(v2, out);
    } else {
      const double den = ss - 2 * st + tt;
      if (num >= den) {
        v_copy(v1, out);
      } else {
        const double pp = num / den;
        lin2(pp, v1, (1 - pp), v2, out);
      }
    }
    return;
  }
  case 2: {
    s = st + sp;
    t = tt + tp;
    if (t > s) {
      const double num = t - s;
      const double den = ss - 2 * st + tt;
      if (num > den) {
        v_copy(v1, out);
      } else {
        const double pp = num / den;
        lin2(pp, v1, (1 - pp), v2, out);
      }
    } else if (t <= 0) {
      v_copy(v2, out);
    } else if (tp >= 0) {
      v_copy(v0, out);
    } else {
      axmy(v0, (tp / tt), c.tv, out);
    }
    return;
  }
  case 3:
    if (tp >= 0) {
      v_copy(v0, out);
    } else if (-tp >= tt) {
      v_copy(v2, out);
    } else {
      axmy(v0, (tp / tt), c.tv, out);
    }
    return;
  case 4:
    if (sp < 0) {
      if (-sp > ss)
        v_copy(v1, out);
      else
        axmy(v0, (sp / ss), c.sv, out);
    } else if (tp < 0) {
      if (-tp > tt)
        v_copy(v2, out);
      else
        axmy(v0, (tp / tt), c.tv, out);
    } else {
      v_copy(v0, out);
    }
    return;
  case 5:
    if (sp >= 0.0)
      v_copy(v0, out);
    else if (-sp >= ss)
      v_copy(v1, out);
    else
      axmy(v0, (sp / ss), c.sv, out);
    return;
  default: /* 6 */
    t = st + tp;
    s = ss + sp;
    if (s > t) {
      const double num = t - s;
      const double den = tt - 2 * st + ss;
      if (num > den) {
        v_copy(v2, out);
      } else {
        t = num / den;
        s = 1 - t;
        lin2(s, v1, t, v2, out);
      }
    } else if (s <= 0) {
      v_copy(v1, out);
    } else if (sp >= 0) {
      v_copy(v0, out);
    } else {
      axmy(v0, (sp / ss), c.sv, out);
    }
    return;
  }
}

/* ------------------------------------------------------------------------ */
/* tetrahedron containment (src/tetrahedron_contain.cpp:11-43)               */
/* ------------------------------------------------------------------------ */

/* T = [e0 e1 e2] as columns, lambda123 = inverse(T) * (p - v0).  The inverse
 * is the adjugate divided element-wise by the determinant (cofactor
 * expansion along the first column); lambda = sum_j col_j(inv) * rhs_j. */
int xo_tet_contains(const double p[3], const double v[12], double* lam_out)
{
  double e0[3], e1[3], e2[3], rhs[3];
  v_sub(v + 3, v, e0);
  v_sub(v + 6, v, e1);
  v_sub(v + 9, v, e2);
  v_sub(p, v, rhs);
  /* column-major matrix: c[col][row] */
  const double* c0 = e0;
  const double* c1 = e1;
  const double* c2 = e2;
  /* adjugate (columns) */
  double a0[3], a1[3], a2[3];
  a0[0] = c1[1] * c2[2] - c2[1] * c1[2];
  a0[1] = c2[1] * c0[2] - c0[1] * c2[2];
  a0[2] = c0[1] * c1[2] - c1[1] * c0[2];
  a1[0] = c1[2] * c2[0] - c2[2] * c1[0];
  a1[1] = c2[2] * c0[0] - c0[2] * c2[0];
  a1[2] = c0[2] * c1[0] - c1[2] * c0[0];
  a2[0] = c1[0] * c2[1] - c2[0] * c1[1];
  a2[1] = c2[0] * c0[1] - c0[0] * c2[1];
  a2[2] = c0[0] * c1[1] - c1[0] * c0[1];
  double det = c0[0] * a0[0] + c0[1] * a1[0] + c0[2] * a2[0];
  double i0[3], i1[3], i2[3];
  for (int k = 0; k < 3; ++k) {
    i0[k] = a0[k] / det;
    i1[k] = a1[k] / det;
    i2[k] = a2[k] / det;
  }
  double l[3];
  for (int k = 0; k < 3; ++k)
    l[k] = i0[k] * rhs[0] + i1[k] * rhs[1] + i2[k] * rhs[2];
  if (g_fault == XO_FAULT_TET_TRANSPOSE) {
    const double* ic[3] = {i0, i1, i2};
    for (int k = 0; k < 3; ++k) l[k] = ic[k][0] * rhs[0] + ic[k][1] * rhs[1] + ic[k][2] * rhs[2];
  }
  /* tetrahedron_contain.cpp:32: 1.0f - (l1 + l2 + l3) */
  double l0 = 1.0 - (l[0] + l[1] + l[2]);
  double b[4] = {l0, l[0], l[1], l[2]};
  if (lam_out) memcpy(lam_out, b, sizeof b);
  const double hi = 1.0 + PLUCKER_ZERO_TOL;
  for (int i = 0; i < 4; ++i) {
    if (b[i] < -PLUCKER_ZERO_TOL || b[i] > hi) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* host BVH (acceleration only: boxes are the reference's dilated boxes and  */
/* every query's answer equals the brute-force answer)                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  double lo[3], hi[3];
  int32_t left, right; /* leaf: left = -(first+1), right = count */
} xo_node;

typedef struct {
  xo_node* nodes;
  int64_t nnodes, cap;
  int32_t* order; /* prim indices in leaf order */
  int depth;
} xo_bvh;

typedef struct {
  const double* blo; /* prim boxes lo (3n) */
  const double* bhi;
  double* cen; /* centroids (3n) */
  xo_bvh* bvh;
} bvh_build_ctx;

static int64_t bvh_push(xo_bvh* b)
{
  if (b->nnodes == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 1024;
    b->nodes = (xo_node*)realloc(b->nodes, (size_t)b->cap * sizeof(xo_node));
  }
  return b->nnodes++;
}

static int64_t bvh_build_rec(bvh_build_ctx* c, int64_t start, int64_t end,
                             int depth)
{
  xo_bvh* b = c->bvh;
  int64_t id = bvh_push(b);
  if (depth > b->depth) b->depth = depth;
  double lo[3] = {INFTY, INFTY, INFTY}, hi[3] = {-INFTY, -INFTY, -INFTY};
  double clo[3] = {INFTY, INFTY, INFTY}, chi[3] = {-INFTY, -INFTY, -INFTY};
  for (int64_t i = start; i < end; ++i) {
    int32_t p = b->order[i];
    for (int k = 0; k < 3; ++k) {
      if (c->blo[3 * p + k] < lo[k]) lo[k] = c->blo[3 * p + k];
      if (c->bhi[3 * p + k] > hi[k]) hi[k] = c->bhi[3 * p + k];
      if (c->cen[3 * p + k] < clo[k]) clo[k] = c->cen[3 * p + k];
      if (c->cen[3 * p + k] > chi[k]) chi[k] = c->cen[3 * p + k];
    }
  }
  xo_node node;
  memcpy(node.lo, lo, sizeof lo);
  memcpy(node.hi, hi, sizeof hi);
  int64_t count = end - start;
  if (count <= 4) {
    node.left = -(int32_t)(start + 1);
    node.right = (int32_t)count;
    b->nodes[id] = node;
    return id;
  }
  int axis = 0;
  double ext = chi[0] - clo[0];
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > ext) {
      ext = chi[k] - clo[k];
      axis = k;
    }
  double mid = 0.5 * (clo[axis] + chi[axis]);
  int64_t i = start, j = end - 1;
  while (i <= j) {
    if (c->cen[3 * b->order[i] + axis] < mid) {
      ++i;
    } else {
      int32_t tmp = b->order[i];
      b->order[i] = b->order[j];
      b->order[j] = tmp;
      --j;
    }
  }
  int64_t split = i;
  if (split == start || split == end) split = start + count / 2;
  int64_t l = bvh_build_rec(c, start, split, depth + 1);
  int64_t r = bvh_build_rec(c, split, end, depth + 1);
  node.left = (int32_t)l;
  node.right = (int32_t)r;
  b->nodes[id] = node;
  return id;
}

static xo_bvh* bvh_build(int64_t n, const double* blo, const double* bhi,
                         double* cen)
{
  xo_bvh* b = (xo_bvh*)calloc(1, sizeof(xo_bvh));
  b->order = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) b->order[i] = (int32_t)i;
  if (n > 0) {
    bvh_build_ctx c = {blo, bhi, cen, b};
    bvh_build_rec(&c, 0, n, 0);
  }
  return b;
}

static void bvh_destroy(xo_bvh* b)
{
  if (!b) return;
  free(b->nodes);
  free(b->order);
  free(b);
}

/* FP64 slab test; NaN-safe (fmin/fmax drop NaN).  Conservative given the
 * >= 1e-3 box dilation. */
static inline int ray_box(const xo_node* nd, const double* o, const double* inv,
                          double tmax)
{
  double tn = 0.0, tf = tmax;
  for (int k = 0; k < 3; ++k) {
    double t1 = (nd->lo[k] - o[k]) * inv[k];
    double t2 = (nd->hi[k] - o[k]) * inv[k];
    double a = fmin(t1, t2), bb = fmax(t1, t2);
    if (a > tn) tn = a;
    if (bb < tf) tf = bb;
  }
  return tn <= tf;
}

static inline double point_box_d2(const xo_node* nd, const double* p)
{
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double e = 0.0;
    if (p[k] < nd->lo[k])
      e = nd->lo[k] - p[k];
    else if (p[k] > nd->hi[k])
      e = p[k] - nd->hi[k];
    d2 += e * e;
  }
  return d2;
}

static inline int point_in_box(const xo_node* nd, const double* p)
{
  return p[0] >= nd->lo[0] && p[0] <= nd->hi[0] && p[1] >= nd->lo[1] &&
         p[1] <= nd->hi[1] && p[2] >= nd->lo[2] && p[2] <= nd->hi[2];
}

/* ------------------------------------------------------------------------ */
/* surface trees                                                             */
/* ------------------------------------------------------------------------ */

struct xo_tree {
  int64_t n;
  double* v;      /* 9n */
  int32_t* face;  /* n */
  int32_t* surf;  /* n */
  uint8_t* flip;  /* n */
  double bump;
  xo_bvh* bvh;
};

xo_tree* xo_tree_create(int64_t n, const double* verts9, const int32_t* face_id,
                        const int32_t* surf_id, const uint8_t* flip,
                        double bump, int bvh)
{
  xo_tree* t = (xo_tree*)calloc(1, sizeof(xo_tree));
  t->n = n;
  size_t nn = (size_t)(n > 0 ? n : 1);
  t->v = (double*)malloc(nn * 9 * sizeof(double));
  t->face = (int32_t*)malloc(nn * sizeof(int32_t));
  t->surf = (int32_t*)malloc(nn * sizeof(int32_t));
  t->flip = (uint8_t*)malloc(nn);
  if (n > 0) {
    memcpy(t->v, verts9, (size_t)n * 9 * sizeof(double));
    memcpy(t->face, face_id, (size_t)n * sizeof(int32_t));
    memcpy(t->surf, surf_id, (size_t)n * sizeof(int32_t));
    memcpy(t->flip, flip, (size_t)n);
  }
  t->bump = bump;
  if (bvh && n > 0) {
    /* TriangleBoundsFunc (triangle_intersect.cpp:36-50): tri box +- bump */
    double* blo = (double*)malloc(nn * 3 * sizeof(double));
    double* bhi = (double*)malloc(nn * 3 * sizeof(double));
    double* cen = (double*)malloc(nn * 3 * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
      const double* v = t->v + 9 * i;
      for (int k = 0; k < 3; ++k) {
        double lo = v[k], hi = v[k];
        for (int j = 1; j < 3; ++j) {
          if (v[3 * j + k] < lo) lo = v[3 * j + k];
          if (v[3 * j + k] > hi) hi = v[3 * j + k];
        }
        blo[3 * i + k] = lo - bump;
        bhi[3 * i + k] = hi + bump;
        cen[3 * i + k] = 0.5 * (lo + hi);
      }
    }
    t->bvh = bvh_build(n, blo, bhi, cen);
    free(blo);
    free(bhi);
    free(cen);
  }
  return t;
}

void xo_tree_destroy(xo_tree* t)
{
  if (!t) return;
  free(t->v);
  free(t->face);
  free(t->surf);
  free(t->flip);
  bvh_destroy(t->bvh);
  free(t);
}

int64_t xo_tree_size(const xo_tree* t) { return t ? t->n : 0; }

typedef struct {
  double t;
  int32_t face, surf;
  double n[3];
  int64_t prim; /* tree position of the committed triangle */
} hit_state;

/* Leaf test for ray_fire / point_in_volume: TriangleIntersectionFunc
 * (triangle_intersect.cpp:52-109) with the smallest-face-id tie rule. */
static inline void leaf_intersect(const xo_tree* tr, int64_t p,
                                  const double* o, const double* d,
                                  int orientation, const int32_t* ex, int k,
                                  hit_state* h)
{
  const double* v = tr->v + 9 * p;
  double t;
  if (!xo_plucker_ray_tri_intersect(v, o, d, h->t, 0.0, 0, 0, &t)) return;
  if (t > h->t) return;
  int32_t face = tr->face[p];
  if (t == h->t && face >= h->face) return; /* tie rule */
  double n[3];
  xo_face_normal(v, n);
  if (tr->flip[p]) {
    n[0] = -n[0];
    n[1] = -n[1];
    n[2] = -n[2];
  }
  /* orientation_cull (triangle_intersect.cpp:12-24) */
  if (orientation != XO_ANY) {
    double dp = v_dot(d, n);
    if (orientation == XO_EXITING && dp < 0.0) return;
    if (orientation == XO_ENTERING && dp >= 0.0) return;
  }
  /* primitive_mask_cull (triangle_intersect.cpp:26-34) */
  for (int j = 0; j < k; ++j)
    if (ex[j] == face) return;
  h->t = t;
  h->face = face;
  h->surf = tr->surf[p];
  v_copy(n, h->n);
  h->prim = p;
}

static void trace_one(const xo_tree* tr, const double* o, const double* d,
                      double tmax, int orientation, const int32_t* ex, int k,
                      int brute, hit_state* h)
{
  h->t = tmax;
  h->face = INT32_MAX;
  h->surf = XO_ID_NONE;
  h->n[0] = h->n[1] = h->n[2] = 0.0;
  h->prim = -1;
  if (brute || !tr->bvh) {
    for (int64_t p = 0; p < tr->n; ++p)
      leaf_intersect(tr, p, o, d, orientation, ex, k, h);
    return;
  }
  double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
  int32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  const xo_bvh* b = tr->bvh;
  while (sp) {
    const xo_node* nd = &b->nodes[stack[--sp]];
    if (!ray_box(nd, o, inv, h->t)) continue;
    if (nd->left < 0) {
      int64_t first = -(int64_t)nd->left - 1;
      for (int64_t i = 0; i < nd->right; ++i)
        leaf_intersect(tr, b->order[first + i], o, d, orientation, ex, k, h);
    } else {
      stack[sp++] = nd->right;
      stack[sp++] = nd->left;
    }
  }
}

#define OMP_FOR _Pragma("omp parallel for schedule(dynamic, 64) num_threads(nt)")

static int resolve_threads(int nthreads)
{
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void xo_ray_fire(const xo_tree* tr, int64_t n, const double* org,
                 const double* dir, const double* tmax, int orientation,
                 const int32_t* excl, int k, int brute, int nthreads,
                 double* t_out, int32_t* surf_out, int32_t* face_out)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    hit_state h;
    trace_one(tr, org + 3 * i, dir + 3 * i, tmax ? tmax[i] : INFTY,
              orientation, excl ? excl + (int64_t)k * i : NULL,
              excl ? k : 0, brute, &h);
    if (h.surf == XO_ID_NONE) {
      t_out[i] = INFTY;
      surf_out[i] = XO_ID_NONE;
      face_out[i] = XO_ID_NONE;
    } else {
      t_out[i] = h.t;
      surf_out[i] = h.surf;
      face_out[i] = h.face;
    }
  }
}

void xo_point_in_volume(const xo_tree* tr, int64_t n, const double* pts,
                        const double* dir, const int32_t* excl, int k,
                        int brute, int nthreads, uint8_t* inside)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  /* ray_tracer.cpp:268: {1. / std::sqrt(2.0), 1 / std::sqrt(2.0), 0.0} */
  const double dflt[3] = {1. / sqrt(2.0), 1 / sqrt(2.0), 0.0};
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    const double* d = dir ? dir + 3 * i : dflt;
    hit_state h;
    trace_one(tr, pts + 3 * i, d, INFTY, XO_ANY,
              excl ? excl + (int64_t)k * i : NULL, excl ? k : 0, brute, &h);
    if (h.surf == XO_ID_NONE)
      inside[i] = 0;
    else
      inside[i] = v_dot(d, h.n) > 0.0; /* ray_tracer.cpp:286 */
  }
}

static int occluded_one(const xo_tree* tr, const double* o, const double* d,
                        int brute)
{
  double t;
  if (brute || !tr->bvh) {
    for (int64_t p = 0; p < tr->n; ++p)
      if (xo_plucker_ray_tri_intersect(tr->v + 9 * p, o, d, INFTY, 0.0, 0, 0,
                                       &t))
        return 1;
    return 0;
  }
  double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
  int32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  const xo_bvh* b = tr->bvh;
  while (sp) {
    const xo_node* nd = &b->nodes[stack[--sp]];
    if (!ray_box(nd, o, inv, INFTY)) continue;
    if (nd->left < 0) {
      int64_t first = -(int64_t)nd->left - 1;
      for (int64_t i = 0; i < nd->right; ++i)
        if (xo_plucker_ray_tri_intersect(tr->v + 9 * b->order[first + i], o, d,
                                         INFTY, 0.0, 0, 0, &t))
          return 1;
    } else {
      stack[sp++] = nd->right;
      stack[sp++] = nd->left;
    }
  }
  return 0;
}

void xo_occluded(const xo_tree* tr, int64_t n, const double* org,
                 const double* dir, int brute, int nthreads, uint8_t* hit,
                 double* dist_out)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    int h = occluded_one(tr, org + 3 * i, dir + 3 * i, brute);
    hit[i] = (uint8_t)h;
    if (dist_out) dist_out[i] = h ? -INFTY : INFTY;
  }
}

static inline void closest_leaf(const xo_tree* tr, int64_t p, const double* q,
                                double* best, int32_t* bface)
{
  double c[3], e[3];
  xo_closest_location_on_triangle(tr->v + 9 * p, q, c);
  v_sub(c, q, e);
  double dist = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  int32_t face = tr->face[p];
  if (dist < *best || (dist == *best && face < *bface)) {
    *best = dist;
    *bface = face;
  }
}

void xo_closest(const xo_tree* tr, int64_t n, const double* pts, int brute,
                int nthreads, double* dist_out, int32_t* face_out)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    const double* q = pts + 3 * i;
    double best = INFTY;
    int32_t bface = INT32_MAX;
    if (brute || !tr->bvh) {
      for (int64_t p = 0; p < tr->n; ++p) closest_leaf(tr, p, q, &best, &bface);
    } else {
      int32_t stack[256];
      int sp = 0;
      stack[sp++] = 0;
      const xo_bvh* b = tr->bvh;
      while (sp) {
        const xo_node* nd = &b->nodes[stack[--sp]];
        double d2 = point_box_d2(nd, q);
        if (best < INFTY && d2 > best * best) continue;
        if (nd->left < 0) {
          int64_t first = -(int64_t)nd->left - 1;
          for (int64_t j = 0; j < nd->right; ++j)
            closest_leaf(tr, b->order[first + j], q, &best, &bface);
        } else {
          stack[sp++] = nd->right;
          stack[sp++] = nd->left;
        }
      }
    }
    if (bface == INT32_MAX) {
      dist_out[i] = INFTY;
      face_out[i] = XO_ID_NONE;
    } else {
      dist_out[i] = best;
      face_out[i] = bface;
    }
  }
}

void xo_find_volume(const xo_tree* const* trees, const int32_t* vol_ids,
                    int nvol, int32_t ipc, int64_t n, const double* pts,
                    const double* dirs, int brute, int nthreads,
                    int32_t* vol_out)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    int32_t found = ipc;
    for (int v = 0; v < nvol; ++v) {
      if (vol_ids[v] == ipc) continue;
      hit_state h;
      const double* d = dirs + 3 * i;
      trace_one(trees[v], pts + 3 * i, d, INFTY, XO_ANY, NULL, 0, brute, &h);
      if (h.surf != XO_ID_NONE && v_dot(d, h.n) > 0.0) {
        found = vol_ids[v];
        break;
      }
    }
    vol_out[i] = found;
  }
}

/* ------------------------------------------------------------------------ */
/* tetrahedral meshes                                                        */
/* ------------------------------------------------------------------------ */

struct xo_tetmesh {
  int64_t nv, ne;
  double* verts;
  int32_t* tets;
  int32_t* adj;
  int32_t fo[12];
  xo_bvh* bvh;
};

static void tet_coords(const xo_tetmesh* m, int32_t e, double* v12)
{
  for (int j = 0; j < 4; ++j) {
    int32_t vi = m->tets[4 * (int64_t)e + j];
    v_copy(m->verts + 3 * (int64_t)vi, v12 + 3 * j);
  }
}

xo_tetmesh* xo_tetmesh_create(int64_t nv, const double* verts, int64_t ne,
                              const int32_t* tets, const int32_t* adj,
                              const int32_t face_order[12], int bvh)
{
  xo_tetmesh* m = (xo_tetmesh*)calloc(1, sizeof(xo_tetmesh));
  m->nv = nv;
  m->ne = ne;
  m->verts = (double*)malloc((size_t)(nv > 0 ? nv : 1) * 3 * sizeof(double));
  m->tets = (int32_t*)malloc((size_t)(ne > 0 ? ne : 1) * 4 * sizeof(int32_t));
  m->adj = (int32_t*)malloc((size_t)(ne > 0 ? ne : 1) * 4 * sizeof(int32_t));
  memcpy(m->verts, verts, (size_t)nv * 3 * sizeof(double));
  memcpy(m->tets, tets, (size_t)ne * 4 * sizeof(int32_t));
  memcpy(m->adj, adj, (size_t)ne * 4 * sizeof(int32_t));
  memcpy(m->fo, face_order, sizeof m->fo);
  if (bvh && ne > 0) {
    size_t nn = (size_t)ne;
    double* blo = (double*)malloc(nn * 3 * sizeof(double));
    double* bhi = (double*)malloc(nn * 3 * sizeof(double));
    double* cen = (double*)malloc(nn * 3 * sizeof(double));
    for (int64_t e = 0; e < ne; ++e) {
      double v[12];
      tet_coords(m, (int32_t)e, v);
      /* BoundingBox::from_points + dilation (bbox.h:85-102):
       * bump = sqrt(w.dot(w)) * 1e-6, no floor (tetrahedron_contain.cpp:47-63) */
      double lo[3] = {INFTY, INFTY, INFTY}, hi[3] = {-INFTY, -INFTY, -INFTY};
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
          lo[k] = fmin(lo[k], v[3 * j + k]);
          hi[k] = fmax(hi[k], v[3 * j + k]);
        }
      double w[3];
      v_sub(hi, lo, w);
      double bump = g_fault == XO_FAULT_NO_DILATION ? 0.0 : sqrt(v_dot(w, w)) * 1e-6;
      for (int k = 0; k < 3; ++k) {
        blo[3 * e + k] = lo[k] - bump;
        bhi[3 * e + k] = hi[k] + bump;
        cen[3 * e + k] = 0.5 * (lo[k] + hi[k]);
      }
    }
    m->bvh = bvh_build(ne, blo, bhi, cen);
    free(blo);
    free(bhi);
    free(cen);
  }
  return m;
}

void xo_tetmesh_destroy(xo_tetmesh* m)
{
  if (!m) return;
  free(m->verts);
  free(m->tets);
  free(m->adj);
  bvh_destroy(m->bvh);
  free(m);
}

static int32_t find_element_one(const xo_tetmesh* m, const double* p,
                                int brute)
{
  int32_t best = INT32_MAX;
  double v[12];
  if (brute || !m->bvh) {
    for (int64_t e = 0; e < m->ne; ++e) {
      tet_coords(m, (int32_t)e, v);
      if (xo_tet_contains(p, v, NULL)) return (int32_t)e;
    }
    return XO_ID_NONE;
  }
  int32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  const xo_bvh* b = m->bvh;
  while (sp) {
    const xo_node* nd = &b->nodes[stack[--sp]];
    if (!point_in_box(nd, p)) continue;
    if (nd->left < 0) {
      int64_t first = -(int64_t)nd->left - 1;
      for (int64_t i = 0; i < nd->right; ++i) {
        int32_t e = b->order[first + i];
        if (e >= best) continue;
        tet_coords(m, e, v);
        if (xo_tet_contains(p, v, NULL)) {
          if (g_fault == XO_FAULT_FIRST_VISITED) return e;
          best = e;
        }
      }
    } else {
      stack[sp++] = nd->right;
      stack[sp++] = nd->left;
    }
  }
  return best == INT32_MAX ? XO_ID_NONE : best;
}

void xo_find_element(const xo_tetmesh* m, int64_t n, const double* pts,
                     int brute, int nthreads, int32_t* elem_out)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i)
    elem_out[i] = find_element_one(m, pts + 3 * i, brute);
}

void xo_next_element(const xo_tetmesh* m, int32_t elem, const double r[3],
                     const double u[3], int32_t* next, double* dist)
{
  double v[12];
  tet_coords(m, elem, v);
  double dists[4];
  for (int i = 0; i < 4; ++i) {
    double tri[9];
    for (int j = 0; j < 3; ++j) v_copy(v + 3 * m->fo[3 * i + j], tri + 3 * j);
    double t;
    int hit = xo_plucker_ray_tri_intersect(tri, r, u, INFTY, 0.0, 1, 1, &t);
    /* std::max(0.0, t) == (0.0 < t) ? t : 0.0 */
    dists[i] = hit ? ((0.0 < t) ? t : 0.0) : INFTY;
  }
  int idx_out = XO_ID_NONE;
  double min_dist = INFTY;
  for (int i = 0; i < 4; ++i) {
    if (dists[i] < min_dist) {
      min_dist = dists[i];
      idx_out = i;
    }
  }
  *next = idx_out == XO_ID_NONE ? XO_ID_NONE
                                : m->adj[4 * (int64_t)elem + idx_out];
  *dist = min_dist;
}

int64_t xo_walk_elements(const xo_tetmesh* m, int32_t elem,
                         const double r_in[3], const double u[3],
                         double distance, int64_t cap, int32_t* elems,
                         double* lens)
{
  double r[3];
  v_copy(r_in, r);
  int64_t cnt = 0;
  while (distance > 0) {
    int32_t next;
    double ex;
    xo_next_element(m, elem, r, u, &next, &ex);
    /* std::min(exit, distance) == (distance < exit) ? distance : exit */
    double s = (distance < ex) ? distance : ex;
    distance -= s;
    if (cnt < cap) {
      elems[cnt] = elem;
      lens[cnt] = s;
    }
    ++cnt;
    r[0] = r[0] + s * u[0];
    r[1] = r[1] + s * u[1];
    r[2] = r[2] + s * u[2];
    elem = next;
    if (elem == XO_ID_NONE) break;
    if (cnt >= MAX_WALK_STEPS) break;
  }
  return cnt;
}

uint64_t xo_mix64(uint64_t x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

static inline uint64_t dbits(double d)
{
  uint64_t b;
  memcpy(&b, &d, sizeof b);
  return b;
}

typedef struct {
  int32_t count;
  uint64_t hash;
  int64_t* tally;
  int32_t* se;
  double* sl;
  int64_t cap;
} seg_sink;

static inline void sink_push(seg_sink* s, int32_t e, double len)
{
  if (s->count < s->cap) {
    s->se[s->count] = e;
    s->sl[s->count] = len;
  }
  s->count++;
  s->hash = xo_mix64(s->hash ^ (uint64_t)(uint32_t)e);
  s->hash = xo_mix64(s->hash ^ dbits(len));
  if (s->tally) {
    long long q = llrint(ldexp(len, 40));
#pragma omp atomic
    s->tally[e] += q;
  }
}

/* walk_elements fused with the sink; returns the accumulate() sum of the
 * produced lengths (xdg.cpp:190-191: ((0.0 + s0) + s1) + ...). */
static double walk_sink(const xo_tetmesh* m, int32_t elem, const double* r_in,
                        const double* u, double distance, seg_sink* s,
                        int32_t* last_elem, int* steps)
{
  double r[3];
  v_copy(r_in, r);
  double sum = 0.0;
  while (distance > 0) {
    int32_t next;
    double ex;
    xo_next_element(m, elem, r, u, &next, &ex);
    double len = (distance < ex) ? distance : ex;
    distance -= len;
    sink_push(s, elem, len);
    *last_elem = elem;
    sum = sum + len;
    r[0] = r[0] + len * u[0];
    r[1] = r[1] + len * u[1];
    r[2] = r[2] + len * u[2];
    elem = next;
    if (elem == XO_ID_NONE) break;
    if (++*steps >= MAX_WALK_STEPS) break;
  }
  return sum;
}

void xo_segments(const xo_tetmesh* m, const xo_tree* ipc,
                 const int32_t* bface_elem, int64_t n, const double* p0,
                 const double* p1, int nthreads, int32_t* count,
                 uint64_t* hash, int64_t* tally_fx, const int64_t* seq_off,
                 int32_t* seq_elem, double* seq_len)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    seg_sink s;
    s.count = 0;
    s.hash = XO_HASH_SEED;
    s.tally = tally_fx;
    if (seq_off) {
      s.se = seq_elem + seq_off[i];
      s.sl = seq_len + seq_off[i];
      s.cap = seq_off[i + 1] - seq_off[i];
    } else {
      s.se = NULL;
      s.sl = NULL;
      s.cap = 0;
    }
    /* xdg.cpp:150-153 */
    double r[3], u[3];
    v_copy(p0 + 3 * i, r);
    v_sub(p1 + 3 * i, p0 + 3 * i, u);
    double distance = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] = u[0] / distance;
    u[1] = u[1] / distance;
    u[2] = u[2] / distance;
    int32_t last = XO_ID_NONE;
    int steps = 0, loops = 0;
    while (distance > 0) {
      if (++loops > MAX_SEG_LOOPS || steps >= MAX_WALK_STEPS) break;
      int32_t cur = find_element_one(m, r, 0);
      if (s.count > 0 && cur == last) cur = XO_ID_NONE;
      if (cur == XO_ID_NONE) {
        /* xdg.cpp:164-185: ray_fire against the implicit complement */
        if (!ipc) break;
        double ro[3];
        ro[0] = r[0] + TINY_BIT * u[0];
        ro[1] = r[1] + TINY_BIT * u[1];
        ro[2] = r[2] + TINY_BIT * u[2];
        hit_state h;
        trace_one(ipc, ro, u, INFTY, XO_EXITING, NULL, 0, 0, &h);
        if (h.surf == XO_ID_NONE || h.t > distance) break;
        double hit_dist = h.t + TINY_BIT;
        r[0] = r[0] + hit_dist * u[0];
        r[1] = r[1] + hit_dist * u[1];
        r[2] = r[2] + hit_dist * u[2];
        distance -= hit_dist;
        int32_t adjacent = bface_elem ? bface_elem[h.face] : XO_ID_NONE;
        if (adjacent == XO_ID_NONE) break;
        cur = adjacent;
      }
      double sum = walk_sink(m, cur, r, u, distance, &s, &last, &steps);
      r[0] = r[0] + sum * u[0];
      r[1] = r[1] + sum * u[1];
      r[2] = r[2] + sum * u[2];
      distance -= sum;
    }
    count[i] = s.count;
    hash[i] = s.hash;
  }
}

/* XDG::segments(volume, start, end) (src/xdg.cpp:200-227).  vol: the
 * volume's surface tree (senses as seen from the volume).  A start outside
 * the element tree enters through the first ENTERING hit (xdg.cpp:211-222);
 * the walk is MeshManager::walk_elements(elem, start, end)
 * (mesh_manager_interface.cpp:139-147): direction and length are taken
 * again from the moved start, with Vec3da::normalize.  flags (nullable):
 * bit 1 = surface hit with no element behind it (the reference warns and
 * returns no segments). */
void xo_segments_volume(const xo_tetmesh* m, const xo_tree* vol, int64_t n,
                        const double* p0, const double* p1, int nthreads,
                        int32_t* count, uint64_t* hash, int64_t* tally_fx,
                        const int64_t* seq_off, int32_t* seq_elem,
                        double* seq_len, uint32_t* flags)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    seg_sink s;
    s.count = 0;
    s.hash = XO_HASH_SEED;
    s.tally = tally_fx;
    if (seq_off) {
      s.se = seq_elem + seq_off[i];
      s.sl = seq_len + seq_off[i];
      s.cap = seq_off[i + 1] - seq_off[i];
    } else {
      s.se = NULL;
      s.sl = NULL;
      s.cap = 0;
    }
    uint32_t f = 0;
    /* xdg.cpp:205-208 */
    const double* a = p0 + 3 * i;
    double start[3], u[3];
    v_copy(a, start);
    v_sub(p1 + 3 * i, a, u);
    v_normalize(u);
    int32_t e = find_element_one(m, a, 0);
    int ok = 1;
    if (e == XO_ID_NONE) {
      /* xdg.cpp:211-222 */
      ok = vol != NULL;
      if (ok) {
        hit_state h;
        trace_one(vol, a, u, INFTY, XO_ENTERING, NULL, 0, 0, &h);
        ok = h.surf != XO_ID_NONE;
        if (ok) {
          const double into = h.t + TINY_BIT;
          double q[3] = {a[0] + u[0] * into, a[1] + u[1] * into,
                         a[2] + u[2] * into};
          e = find_element_one(m, q, 0);
          if (e == XO_ID_NONE) {
            ok = 0;
            f |= 2u;
          } else {
            start[0] = start[0] + u[0] * h.t;
            start[1] = start[1] + u[1] * h.t;
            start[2] = start[2] + u[2] * h.t;
          }
        }
      }
    }
    if (ok) {
      /* mesh_manager_interface.cpp:139-147 */
      double w[3];
      v_sub(p1 + 3 * i, start, w);
      const double distance = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      v_normalize(w);
      int32_t last = XO_ID_NONE;
      int steps = 0;
      walk_sink(m, e, start, w, distance, &s, &last, &steps);
      if (steps >= MAX_WALK_STEPS) f |= 1u;
    }
    count[i] = s.count;
    hash[i] = s.hash;
    if (flags) flags[i] = f;
  }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and uniforms            */
/* ------------------------------------------------------------------------ */

void xo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void xo_uniform2(uint32_t seed, uint32_t stream, uint64_t item, uint32_t draw,
                 double* ua, double* ub)
{
  uint32_t ctr[4] = {(uint32_t)item, (uint32_t)(item >> 32), draw, 0u};
  uint32_t key[2] = {seed, stream};
  uint32_t o[4];
  xo_philox4x32_10(ctr, key, o);
  uint64_t a = ((uint64_t)o[0] << 32) | o[1];
  uint64_t b = ((uint64_t)o[2] << 32) | o[3];
  *ua = (double)(a >> 11) * 0x1.0p-53;
  *ub = (double)(b >> 11) * 0x1.0p-53;
}

/* ln(x) for positive normal x: x = m * 2^e with m in [sqrt(1/2), sqrt(2));
 * ln(m) = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716; atanh series to s^23;
 * ln2 split into a 32-bit head (exact e*ln2_hi) and a tail. */
#define LN2_HI 0x1.62e42fefa0000p-1
#define LN2_LO 0x1.cf79abc9e3b3ap-40
double xo_log(double x)
{
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint64_t b = dbits(x);
  int e = (int)((b >> 52) & 0x7ff);
  if (e == 0) { /* subnormal: scale up */
    x = x * 0x1.0p54;
    b = dbits(x);
    e = (int)((b >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &b, sizeof m);
  if (m > 0x1.6a09e667f3bcdp+0) { /* sqrt(2) */
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  /* P(z) = sum_{k=0..11} z^k / (2k+1), Horner, no FMA */
  double p = 1.0 / 23.0;
  p = 1.0 / 21.0 + z * p;
  p = 1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = 1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = 1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = 1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = 1.0 / 3.0 + z * p;
  /* ln(m) = 2s + 2s*z*p  (split keeps the leading term exact-ish) */
  double two_s = s + s;
  double lnm = two_s + two_s * (z * p);
  double de = (double)e;
  return de * LN2_HI + (de * LN2_LO + lnm);
}

/* ------------------------------------------------------------------------ */
/* batched particle random walk                                             */
/* ------------------------------------------------------------------------ */

/* Marsaglia (1972) isotropic direction, as tools/ray_benchmark.h:24-40: one
 * Philox block per attempt, the draw counter advances once per attempt. */
static int iso_dir_draw(uint32_t seed, uint32_t stream, uint64_t item,
                        uint32_t* dc, double* u)
{
  for (int a = 0; a < 1000; ++a) {
    double ua, ub;
    xo_uniform2(seed, stream, item, (*dc)++, &ua, &ub);
    double x1 = ua * 2.0 - 1.0;
    double x2 = ub * 2.0 - 1.0;
    double s = x1 * x1 + x2 * x2;
    if (s > 0.0 && s < 1.0) {
      double t = 2.0 * sqrt(1.0 - s);
      u[0] = x1 * t;
      u[1] = x2 * t;
      u[2] = 1.0 - 2.0 * s;
      return 1;
    }
  }
  return 0;
}

static int slot_of(const xo_transport_problem* p, int32_t vol)
{
  for (int v = 0; v < p->nvol; ++v)
    if (p->vol_ids[v] == vol) return v;
  return -1;
}

static inline void hist_push(int32_t* hist, int* n, int32_t face)
{
  if (*n == XO_TRANSPORT_HISTORY) {
    for (int j = 1; j < XO_TRANSPORT_HISTORY; ++j) hist[j - 1] = hist[j];
    *n -= 1;
  }
  hist[(*n)++] = face;
}

void xo_transport(const xo_transport_problem* p, uint64_t first, int64_t n,
                  int nthreads, int64_t* tally_fx, int32_t* events,
                  uint8_t* status, uint64_t* hash)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t item = first + (uint64_t)i;
    /* start: uniform in the box from blocks 0 and 1 (same protocol as the
     * host box_points generator), then the direction from block 2 on */
    uint32_t dc = 0;
    double ua, ub, uc, unused;
    double r[3], u[3];
    xo_uniform2(p->seed, p->stream, item, dc++, &ua, &ub);
    xo_uniform2(p->seed, p->stream, item, dc++, &uc, &unused);
    r[0] = p->lo[0] + (p->hi[0] - p->lo[0]) * ua;
    r[1] = p->lo[1] + (p->hi[1] - p->lo[1]) * ub;
    r[2] = p->lo[2] + (p->hi[2] - p->lo[2]) * uc;
    uint8_t st = XO_TR_ALIVE;
    int32_t ev = 0;
    uint64_t hh = XO_HASH_SEED;
    if (!iso_dir_draw(p->seed, p->stream, item, &dc, u)) st = XO_TR_ERROR;
    /* XDG::find_volume (xdg.cpp:113-128) with the particle's direction */
    int32_t vol = p->ipc;
    if (st == XO_TR_ALIVE) {
      for (int v = 0; v < p->nvol; ++v) {
        if (p->vol_ids[v] == p->ipc) continue;
        hit_state h;
        trace_one(p->trees[v], r, u, INFTY, XO_ANY, NULL, 0, 0, &h);
        if (h.surf != XO_ID_NONE && v_dot(u, h.n) > 0.0) {
          vol = p->vol_ids[v];
          break;
        }
      }
      if (vol == p->ipc) st = XO_TR_LEAKED;
    }
    int32_t hist[XO_TRANSPORT_HISTORY];
    int hn = 0;
    while (st == XO_TR_ALIVE) {
      const int slot = slot_of(p, vol);
      if (slot < 0) {
        st = XO_TR_ERROR;
        break;
      }
      /* Particle::surf_dist (particle_sim.h:41-54) */
      hit_state h;
      trace_one(p->trees[slot], r, u, INFTY, XO_EXITING, hist, hn, 0, &h);
      if (h.surf == XO_ID_NONE) {
        st = XO_TR_LOST;
        break;
      }
      if (h.t == 0.0) {
        st = XO_TR_STUCK;
        break;
      }
      hist_push(hist, &hn, h.face); /* ray_tracer.cpp:324 appends the hit */
      /* sample_collision_distance (particle_sim.h:56-58) */
      xo_uniform2(p->seed, p->stream, item, dc++, &ua, &ub);
      const double dcol = -xo_log(1.0 - ua) * p->mfp;
      /* advance (particle_sim.h:70-83): r += s * u, tally the flight */
      const double s = dcol < h.t ? dcol : h.t;
      r[0] = r[0] + s * u[0];
      r[1] = r[1] + s * u[1];
      r[2] = r[2] + s * u[2];
      if (tally_fx) {
#pragma omp atomic
        tally_fx[slot] += (int64_t)llrint(ldexp(s, 32));
      }
      hh = xo_mix64(hh ^ (uint64_t)(uint32_t)vol);
      hh = xo_mix64(hh ^ dbits(s));
      ev += 1;
      if (dcol < h.t) {
        /* collide (particle_sim.h:60-68), absorption first */
        if (ub < p->p_absorb) {
          st = XO_TR_ABSORBED;
          break;
        }
        if (!iso_dir_draw(p->seed, p->stream, item, &dc, u)) {
          st = XO_TR_ERROR;
          break;
        }
        hn = 0;
      } else {
        /* cross_surface (particle_sim.h:85-130) */
        const int32_t sid = h.surf;
        const int bc = (p->surf_bc && sid >= 0 && sid < p->nsurf_ids)
                           ? p->surf_bc[sid]
                           : XO_BC_TRANSMISSION;
        if (bc == XO_BC_VACUUM) {
          st = XO_TR_VACUUM;
          break;
        }
        if (bc == XO_BC_REFLECTING) {
          /* particle_sim.h:90-104 with XDG::surface_normal (xdg.cpp:272-290)
           * = face_normal of the last history entry, the face just hit */
          double nrm[3];
          xo_face_normal(p->trees[slot]->v + 9 * h.prim, nrm);
          const double proj = nrm[0] * u[0] + nrm[1] * u[1] + nrm[2] * u[2];
          const double mag =
              sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
          const double sc = 2.0 * proj / mag;
          u[0] = u[0] - sc * nrm[0];
          u[1] = u[1] - sc * nrm[1];
          u[2] = u[2] - sc * nrm[2];
          v_normalize(u);
          hist[0] = h.face; /* history_ = {history_.back()} */
          hn = 1;
          if (ev >= p->max_events) {
            st = XO_TR_CAPPED;
            break;
          }
          continue;
        }
        /* transmission: MeshManager::next_volume
         * (mesh_manager_interface.cpp:202-214) */
        int32_t nv = XO_ID_NONE;
        if (sid >= 0 && sid < p->nsurf_ids) {
          if (p->sense_fwd[sid] == vol)
            nv = p->sense_rev[sid];
          else if (p->sense_rev[sid] == vol)
            nv = p->sense_fwd[sid];
        }
        if (nv == XO_ID_NONE) {
          st = XO_TR_ERROR;
          break;
        }
        vol = nv;
        if (vol == p->ipc) {
          st = XO_TR_LEAKED;
          break;
        }
      }
      if (ev >= p->max_events) {
        st = XO_TR_CAPPED;
        break;
      }
    }
    if (events) events[i] = ev;
    if (status) status[i] = st;
    if (hash) hash[i] = hh;
  }
}

/* ------------------------------------------------------------------------ */
/* adjoint of the ray distance                                               */
/* ------------------------------------------------------------------------ */

void xo_adjoint(int64_t nv, const double* verts, int64_t ntri,
                const int32_t* tris, int64_t n, const double* org,
                const double* dir, const double* t, const int32_t* face,
                const double* dldt, double* dldv, double* dldo, double* dldd)
{
  (void)nv;
  for (int64_t i = 0; i < n; ++i) {
    double* go = dldo ? dldo + 3 * i : NULL;
    double* gd = dldd ? dldd + 3 * i : NULL;
    const int32_t f = face[i];
    const double ti = t[i];
    int ok = f >= 0 && f < ntri && ti < INFTY;
    double e1[3], e2[3], nn[3], q[3], a1[3], a2[3], den = 0.0;
    const double *v0 = NULL, *v1 = NULL, *v2 = NULL;
    if (ok) {
      v0 = verts + 3 * (int64_t)tris[3 * (int64_t)f];
      v1 = verts + 3 * (int64_t)tris[3 * (int64_t)f + 1];
      v2 = verts + 3 * (int64_t)tris[3 * (int64_t)f + 2];
      v_sub(v1, v0, e1);
      v_sub(v2, v0, e2);
      v_cross(e1, e2, nn);
      den = v_dot(nn, dir + 3 * i);
      ok = den != 0.0;
    }
    if (!ok) {
      if (go) go[0] = go[1] = go[2] = 0.0;
      if (gd) gd[0] = gd[1] = gd[2] = 0.0;
      continue;
    }
    const double w = dldt[i] / den;
    const double* o = org + 3 * i;
    const double* d = dir + 3 * i;
    /* q = v0 - (o + t d), in the triangle's plane */
    for (int k = 0; k < 3; ++k) q[k] = v0[k] - (o[k] + ti * d[k]);
    v_cross(e2, q, a1); /* dt/dv1 * den */
    v_cross(q, e1, a2); /* dt/dv2 * den */
    if (dldv) {
      const int64_t i0 = tris[3 * (int64_t)f], i1 = tris[3 * (int64_t)f + 1],
                    i2 = tris[3 * (int64_t)f + 2];
      for (int k = 0; k < 3; ++k) {
        dldv[3 * i1 + k] += w * a1[k];
        dldv[3 * i2 + k] += w * a2[k];
        dldv[3 * i0 + k] += w * (nn[k] - a1[k] - a2[k]);
      }
    }
    for (int k = 0; k < 3; ++k) {
      if (go) go[k] = -w * nn[k];
      if (gd) gd[k] = -w * ti * nn[k];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* element-by-element particle walk (tools/walk_elements.h:53-117)          */
/* ------------------------------------------------------------------------ */

static inline void prim_push(int32_t* prims, int* n, int32_t face)
{
  if (*n == XO_TRANSPORT_HISTORY) {
    for (int j = 1; j < XO_TRANSPORT_HISTORY; ++j) prims[j - 1] = prims[j];
    *n -= 1;
  }
  prims[(*n)++] = face;
}

void xo_mesh_walk(const xo_tetmesh* m, const xo_tree* ipc, const double lo[3],
                  const double hi[3], uint32_t seed, uint32_t stream,
                  double mfp, int32_t max_events, uint64_t first, int64_t n,
                  int nthreads, double* distance, int32_t* events,
                  uint8_t* status, uint64_t* hash, int64_t* tally_fx)
{
  int nt = resolve_threads(nthreads);
  (void)nt;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t item = first + (uint64_t)i;
    uint32_t dc = 0;
    uint8_t st = XO_TR_ALIVE;
    double r[3], u[3];
    int32_t e = XO_ID_NONE;
    /* sample a location within the model (walk_elements.h:64-68) */
    for (int a = 0; a < XO_MESH_WALK_STARTS && e == XO_ID_NONE; ++a) {
      double ua, ub, uc, unused;
      xo_uniform2(seed, stream, item, dc++, &ua, &ub);
      xo_uniform2(seed, stream, item, dc++, &uc, &unused);
      r[0] = lo[0] + (hi[0] - lo[0]) * ua;
      r[1] = lo[1] + (hi[1] - lo[1]) * ub;
      r[2] = lo[2] + (hi[2] - lo[2]) * uc;
      e = find_element_one(m, r, 0);
    }
    if (e == XO_ID_NONE) st = XO_TR_ERROR;
    /* u = rand_dir(); u.normalize() (:70-71) */
    if (st == XO_TR_ALIVE) {
      if (!iso_dir_draw(seed, stream, item, &dc, u))
        st = XO_TR_ERROR;
      else
        v_normalize(u);
    }
    double dist = 0.0;
    int32_t nev = 0;
    uint64_t hh = XO_HASH_SEED;
    int32_t prims[XO_TRANSPORT_HISTORY];
    int np = 0;
    while (st == XO_TR_ALIVE && e != XO_ID_NONE) {
      int32_t next;
      double ex;
      xo_next_element(m, e, r, u, &next, &ex);
      double ua, ub;
      xo_uniform2(seed, stream, item, dc++, &ua, &ub);
      const double coll = -xo_log(1.0 - ua) * mfp;
      double s;
      const int32_t here = e;
      if (coll < ex) {
        s = coll;
        r[0] = r[0] + u[0] * coll;
        r[1] = r[1] + u[1] * coll;
        r[2] = r[2] + u[2] * coll;
        dist += coll;
        if (!iso_dir_draw(seed, stream, item, &dc, u)) st = XO_TR_ERROR;
      } else {
        s = ex;
        r[0] = r[0] + u[0] * ex;
        r[1] = r[1] + u[1] * ex;
        r[2] = r[2] + u[2] * ex;
        dist += ex;
        e = next;
      }
      hh = xo_mix64(hh ^ (uint64_t)(uint32_t)here);
      hh = xo_mix64(hh ^ dbits(s));
      if (tally_fx) {
        long long q = llrint(ldexp(s, 40));
#pragma omp atomic
        tally_fx[here] += q;
      }
      /* re-entry through the implicit complement (:92-102) */
      int tries = 0;
      while (st == XO_TR_ALIVE && e == XO_ID_NONE && ipc) {
        if (++tries > XO_MESH_WALK_REENTRIES) {
          st = XO_TR_LOST;
          break;
        }
        hit_state h;
        trace_one(ipc, r, u, INFTY, XO_EXITING, prims, np, 0, &h);
        if (h.surf == XO_ID_NONE) break;
        prim_push(prims, &np, h.face); /* ray_tracer.cpp:324 */
        r[0] = r[0] + u[0] * h.t;
        r[1] = r[1] + u[1] * h.t;
        r[2] = r[2] + u[2] * h.t;
        dist += h.t;
        double q[3] = {r[0] + u[0] * TINY_BIT, r[1] + u[1] * TINY_BIT,
                       r[2] + u[2] * TINY_BIT};
        e = find_element_one(m, q, 0);
      }
      np = 0;
      ++nev;
      if (st == XO_TR_ALIVE && e != XO_ID_NONE && nev >= max_events)
        st = XO_TR_CAPPED;
    }
    if (st == XO_TR_ALIVE) st = XO_TR_LEAKED;
    if (distance) distance[i] = dist;
    if (events) events[i] = nev;
    if (status) status[i] = st;
    if (hash) hash[i] = hh;
  }
}
