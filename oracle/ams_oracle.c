/*
 * ams_oracle.c -- FP64 CPU restatement of AutoMS's channel model.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; the product library (automs_amd/csrc) never links or calls it.
 *
 * What it restates (reference files are cited by path:line):
 *   - closest hit over scene triangles and reception spheres
 *       optixPathTracer.cu:92-105 (rayTrace, tmin 1e-5), :419-444 (sphere IS)
 *   - forward path discovery, reflect / transmit branch tree
 *       optixPathTracer.cu:142-190 (raygen stage 1), :285-331 (closest-hit fwd)
 *   - backward image-method evaluation of (path, tile) tasks
 *       optixPathTracer.cu:191-273 (raygen stage 2), :332-417 (closest-hit bwd)
 *   - beam patterns  optixPathTracer.cu:107-136
 *   - Rx tiling      optixPathTracer.cu:83-90
 *   - neighbour expansion rounds  optixPathTracer.cpp:229-354
 *
 * Documented departures from the reference, shared with the HIP path:
 *   - rays are Fibonacci-sphere directions (per element or distributed),
 *     not TEA/LCG stratified samples; no min_hc / reserve_size gating;
 *   - a path is keyed by (hit count, reflect mask, reflection-plane ids),
 *     not by the bitwise float image transform (optixPathTracer.h:28-38);
 *   - the +-r AABB of a tiny Tx/Rx sphere is replaced by an acceptance radius
 *     max(r, 1e-4 m) on the perpendicular miss distance;
 *   - the plane of the previous interaction is excluded from the next query
 *     (a straight ray cannot meet one plane twice);
 *   - in the backward pass a reflection must happen on the path's plane;
 *   - an optional per-material reflection phase (default 0 = reference).
 *
 * Parity status of each function is listed in the repository design notes;
 * every function here is pinned by the tests in tests/test_oracle_*.py
 * (closed forms, brute force, invariants).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAX_HIT 5
#define OR_NO_PLANE 0xFFFFFu
#define OR_TMIN 1e-5
#define OR_ACC_MIN 1e-4

typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len(v3 a) { return sqrt(dot(a, a)); }
static inline v3 nrm(v3 a) { return scl(a, 1.0 / len(a)); }
static inline double vget(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* ------------------------------------------------------------------ BVH -- */
typedef struct {
  double lo[3], hi[3];
  int left, right, first, count;
} bnode;

typedef struct {
  bnode *nodes;
  int n_nodes;
  int *idx;
} bvh_t;

typedef struct {
  const double *lo, *hi; /* per primitive boxes [n][3] */
  const double *c;       /* centroids [n][3] */
  bvh_t *b;
} bvh_build_ctx;

static int bvh_build_rec(bvh_build_ctx *ctx, int first, int count) {
  bvh_t *b = ctx->b;
  int me = b->n_nodes++;
  bnode *n = &b->nodes[me];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = first; i < first + count; i++) {
    int p = b->idx[i];
    for (int k = 0; k < 3; k++) {
      if (ctx->lo[3 * p + k] < lo[k]) lo[k] = ctx->lo[3 * p + k];
      if (ctx->hi[3 * p + k] > hi[k]) hi[k] = ctx->hi[3 * p + k];
      if (ctx->c[3 * p + k] < clo[k]) clo[k] = ctx->c[3 * p + k];
      if (ctx->c[3 * p + k] > chi[k]) chi[k] = ctx->c[3 * p + k];
    }
  }
  memcpy(n->lo, lo, sizeof lo);
  memcpy(n->hi, hi, sizeof hi);
  n->left = n->right = -1;
  n->first = first;
  n->count = count;
  if (count <= 4) return me;
  int ax = 0;
  double ext = chi[0] - clo[0];
  for (int k = 1; k < 3; k++)
    if (chi[k] - clo[k] > ext) { ext = chi[k] - clo[k]; ax = k; }
  if (ext <= 0) return me;
  double mid = 0.5 * (clo[ax] + chi[ax]);
  int i = first, j = first + count - 1;
  while (i <= j) {
    if (ctx->c[3 * b->idx[i] + ax] < mid) i++;
    else { int t = b->idx[i]; b->idx[i] = b->idx[j]; b->idx[j] = t; j--; }
  }
  int nl = i - first;
  if (nl == 0 || nl == count) nl = count / 2;
  int l = bvh_build_rec(ctx, first, nl);
  int r = bvh_build_rec(ctx, first + nl, count - nl);
  n = &b->nodes[me];
  n->left = l;
  n->right = r;
  n->count = 0;
  return me;
}

static void bvh_build(bvh_t *b, int n, const double *lo, const double *hi) {
  double *c = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  for (int i = 0; i < 3 * n; i++) c[i] = 0.5 * (lo[i] + hi[i]);
  b->idx = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) b->idx[i] = i;
  b->nodes = (bnode *)malloc(sizeof(bnode) * (2 * n + 1));
  b->n_nodes = 0;
  bvh_build_ctx ctx = {lo, hi, c, b};
  if (n > 0) bvh_build_rec(&ctx, 0, n);
  free(c);
}

static void bvh_free(bvh_t *b) {
  free(b->nodes);
  free(b->idx);
  b->nodes = NULL;
  b->idx = NULL;
}

/* slab test: does the ray o + t d hit the box for some t in [tmin, tmax] */
static inline int box_hit(const bnode *n, v3 o, v3 inv, double tmin, double tmax) {
  double t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; k++) {
    double oo = vget(o, k), iv = vget(inv, k);
    double a = (n->lo[k] - oo) * iv, b = (n->hi[k] - oo) * iv;
    if (isnan(a)) a = -INFINITY; /* 0 * inf: ray parallel and on the slab face */
    if (isnan(b)) b = INFINITY;
    if (a > b) { double t = a; a = b; b = t; }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
    if (t0 > t1) return 0;
  }
  return 1;
}

/* ---------------------------------------------------------------- scene -- */
typedef struct {
  int n_tri;
  v3 *v0, *v1, *v2, *e1, *e2, *n0; /* n0 = normalize(cross(v1-v0, v2-v0)) */
  double *maxedge;
  int *mat, *plane;
  int n_mat;
  double *mat_prop; /* [n_mat][19] */
  double *refl_phase;
  v3 *pl_n;       /* canonical plane normal of each triangle's plane */
  double *pl_d;   /* plane offset: p . pl_n = pl_d */
  bvh_t bvh;
} or_scene;

typedef struct {
  int n;
  v3 *c;
  double r, acc; /* sphere radius and acceptance radius on the miss distance */
  bvh_t bvh;
} or_spheres;

/* canonical sign: first component with |x| > 1e-9 is positive */
static v3 canon(v3 n) {
  double c[3] = {n.x, n.y, n.z};
  for (int k = 0; k < 3; k++) {
    if (fabs(c[k]) > 1e-9) {
      if (c[k] < 0) return scl(n, -1.0);
      return n;
    }
  }
  return n;
}

typedef struct { long long q[4]; int tri; } plkey;
static int plkey_cmp(const void *a, const void *b) {
  const plkey *x = (const plkey *)a, *y = (const plkey *)b;
  for (int k = 0; k < 4; k++) {
    if (x->q[k] < y->q[k]) return -1;
    if (x->q[k] > y->q[k]) return 1;
  }
  return x->tri - y->tri;
}

void *or_scene_create(int n_tri, const float *tri, const int *mat, int n_mat,
                      const float *mat_prop, const float *refl_phase) {
  or_scene *s = (or_scene *)calloc(1, sizeof(or_scene));
  s->n_tri = n_tri;
  s->v0 = malloc(sizeof(v3) * (n_tri + 1));
  s->v1 = malloc(sizeof(v3) * (n_tri + 1));
  s->v2 = malloc(sizeof(v3) * (n_tri + 1));
  s->e1 = malloc(sizeof(v3) * (n_tri + 1));
  s->e2 = malloc(sizeof(v3) * (n_tri + 1));
  s->n0 = malloc(sizeof(v3) * (n_tri + 1));
  s->pl_n = malloc(sizeof(v3) * (n_tri + 1));
  s->pl_d = malloc(sizeof(double) * (n_tri + 1));
  s->maxedge = malloc(sizeof(double) * (n_tri + 1));
  s->mat = malloc(sizeof(int) * (n_tri + 1));
  s->plane = malloc(sizeof(int) * (n_tri + 1));
  s->n_mat = n_mat;
  s->mat_prop = malloc(sizeof(double) * 19 * (n_mat + 1));
  s->refl_phase = malloc(sizeof(double) * (n_mat + 1));
  for (int i = 0; i < 19 * n_mat; i++) s->mat_prop[i] = mat_prop[i];
  for (int i = 0; i < n_mat; i++) s->refl_phase[i] = refl_phase ? refl_phase[i] : 0.0;
  double *lo = malloc(sizeof(double) * 3 * (n_tri + 1));
  double *hi = malloc(sizeof(double) * 3 * (n_tri + 1));
  plkey *keys = malloc(sizeof(plkey) * (n_tri + 1));
  for (int i = 0; i < n_tri; i++) {
    const float *t = tri + 9 * i;
    s->v0[i] = mk(t[0], t[1], t[2]);
    s->v1[i] = mk(t[3], t[4], t[5]);
    s->v2[i] = mk(t[6], t[7], t[8]);
    s->e1[i] = sub(s->v1[i], s->v0[i]);
    s->e2[i] = sub(s->v2[i], s->v0[i]);
    v3 c = cross(s->e1[i], s->e2[i]);
    double cl = len(c);
    s->n0[i] = cl > 0 ? scl(c, 1.0 / cl) : mk(0, 0, 1);
    /* optixPathTracer.cu:369-370: longest edge of the hit triangle */
    double a = len(sub(s->v0[i], s->v1[i])), b = len(sub(s->v1[i], s->v2[i])),
           cc = len(sub(s->v2[i], s->v0[i]));
    s->maxedge[i] = fmax(fmax(a, b), cc);
    s->mat[i] = mat[i];
    v3 pn = canon(s->n0[i]);
    double pd = dot(pn, s->v0[i]);
    keys[i].q[0] = llround(pn.x * 1e6);
    keys[i].q[1] = llround(pn.y * 1e6);
    keys[i].q[2] = llround(pn.z * 1e6);
    keys[i].q[3] = llround(pd * 1e5);
    keys[i].tri = i;
    for (int k = 0; k < 3; k++) {
      double x0 = vget(s->v0[i], k), x1 = vget(s->v1[i], k), x2 = vget(s->v2[i], k);
      lo[3 * i + k] = fmin(x0, fmin(x1, x2));
      hi[3 * i + k] = fmax(x0, fmax(x1, x2));
    }
  }
  /* plane id = smallest triangle index among triangles of the same plane */
  qsort(keys, n_tri, sizeof(plkey), plkey_cmp);
  for (int i = 0; i < n_tri;) {
    int j = i;
    while (j < n_tri && memcmp(keys[j].q, keys[i].q, sizeof keys[i].q) == 0) j++;
    int rep = keys[i].tri; /* sorted by tri within a group: smallest first */
    for (int k = i; k < j; k++) s->plane[keys[k].tri] = rep;
    i = j;
  }
  for (int i = 0; i < n_tri; i++) {
    int p = s->plane[i];
    s->pl_n[i] = canon(s->n0[p]);
    s->pl_d[i] = dot(s->pl_n[i], s->v0[p]);
  }
  free(keys);
  bvh_build(&s->bvh, n_tri, lo, hi);
  free(lo);
  free(hi);
  return s;
}

void or_scene_destroy(void *p) {
  or_scene *s = (or_scene *)p;
  if (!s) return;
  free(s->v0); free(s->v1); free(s->v2); free(s->e1); free(s->e2); free(s->n0);
  free(s->pl_n); free(s->pl_d); free(s->maxedge); free(s->mat); free(s->plane);
  free(s->mat_prop); free(s->refl_phase);
  bvh_free(&s->bvh);
  free(s);
}

void or_scene_plane_ids(void *p, int *out) {
  or_scene *s = (or_scene *)p;
  memcpy(out, s->plane, sizeof(int) * s->n_tri);
}

static or_spheres *spheres_create(int n, const float *c, double r) {
  or_spheres *sp = (or_spheres *)calloc(1, sizeof(or_spheres));
  sp->n = n;
  sp->r = r;
  sp->acc = fmax(r, OR_ACC_MIN);
  sp->c = malloc(sizeof(v3) * (n + 1));
  double *lo = malloc(sizeof(double) * 3 * (n + 1)), *hi = malloc(sizeof(double) * 3 * (n + 1));
  for (int i = 0; i < n; i++) {
    sp->c[i] = mk(c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    for (int k = 0; k < 3; k++) {
      lo[3 * i + k] = vget(sp->c[i], k) - sp->acc;
      hi[3 * i + k] = vget(sp->c[i], k) + sp->acc;
    }
  }
  bvh_build(&sp->bvh, n, lo, hi);
  free(lo);
  free(hi);
  return sp;
}

static void spheres_destroy(or_spheres *sp) {
  if (!sp) return;
  free(sp->c);
  bvh_free(&sp->bvh);
  free(sp);
}

/* tie tolerance in metres for counting near-boundary events */
static double g_tie_eps = 1e-6;
void or_set_tie_eps(double e) { g_tie_eps = e; }

/* Moller-Trumbore in FP64 with closed edges: the plane distance t (-1 for a
 * ray parallel to the plane) and the signed margin, the smallest
 * barycentric coordinate times the longest edge (metres inside the nearest
 * edge, approximately; negative outside).  A hit needs margin >= 0. */
static inline double tri_hit(const or_scene *s, int i, v3 o, v3 d, double *margin) {
  v3 e1 = s->e1[i], e2 = s->e2[i];
  v3 pv = cross(d, e2);
  double det = dot(e1, pv);
  *margin = -INFINITY;
  if (det == 0.0) return -1.0;
  double inv = 1.0 / det;
  v3 tv = sub(o, s->v0[i]);
  double u = dot(tv, pv) * inv;
  v3 qv = cross(tv, e1);
  double v = dot(d, qv) * inv;
  double t = dot(e2, qv) * inv;
  double w = 1.0 - u - v;
  double m = fmin(u, fmin(v, w));
  *margin = m * s->maxedge[i];
  return t;
}

typedef struct {
  double t;
  int tri;
  double margin;    /* edge distance of the accepted hit */
  double near_miss; /* smallest t of a triangle missed by less than the tie tolerance */
  int edge_tie;     /* the accepted hit lies within the tie tolerance of an edge */
  int tmin_tie;     /* a triangle meets the ray within the tie tolerance of tmin */
} tri_q;

/* closest triangle with t in [tmin, tmax], skipping triangles of plane skip_pl */
static tri_q closest_tri(const or_scene *s, v3 o, v3 d, double tmin, double tmax, int skip_pl) {
  tri_q q = {INFINITY, -1, INFINITY, INFINITY, 0, 0};
  if (s->n_tri == 0) return q;
  v3 inv = mk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  int stack[128], sp = 0;
  stack[sp++] = 0;
  double best = tmax;
  while (sp) {
    const bnode *n = &s->bvh.nodes[stack[--sp]];
    if (!box_hit(n, o, inv, tmin, best)) continue;
    if (n->count > 0) {
      for (int k = n->first; k < n->first + n->count; k++) {
        int i = s->bvh.idx[k];
        if (s->plane[i] == skip_pl) continue;
        double m;
        double t = tri_hit(s, i, o, d, &m);
        /* A perturbation of the ray by g_tie_eps moves its meeting point with
         * the plane by g_tie_eps / |cos| within the plane and along the ray:
         * the tie band of a grazing leg is wider. */
        double band = g_tie_eps / fmax(fabs(dot(s->n0[i], d)), 1e-3);
        if (m > -band && fabs(t - tmin) < band) q.tmin_tie = 1;
        if (m < 0.0) {
          if (m > -band && t >= tmin && t < q.near_miss) q.near_miss = t;
          continue;
        }
        if (t >= tmin && t <= best) {
          if (t < best || (t == best && i < q.tri)) {
            best = t;
            q.t = t;
            q.tri = i;
            q.margin = m;
            q.edge_tie = m < band;
          }
        }
      }
    } else {
      stack[sp++] = n->left;
      stack[sp++] = n->right;
    }
  }
  return q;
}

typedef struct {
  double t;
  int idx;
  double margin; /* acc - perp of the accepted sphere */
  double near_t; /* smallest b of a sphere missed by less than the tie tolerance */
} sph_q;

/* closest reception sphere: optixPathTracer.cu:436-444 with the acceptance
 * radius above; reported at t = b (closest approach). */
static sph_q closest_sphere(const or_spheres *sp, v3 o, v3 d, double tmin, double tmax) {
  sph_q q = {INFINITY, -1, INFINITY, INFINITY};
  if (!sp || sp->n == 0) return q;
  v3 inv = mk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  int stack[128], top = 0;
  stack[top++] = 0;
  double best = tmax, acc2 = sp->acc * sp->acc, r2 = sp->r * sp->r;
  while (top) {
    const bnode *n = &sp->bvh.nodes[stack[--top]];
    if (!box_hit(n, o, inv, tmin, best + sp->acc)) continue;
    if (n->count > 0) {
      for (int k = n->first; k < n->first + n->count; k++) {
        int i = sp->bvh.idx[k];
        v3 O = sub(sp->c[i], o);
        double b = dot(O, d);
        double c = dot(O, O);
        double perp2 = c - b * b;
        if (perp2 >= acc2 && b >= tmin && c > r2 && sqrt(perp2) - sp->acc < g_tie_eps && b < q.near_t)
          q.near_t = b;
        if (perp2 < acc2 && b >= tmin && b <= best && c > r2) {
          if (b < best || (b == best && i < q.idx)) {
            best = b;
            q.t = b;
            q.idx = i;
            q.margin = sp->acc - sqrt(fmax(perp2, 0.0));
          }
        }
      }
    } else {
      stack[top++] = n->left;
      stack[top++] = n->right;
    }
  }
  return q;
}

/* ------------------------------------------------------- beam patterns -- */
/* optixPathTracer.cu:107-136 */
double or_beam_pattern(const double *d1, const double *d2, int type) {
  v3 a = mk(d1[0], d1[1], d1[2]), b = mk(d2[0], d2[1], d2[2]);
  if (type == 0) return 1.0;
  double la = len(a), lb = len(b);
  double c = dot(a, b) / (la * lb);
  switch (type) {
  case 1:
    return c < 0 ? 0.0 : c * 2.4495;
  case 2: {
    double th = acos(fmax(-1.0, fmin(1.0, c)));
    if (th > M_PI * 2 / 3) return 0.1;
    if (th > M_PI * 5 / 12)
      return pow(10.0, (M_PI * 2 / 3 - th) / M_PI * 4 * (log10(cos(M_PI * 5 / 12)) + 1) - 1) * 2.4495;
    return c * 2.4495;
  }
  case 3:
    return c < 0 ? 0.0 : 1.4142;
  default:
    return NAN; /* the reference leaves this undefined (no default return) */
  }
}

/* optixPathTracer.cu:304-306: nearest 5-degree bin of the incidence angle */
static inline int theta_bin(v3 d, v3 N) {
  double c = dot(d, N) / (len(N) * len(d));
  double th = acos(fmin(1.0, fabs(c)));
  int id = (int)((th / (M_PI / 2) * 36 + 1) / 2);
  if (id > 18) id = 18;
  return id;
}

/* The two bins either side of the nearest bin edge when the incidence angle
 * lies within g_tie_ang radians of that edge (an FP32 angle may round to the
 * other bin); returns 0 otherwise.  Edges are at (k - 1/2) 5 degrees. */
static double g_tie_ang = 1e-5;
void or_set_tie_ang(double a) { g_tie_ang = a; }
static inline int theta_edge(v3 d, v3 N, int *lo, int *hi) {
  double c = dot(d, N) / (len(N) * len(d));
  double th = acos(fmin(1.0, fabs(c)));
  double x = th / (M_PI / 36) + 0.5; /* bin edges at integer x */
  double k = floor(x + 0.5);
  if (fabs(x - k) * (M_PI / 36) >= g_tie_ang || k < 1 || k > 18) return 0;
  *lo = (int)k - 1;
  *hi = (int)k;
  return 1;
}

/* -------------------------------------------------------------- keys ---- */
/* key layout (two u64, compared as (hi, lo)):
 * hi = h:3 | b:5 | p0:20 | p1:20 | p2[19:4]:16
 * lo = p2[3:0]:4 | p3:20 | p4:20 | tile:20                                 */
static inline void key_pack(int h, int b, const unsigned *pl, unsigned tile, uint64_t *hi, uint64_t *lo) {
  *hi = ((uint64_t)h << 61) | ((uint64_t)b << 56) | ((uint64_t)pl[0] << 36) |
        ((uint64_t)pl[1] << 16) | ((uint64_t)pl[2] >> 4);
  *lo = ((uint64_t)(pl[2] & 0xF) << 60) | ((uint64_t)pl[3] << 40) | ((uint64_t)pl[4] << 20) |
        (uint64_t)(tile & 0xFFFFF);
}
static inline void key_unpack(uint64_t hi, uint64_t lo, int *h, int *b, unsigned *pl, unsigned *tile) {
  *h = (int)(hi >> 61);
  *b = (int)((hi >> 56) & 31);
  pl[0] = (unsigned)((hi >> 36) & 0xFFFFF);
  pl[1] = (unsigned)((hi >> 16) & 0xFFFFF);
  pl[2] = (unsigned)(((hi & 0xFFFF) << 4) | (lo >> 60));
  pl[3] = (unsigned)((lo >> 40) & 0xFFFFF);
  pl[4] = (unsigned)((lo >> 20) & 0xFFFFF);
  *tile = (unsigned)(lo & 0xFFFFF);
}

typedef struct { uint64_t hi, lo; } key128;
static int key_cmp(const void *a, const void *b) {
  const key128 *x = (const key128 *)a, *y = (const key128 *)b;
  if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
  if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
  return 0;
}

static long key_unique(key128 *k, long n) {
  if (n == 0) return 0;
  qsort(k, n, sizeof(key128), key_cmp);
  long m = 1;
  for (long i = 1; i < n; i++)
    if (k[i].hi != k[m - 1].hi || k[i].lo != k[m - 1].lo) k[m++] = k[i];
  return m;
}

/* optixPathTracer.cu:83-90 */
static inline unsigned tile_of(int j, int n1, int gs) {
  int n1g = (n1 - 1) / gs + 1;
  return (unsigned)((j / n1) / gs * n1g + (j % n1) / gs);
}

/* Fibonacci-sphere direction k of n (FP64, fractional angle range-reduced) */
void or_fib_dir(long k, long n, double *out) {
  const double g = 0.38196601125010515179541316563436; /* (3 - sqrt 5) / 2 */
  double z = 1.0 - (2.0 * (double)k + 1.0) / (double)n;
  double r = sqrt(fmax(0.0, 1.0 - z * z));
  double f = fmod((double)k * g, 1.0);
  /* k*g is exact to ~k*1e-16 cycles, below 1e-9 rad for k < 2^24 */
  double ph = 2.0 * M_PI * f;
  out[0] = r * cos(ph);
  out[1] = r * sin(ph);
  out[2] = z;
}

/* ------------------------------------------------------ forward pass ---- */
typedef struct {
  v3 o, d;
  int ttl, b, nref, last_pl;
  unsigned pl[OR_MAX_HIT];
} fbranch;

typedef struct {
  long n_rays, n_queries, n_records, n_ties;
} or_fwd_stats;

typedef struct {
  key128 *k;
  long n, cap;
} kbuf;

static void kbuf_push(kbuf *kb, uint64_t hi, uint64_t lo) {
  if (kb->n == kb->cap) {
    kb->cap = kb->cap ? 2 * kb->cap : 4096;
    kb->k = realloc(kb->k, sizeof(key128) * kb->cap);
  }
  kb->k[kb->n].hi = hi;
  kb->k[kb->n].lo = lo;
  kb->n++;
}

/* 1: or_discover drops the records of rays with a near tie, and
 * or_discover_ref those of launch threads with one (test bookkeeping) */
static int g_skip_ties = 0;
void or_set_skip_ties(int on) { g_skip_ties = on; }

/* A query's outcome is a near tie when a sphere or triangle decision or the
 * order of the two hits could go the other way under a perturbation of
 * g_tie_eps metres: the accepted hits' margins, near misses up to the event
 * that ends the query, and a sphere hit just before the triangle hit. */
static inline int query_tie(const tri_q *tq, const sph_q *sq) {
  double eps = g_tie_eps;
  int sph = sq->idx >= 0 && sq->t < tq->t;
  double t_end = sph ? sq->t : tq->t;
  if (sph && (sq->margin < eps || tq->t - sq->t < eps)) return 1;
  if (isfinite(sq->near_t) && sq->near_t <= t_end + eps) return 1;
  if (tq->tri >= 0 && tq->edge_tie) return 1;
  if (tq->tmin_tie) return 1;
  if (isfinite(tq->near_miss) && tq->near_miss <= tq->t + eps) return 1;
  return 0;
}

/* One primary ray: optixPathTracer.cu:150-189 with closest-hit forward
 * :285-331.  Appends (key, tile) records; optional per-event log. */
static void forward_ray(const or_scene *s, const or_spheres *rx, v3 o, v3 d, int K, int n1, int gs,
                        kbuf *out, long *nq, long *nties, int *log, int log_cap, int *log_n) {
  fbranch st[64];
  int sp = 0;
  fbranch cur;
  cur.o = o;
  cur.d = d;
  cur.ttl = K + 1;
  cur.b = 0;
  cur.nref = 0;
  cur.last_pl = -1;
  for (int i = 0; i < OR_MAX_HIT; i++) cur.pl[i] = OR_NO_PLANE;
  st[sp++] = cur;
  int tie = 0;
  while (sp) {
    cur = st[--sp];
    while (cur.ttl > 0) {
      tri_q tq = closest_tri(s, cur.o, cur.d, OR_TMIN, INFINITY, cur.last_pl);
      sph_q sq = closest_sphere(rx, cur.o, cur.d, OR_TMIN, tq.t);
      (*nq)++;
      if (query_tie(&tq, &sq)) tie = 1;
      if (sq.idx >= 0 && sq.t < tq.t) {
        int h = K + 1 - cur.ttl;
        uint64_t hi, lo;
        key_pack(h, cur.b, cur.pl, tile_of(sq.idx, n1, gs), &hi, &lo);
        kbuf_push(out, hi, lo);
        if (log && *log_n + 2 <= log_cap) { log[(*log_n)++] = 1; log[(*log_n)++] = sq.idx; }
        cur.o = add(cur.o, scl(cur.d, sq.t));
        cur.last_pl = -1;
        continue;
      }
      if (tq.tri < 0) break; /* miss: optixPathTracer.cu:276-279 */
      int tri = tq.tri;
      if (log && *log_n + 2 <= log_cap) { log[(*log_n)++] = 0; log[(*log_n)++] = tri; }
      cur.o = add(cur.o, scl(cur.d, tq.t));
      cur.last_pl = s->plane[tri];
      if (cur.ttl > 1) {
        v3 N0 = s->n0[tri];
        v3 N = dot(N0, scl(cur.d, -1.0)) < 0 ? scl(N0, -1.0) : N0; /* faceforward */
        int m = s->mat[tri];
        int th = theta_bin(cur.d, N);
        int pen = s->mat_prop[th + 19 * m] < 0;
        int blo, bhi;
        if (theta_edge(cur.d, N, &blo, &bhi) && (s->mat_prop[blo + 19 * m] < 0) != (s->mat_prop[bhi + 19 * m] < 0))
          tie = 1; /* penetrability of an angle at a bin edge */
        cur.ttl--;
        int h = K + 1 - cur.ttl;
        fbranch ch = cur;
        ch.b = cur.b | (1 << (h - 1));
        ch.d = sub(cur.d, scl(N, 2.0 * dot(cur.d, N)));
        if (ch.nref < OR_MAX_HIT) ch.pl[ch.nref] = (unsigned)s->plane[tri];
        ch.nref++;
        if (pen && sp < 64) st[sp++] = cur; /* transmitted parent continues */
        cur = ch;
      }
      /* ttl == 1: the scene is pass-through, optixPathTracer.cu:294 */
    }
  }
  if (tie) (*nties)++;
}

/* Discovery over rays [ray_begin, ray_end).  per_element != 0: ray r uses
 * element r / fib_n and direction r % fib_n of a fib_n-point set; else ray r
 * uses element r % n_src and direction r of fib_n.  Returns the number of
 * unique (key, tile) records written to out (at most cap; -1 if cap is too
 * small, with the needed count in stats->n_records). */
long or_discover(void *scene, const float *src, int n_src, const float *dst, int n0_dst, int n1_dst,
                 double dst_r, int max_reflection, int gridsize, long fib_n, int per_element,
                 long ray_begin, long ray_end, uint64_t *out, long cap, long *stats) {
  or_scene *s = (or_scene *)scene;
  or_spheres *rx = spheres_create(n0_dst * n1_dst, dst, dst_r);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  kbuf *bufs = calloc(nt, sizeof(kbuf));
  long nq = 0, nties = 0;
#pragma omp parallel reduction(+ : nq, nties)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    kbuf *kb = &bufs[tid];
#pragma omp for schedule(dynamic, 1024)
    for (long r = ray_begin; r < ray_end; r++) {
      long e = per_element ? r / fib_n : r % n_src;
      long k = per_element ? r % fib_n : r;
      double dd[3];
      or_fib_dir(k, fib_n, dd);
      v3 o = mk(src[3 * e], src[3 * e + 1], src[3 * e + 2]);
      long before = kb->n, ties_before = nties;
      forward_ray(s, rx, o, mk(dd[0], dd[1], dd[2]), max_reflection, n1_dst, gridsize, kb, &nq, &nties,
                  NULL, 0, NULL);
      if (g_skip_ties && nties > ties_before) kb->n = before;
      /* drop consecutive duplicates of this ray */
      long m = before;
      for (long i = before; i < kb->n; i++) {
        int dup = 0;
        for (long j = before; j < m; j++)
          if (kb->k[j].hi == kb->k[i].hi && kb->k[j].lo == kb->k[i].lo) { dup = 1; break; }
        if (!dup) kb->k[m++] = kb->k[i];
      }
      kb->n = m;
      if (kb->n > (1 << 22)) kb->n = key_unique(kb->k, kb->n);
    }
  }
  long total = 0;
  for (int t = 0; t < nt; t++) total += bufs[t].n;
  key128 *all = malloc(sizeof(key128) * (total + 1));
  long off = 0;
  for (int t = 0; t < nt; t++) {
    memcpy(all + off, bufs[t].k, sizeof(key128) * bufs[t].n);
    off += bufs[t].n;
    free(bufs[t].k);
  }
  free(bufs);
  long m = key_unique(all, total);
  if (stats) {
    stats[0] = ray_end - ray_begin;
    stats[1] = nq;
    stats[2] = m;
    stats[3] = nties;
  }
  long ret = m;
  if (m > cap) ret = -1;
  else
    for (long i = 0; i < m; i++) { out[2 * i] = all[i].hi; out[2 * i + 1] = all[i].lo; }
  free(all);
  spheres_destroy(rx);
  return ret;
}

/* single-ray trace with the event log: entries (kind, index) with kind 0 =
 * triangle, 1 = Rx sphere.  Returns the number of ints written; the tie flag
 * is returned through *tie. */
int or_trace_ray_log(void *scene, const double *o, const double *d, const float *dst, int n_dst,
                     double dst_r, int max_reflection, int n1_dst, int gridsize, int *log, int log_cap,
                     int *tie) {
  or_scene *s = (or_scene *)scene;
  or_spheres *rx = spheres_create(n_dst, dst, dst_r);
  kbuf kb = {0};
  long nq = 0, nt = 0;
  int n = 0;
  forward_ray(s, rx, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), max_reflection, n1_dst, gridsize, &kb,
              &nq, &nt, log, log_cap, &n);
  free(kb.k);
  spheres_destroy(rx);
  if (tie) *tie = (int)nt;
  return n;
}

/* closest triangle for tests (brute force if use_bvh == 0) */
int or_closest_hit(void *scene, const double *o, const double *d, double tmin, int use_bvh, double *t_out) {
  or_scene *s = (or_scene *)scene;
  v3 O = mk(o[0], o[1], o[2]), D = mk(d[0], d[1], d[2]);
  if (use_bvh) {
    tri_q q = closest_tri(s, O, D, tmin, INFINITY, -1);
    *t_out = q.t;
    return q.tri;
  }
  double best = INFINITY;
  int bi = -1;
  for (int i = 0; i < s->n_tri; i++) {
    double m;
    double t = tri_hit(s, i, O, D, &m);
    if (m >= 0.0 && t >= tmin && (t < best || (t == best && i < bi))) { best = t; bi = i; }
  }
  *t_out = best;
  return bi;
}

/* ----------------------------------------------------- backward pass ---- */
typedef struct {
  double wave_len, src_r;
  v3 src_dir, dst_dir;
  int src_bptype, dst_bptype;
} bw_par;

/* image of p across the reflection planes of a path, first plane first */
static v3 image_of(const or_scene *s, const unsigned *pl, int nref, v3 p) {
  for (int k = 0; k < nref; k++) {
    int t = (int)pl[k];
    v3 n = s->pl_n[t];
    double dd = s->pl_d[t];
    p = sub(p, scl(n, 2.0 * (dot(p, n) - dd)));
  }
  return p;
}

static inline double bp(v3 a, v3 b, int type) {
  double x[3] = {a.x, a.y, a.z}, y[3] = {b.x, b.y, b.z};
  return or_beam_pattern(x, y, type);
}

/* 1 when the angle of a to b lies within g_tie_ang of a jump of the beam
 * pattern: type 2 at 2 pi / 3 (0.24495 -> 0.1), type 3 at pi / 2 */
static inline int bp_tie(v3 a, v3 b, int type) {
  if (type != 2 && type != 3) return 0;
  double c = dot(a, b) / (len(a) * len(b));
  double th = acos(fmax(-1.0, fmin(1.0, c)));
  double jump = type == 2 ? M_PI * 2 / 3 : M_PI / 2;
  return fabs(th - jump) < g_tie_ang;
}

/* One backward chain: optixPathTracer.cu:223-271 with :332-417.  Returns 1
 * on success and the complex contribution A e^{-j phi} e^{j psi}. */
static int backward_chain(const or_scene *s, const bw_par *P, int h, int bmask, const unsigned *pl, v3 si,
                          v3 rj, double *re, double *im, double *L_out, int *tie, long *nq) {
  int nref = 0;
  for (int k = 0; k < h; k++) nref += (bmask >> k) & 1;
  v3 I = image_of(s, pl, nref, si);
  v3 o = rj;
  v3 d = nrm(sub(I, rj));
  double A = bp(d, P->dst_dir, P->dst_bptype);
  if (bp_tie(d, P->dst_dir, P->dst_bptype)) *tie = 1;
  double L = 0.0, D = 0.0, Dres = 0.0, far = 1e6, psi = 0.0;
  int ttl = h + 1, ri = 0, last_pl = -1;
  double acc = fmax(P->src_r, OR_ACC_MIN), acc2 = acc * acc, r2 = P->src_r * P->src_r;
  /* reflection planes are consumed from the Rx side: the last forward plane first */
  while (ttl > 0) {
    tri_q tq = closest_tri(s, o, d, OR_TMIN, INFINITY, last_pl);
    (*nq)++;
    v3 O = sub(si, o);
    double bs = dot(O, d), c = dot(O, O), perp2 = c - bs * bs;
    int sh = perp2 < acc2 && bs >= OR_TMIN && c > r2;
    /* near ties: the own Tx sphere's boundary, its order against the
     * triangle hit, the triangle hit's edge and near-missed triangles */
    if (bs >= OR_TMIN && c > r2 && fabs(sqrt(fmax(perp2, 0)) - acc) < g_tie_eps && bs <= tq.t + g_tie_eps)
      *tie = 1;
    if (sh && fabs(tq.t - bs) < g_tie_eps) *tie = 1;
    if (isfinite(tq.near_miss) && tq.near_miss <= ((sh && bs < tq.t) ? bs : tq.t) + g_tie_eps) *tie = 1;
    if (sh && bs < tq.t) {
      if (ttl != 1) return 0; /* own Tx reached early: optixPathTracer.cu:396-415 */
      L += bs;
      D += bs;
      double f = bp(scl(d, -1.0), P->src_dir, P->src_bptype) * P->wave_len / (4 * M_PI * (Dres + D));
      if (bp_tie(scl(d, -1.0), P->src_dir, P->src_bptype)) *tie = 1;
      A *= f;
      double ph = 2 * M_PI * fmod(L / P->wave_len, 1.0) - psi;
      *re = A * cos(ph);
      *im = -A * sin(ph);
      *L_out = L;
      return 1;
    }
    if (tq.tmin_tie) *tie = 1;
    if (tq.tri < 0) return 0; /* miss */
    if (tq.edge_tie) *tie = 1;
    int tri = tq.tri;
    o = add(o, scl(d, tq.t));
    L += tq.t;
    D += tq.t;
    last_pl = s->plane[tri];
    if (ttl == 1) return 0; /* blocked: optixPathTracer.cu:348-349 */
    v3 N0 = s->n0[tri];
    v3 N = dot(N0, scl(d, -1.0)) < 0 ? scl(N0, -1.0) : N0;
    int m = s->mat[tri];
    int th = theta_bin(d, N);
    double loss = fabs(s->mat_prop[th + 19 * m]);
    int blo, bhi;
    if (theta_edge(d, N, &blo, &bhi) && fabs(s->mat_prop[blo + 19 * m]) != fabs(s->mat_prop[bhi + 19 * m]))
      *tie = 1; /* the loss of an angle at a bin edge */
    ttl--;
    int k = h + 1 - ttl;
    int refl = (bmask >> (k - 1)) & 1;
    if (refl) {
      /* the k-th interaction reflects on the path's plane (Rx side first) */
      if (ri >= nref || (unsigned)s->plane[tri] != pl[nref - 1 - ri]) return 0;
      ri++;
    }
    double fspl = 1.0;
    double dtol = g_tie_eps + 1e-6 * D; /* FP32 path lengths */
    if (far < 1e6 && fabs(D - far) < dtol) *tie = 1;
    if (fabs(D - 2.0 * s->maxedge[tri] * s->maxedge[tri] / P->wave_len) < dtol) *tie = 1;
    if (D > far) { /* Fraunhofer split, optixPathTracer.cu:369-391 */
      fspl = P->wave_len / (4 * M_PI * Dres);
      Dres = 0.0;
    }
    A *= loss * fspl;
    Dres += D;
    double fk = 2.0 * s->maxedge[tri] * s->maxedge[tri] / P->wave_len;
    far = (D > fk) ? fk : 1e6;
    D = 0.0;
    if (refl) {
      d = sub(d, scl(N, 2.0 * dot(d, N)));
      psi += s->refl_phase[m];
    }
  }
  return 0;
}

typedef struct {
  long n_chains, n_success, n_queries, n_ties;
} or_eval_stats;

/* when set, or_evaluate marks tie_mask[i * n_dst + j] for every (Tx i, Rx j)
 * with a near-tie chain (test bookkeeping) */
static unsigned char *g_tie_mask = NULL;
static long g_tie_mask_n = 0;
void or_set_tie_mask(unsigned char *mask, long n) {
  g_tie_mask = mask;
  g_tie_mask_n = n;
}

/* Evaluate tasks.  paths: [n_paths][2] u64 keys (tile field ignored);
 * task_path/task_tile: [n_tasks]; H: [n_src][n_dst][2] FP64, accumulated;
 * task_ok: [n_tasks] set to 1 if any chain of the task succeeds. */
void or_evaluate(void *scene, const float *src, int n_src, const float *dst, int n0_dst, int n1_dst,
                 double wave_len, double src_r, const float *src_dir, const float *dst_dir,
                 int src_bptype, int dst_bptype, int gridsize, const uint64_t *paths, int n_paths,
                 const int *task_path, const int *task_tile, long n_tasks, int flags, double *H,
                 unsigned char *task_ok, long *stats) {
  or_scene *s = (or_scene *)scene;
  bw_par P;
  P.wave_len = wave_len;
  P.src_r = src_r;
  P.src_dir = mk(src_dir[0], src_dir[1], src_dir[2]);
  P.dst_dir = mk(dst_dir[0], dst_dir[1], dst_dir[2]);
  P.src_bptype = src_bptype;
  P.dst_bptype = dst_bptype;
  int n_dst = n0_dst * n1_dst;
  int n1g = (n1_dst - 1) / gridsize + 1;
  long nchains = 0, nsucc = 0, nq = 0, nties = 0;
  (void)n_paths;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nchains, nsucc, nq, nties)
  for (long t = 0; t < n_tasks; t++) {
    int p = task_path[t], g = task_tile[t];
    int h, b;
    unsigned pl[5], tile;
    key_unpack(paths[2 * p], paths[2 * p + 1], &h, &b, pl, &tile);
    int gi = g / n1g, gj = g % n1g;
    int ok = 0;
    for (int q = 0; q < gridsize * gridsize; q++) {
      int di = gi * gridsize + q / gridsize, dj = gj * gridsize + q % gridsize;
      /* optixPathTracer.cu:216-220 checks only dst < n_dst, so a column past
       * the grid wraps into the next row; flags bit 0 reproduces that. */
      if (dj >= n1_dst && !(flags & 1)) continue;
      int j = di * n1_dst + dj;
      if (di >= n0_dst || j >= n_dst) continue;
      v3 rj = mk(dst[3 * j], dst[3 * j + 1], dst[3 * j + 2]);
      for (int i = 0; i < n_src; i++) {
        v3 si = mk(src[3 * i], src[3 * i + 1], src[3 * i + 2]);
        double re, im, L;
        int tie = 0;
        nchains++;
        if (backward_chain(s, &P, h, b, pl, si, rj, &re, &im, &L, &tie, &nq)) {
          nsucc++;
          ok = 1;
          double *hp = H + 2 * ((long)i * n_dst + j);
#pragma omp atomic
          hp[0] += re;
#pragma omp atomic
          hp[1] += im;
        }
        nties += tie;
        if (tie && g_tie_mask && (long)i * n_dst + j < g_tie_mask_n) g_tie_mask[(long)i * n_dst + j] = 1;
      }
    }
    if (task_ok) task_ok[t] = (unsigned char)ok;
  }
  if (stats) {
    stats[0] = nchains;
    stats[1] = nsucc;
    stats[2] = nq;
    stats[3] = nties;
  }
}

/* One chain for tests: returns success and the contribution / path length. */
int or_backward_one(void *scene, const double *si, const double *rj, int h, int bmask, const unsigned *pl,
                    double wave_len, double src_r, const double *src_dir, const double *dst_dir,
                    int src_bptype, int dst_bptype, double *out /* re, im, L */) {
  or_scene *s = (or_scene *)scene;
  bw_par P;
  P.wave_len = wave_len;
  P.src_r = src_r;
  P.src_dir = mk(src_dir[0], src_dir[1], src_dir[2]);
  P.dst_dir = mk(dst_dir[0], dst_dir[1], dst_dir[2]);
  P.src_bptype = src_bptype;
  P.dst_bptype = dst_bptype;
  int tie = 0;
  long nq = 0;
  double re = 0, im = 0, L = 0;
  int ok = backward_chain(s, &P, h, bmask, pl, mk(si[0], si[1], si[2]), mk(rj[0], rj[1], rj[2]), &re, &im,
                          &L, &tie, &nq);
  out[0] = re;
  out[1] = im;
  out[2] = L;
  return ok;
}

/* --------------------------------------------- neighbour expansion ---- */
/* optixPathTracer.cpp:229-354 restated: rounds of tasks until a fixpoint.
 * The traced (path, tile) set is a hash set of 64-bit (path << 24 | tile). */
typedef struct {
  uint64_t *slot;
  long cap, n;
} hset;

static inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

static int hset_insert(hset *h, uint64_t k) { /* returns 1 if newly inserted; k != ~0 */
  if (2 * (h->n + 1) > h->cap) {
    long oc = h->cap;
    uint64_t *os = h->slot;
    h->cap = oc ? 2 * oc : 1 << 16;
    h->slot = malloc(sizeof(uint64_t) * h->cap);
    memset(h->slot, 0xFF, sizeof(uint64_t) * h->cap);
    h->n = 0;
    for (long i = 0; i < oc; i++)
      if (os[i] != ~0ULL) hset_insert(h, os[i]);
    free(os);
  }
  long m = h->cap - 1, i = (long)(mix64(k) & (uint64_t)m);
  while (h->slot[i] != ~0ULL) {
    if (h->slot[i] == k) return 0;
    i = (i + 1) & m;
  }
  h->slot[i] = k;
  h->n++;
  return 1;
}

/* Full trace: discovery, dedup, task rounds.  H_out [n_src][n_dst][2] FP64 is
 * accumulated.  stats: [0] rays [1] fwd queries [2] unique records
 * [3] fwd ties [4] paths [5] tasks total [6] rounds [7] chains [8] successes
 * [9] bwd queries [10] bwd ties.  Returns 0, or -1 if records exceed the
 * internal cap. */
int or_trace_channel(void *scene, const float *src, int n0_src, int n1_src, const float *dst, int n0_dst,
                     int n1_dst, double wave_len, int max_reflection, double src_r, double dst_r,
                     const float *src_dir, const float *dst_dir, int src_bptype, int dst_bptype,
                     int gridsize, long fib_n, int per_element, int flags, const uint64_t *seed,
                     long n_seed, double *H_out, long *stats, uint64_t *paths_out, long paths_cap,
                     long *n_paths_out) {
  int n_src = n0_src * n1_src, n_dst = n0_dst * n1_dst;
  long n_rays = per_element ? (long)n_src * fib_n : fib_n;
  long fst[4] = {0, 0, 0, 0};
  uint64_t *rec;
  long m;
  if (seed) {
    /* start from given (path, tile) records instead of discovery */
    rec = malloc(sizeof(uint64_t) * 2 * (n_seed + 1));
    memcpy(rec, seed, sizeof(uint64_t) * 2 * n_seed);
    m = key_unique((key128 *)rec, n_seed);
    fst[2] = m;
  } else {
    long cap = 1L << 26;
    rec = malloc(sizeof(uint64_t) * 2 * cap);
    m = or_discover(scene, src, n_src, dst, n0_dst, n1_dst, dst_r, max_reflection, gridsize, fib_n,
                    per_element, 0, n_rays, rec, cap, fst);
    if (m < 0) { free(rec); return -1; }
  }
  /* unique path keys (records are sorted by (path, tile)) */
  int *rec_path = malloc(sizeof(int) * (m + 1));
  uint64_t *paths = malloc(sizeof(uint64_t) * 2 * (m + 1));
  long np = 0;
  for (long i = 0; i < m; i++) {
    uint64_t hi = rec[2 * i], lo = rec[2 * i + 1] & ~0xFFFFFULL;
    if (np == 0 || paths[2 * (np - 1)] != hi || paths[2 * (np - 1) + 1] != lo) {
      paths[2 * np] = hi;
      paths[2 * np + 1] = lo;
      np++;
    }
    rec_path[i] = (int)(np - 1);
  }
  int n1g = (n1_dst - 1) / gridsize + 1, n0g = (n0_dst - 1) / gridsize + 1;
  long n_grid = (long)n0g * n1g;
  hset traced = {0};
  long ntask = m;
  int *tp = malloc(sizeof(int) * (ntask + 1)), *tt = malloc(sizeof(int) * (ntask + 1));
  /* round 1 tasks ordered by tile then path, as optixPathTracer.cpp:325-330 */
  long k = 0;
  {
    /* counting sort by tile */
    long *cnt = calloc(n_grid + 1, sizeof(long));
    for (long i = 0; i < m; i++) cnt[(rec[2 * i + 1] & 0xFFFFF) + 1]++;
    for (long g = 0; g < n_grid; g++) cnt[g + 1] += cnt[g];
    for (long i = 0; i < m; i++) {
      long g = (long)(rec[2 * i + 1] & 0xFFFFF);
      long pos = cnt[g]++;
      tp[pos] = rec_path[i];
      tt[pos] = (int)g;
    }
    free(cnt);
    for (long i = 0; i < m; i++) hset_insert(&traced, ((uint64_t)tp[i] << 24) | (uint64_t)tt[i]);
    k = m;
  }
  free(rec);
  free(rec_path);
  long tasks_total = 0, ch = 0, su = 0, bq = 0, bt = 0;
  int rounds = 0;
  while (k > 0) {
    rounds++;
    unsigned char *ok = malloc(k + 1);
    long est[4];
    or_evaluate(scene, src, n_src, dst, n0_dst, n1_dst, wave_len, src_r, src_dir, dst_dir, src_bptype,
                dst_bptype, gridsize, paths, (int)np, tp, tt, k, flags, H_out, ok, est);
    tasks_total += k;
    ch += est[0];
    su += est[1];
    bq += est[2];
    bt += est[3];
    /* expansion: optixPathTracer.cpp:264-284 (linear bounds only) */
    long ncap = 4 * k + 16, nk = 0;
    int *np_ = malloc(sizeof(int) * ncap), *nt_ = malloc(sizeof(int) * ncap);
    /* new tasks are grouped per tile in tile order (optixPathTracer.cpp:325) */
    for (long i = 0; i < k; i++) {
      if (!ok[i]) continue;
      long g = tt[i];
      long nb[4] = {g - 1, g + 1, g - n1g, g + n1g};
      for (int q = 0; q < 4; q++) {
        if (nb[q] < 0 || nb[q] >= n_grid) continue;
        if (hset_insert(&traced, ((uint64_t)tp[i] << 24) | (uint64_t)nb[q])) {
          np_[nk] = tp[i];
          nt_[nk] = (int)nb[q];
          nk++;
        }
      }
    }
    free(ok);
    free(tp);
    free(tt);
    if (flags & 2) nk = 0; /* evaluate the given tasks only */
    /* stable order by tile */
    long *cnt = calloc(n_grid + 1, sizeof(long));
    for (long i = 0; i < nk; i++) cnt[nt_[i] + 1]++;
    for (long g = 0; g < n_grid; g++) cnt[g + 1] += cnt[g];
    tp = malloc(sizeof(int) * (nk + 1));
    tt = malloc(sizeof(int) * (nk + 1));
    for (long i = 0; i < nk; i++) {
      long pos = cnt[nt_[i]]++;
      tp[pos] = np_[i];
      tt[pos] = nt_[i];
    }
    free(cnt);
    free(np_);
    free(nt_);
    k = nk;
  }
  free(tp);
  free(tt);
  free(traced.slot);
  if (stats) {
    stats[0] = fst[0]; stats[1] = fst[1]; stats[2] = fst[2]; stats[3] = fst[3];
    stats[4] = np; stats[5] = tasks_total; stats[6] = rounds; stats[7] = ch;
    stats[8] = su; stats[9] = bq; stats[10] = bt;
  }
  if (n_paths_out) *n_paths_out = np;
  if (paths_out) {
    long c = np < paths_cap ? np : paths_cap;
    memcpy(paths_out, paths, sizeof(uint64_t) * 2 * c);
  }
  free(paths);
  (void)n_dst;
  return 0;
}

/* ------------------------------------ reference-compatible discovery ---- */
/* optixPathTracer.cu:142-190 with random.h's generators: TEA (4 rounds) seeds
 * a Numerical-Recipes LCG per launch index; the ray's Tx element and the
 * stratified direction of rand_dir are computed in f32 as the reference does,
 * then traced in FP64; the 32-slot branch tree is walked in slot order with
 * per-tile min_hc gating and the reserve cap. */
/* TEA (Wheeler & Needham 1994) with `rounds` rounds and key k; random.h:31-45
 * seeds with 4 rounds and the key below. */
void or_tea(unsigned v0, unsigned v1, const unsigned *k, int rounds, unsigned *out) {
  unsigned s0 = 0;
  for (int n = 0; n < rounds; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + k[0]) ^ (v1 + s0) ^ ((v1 >> 5) + k[1]);
    v1 += ((v0 << 4) + k[2]) ^ (v0 + s0) ^ ((v0 >> 5) + k[3]);
  }
  out[0] = v0;
  out[1] = v1;
}
static const unsigned or_tea_key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
static unsigned or_tea4(unsigned v0, unsigned v1) {
  unsigned o[2];
  or_tea(v0, v1, or_tea_key, 4, o);
  return o[0];
}
/* Numerical Recipes LCG and its [0, 1) float, random.h:47-67 */
unsigned or_lcg(unsigned st) { return 1664525u * st + 1013904223u; }
static float or_rnd(unsigned *st) {
  *st = or_lcg(*st);
  return (float)(*st & 0x00FFFFFFu) / (float)0x01000000u;
}
/* the sample-th ray of launch index t = ix * launch_y + iy: Tx element and
 * rand_dir's stratified direction (optixPathTracer.cu:56-68,143,151) */
static int or_ref_sample(unsigned *seed, int ix, int iy, int launch_x, int launch_y, int n_src, float *d) {
  int e = (int)(or_rnd(seed) * (float)n_src);
  if (e >= n_src) e = n_src - 1;
  float r1 = (or_rnd(seed) + (float)ix) / (float)launch_x;
  float r2 = (or_rnd(seed) + (float)iy) / (float)launch_y;
  float th = 2.0f * (float)M_PI * r1, rr = sqrtf(r2);
  float u = rr * cosf(th), v = rr * sinf(th), q = sqrtf(fmaxf(0.0f, 1.0f - rr * rr));
  d[0] = 2.0f * u * q;
  d[1] = 2.0f * v * q;
  d[2] = 1.0f - 2.0f * rr * rr;
  return e;
}
/* every ray of a launch: elem[t * samples + s], dir[3 * (t * samples + s)] */
void or_ref_rays(int launch_x, int launch_y, int samples, int n_src, int *elem, float *dir) {
  for (long t = 0; t < (long)launch_x * launch_y; t++) {
    unsigned seed = or_tea4((unsigned)t, 0u);
    for (int smp = 0; smp < samples; smp++) {
      long j = t * samples + smp;
      elem[j] = or_ref_sample(&seed, (int)(t / launch_y), (int)(t % launch_y), launch_x, launch_y, n_src, dir + 3 * j);
    }
  }
}

/* One launch thread t of reference-mode discovery: its samples, each a
 * 32-slot branch tree walked in slot order, with per-thread min_hc gating and
 * the reserve cap (optixPathTracer.cu:143-189).  Keys go to kb; when log is
 * given, every Rx event appends (t, sample, h, tile, recorded) to it. */
static long ref_thread(const or_scene *s, const or_spheres *rx, const float *src, int n_src, int n1_dst,
                       int gridsize, int n_grid, int K, int launch_x, int launch_y, int samples, int reserve,
                       int gating, long t, int *min_hc, fbranch *slot, kbuf *kb, int *log, long log_cap,
                       long *log_n, long *n_tie_threads) {
  long nq = 0, kb_before = kb->n;
  int tie = 0;
  int ix = (int)(t / launch_y), iy = (int)(t % launch_y);
  unsigned seed = or_tea4((unsigned)t, 0u);
  memset(min_hc, 0, sizeof(int) * n_grid);
  int res_cnt = 0;
  for (int smp = 0; smp < samples; smp++) {
    float dd[3];
    int e = or_ref_sample(&seed, ix, iy, launch_x, launch_y, n_src, dd);
    for (int k = 0; k < (1 << OR_MAX_HIT); k++) slot[k].ttl = 0;
    slot[0].o = mk(src[3 * e], src[3 * e + 1], src[3 * e + 2]);
    slot[0].d = nrm(mk(dd[0], dd[1], dd[2]));
    slot[0].ttl = K + 1;
    slot[0].b = 0;
    slot[0].nref = 0;
    slot[0].last_pl = -1;
    for (int k = 0; k < OR_MAX_HIT; k++) slot[0].pl[k] = OR_NO_PLANE;
    for (int i = 0; i < (1 << OR_MAX_HIT); i++) {
      fbranch cur = slot[i];
      while (cur.ttl > 0) {
        tri_q tq = closest_tri(s, cur.o, cur.d, OR_TMIN, INFINITY, cur.last_pl);
        sph_q sq = closest_sphere(rx, cur.o, cur.d, OR_TMIN, tq.t);
        nq++;
        if (query_tie(&tq, &sq)) tie = 1;
        if (sq.idx >= 0 && sq.t < tq.t) {
          int grid = (int)tile_of(sq.idx, n1_dst, gridsize);
          int h = K + 1 - cur.ttl;
          int rec = 0;
          if (!gating || res_cnt < reserve) {
            if (!gating || min_hc[grid] <= h) {
              min_hc[grid] = h + 1;
              uint64_t hi, lo;
              key_pack(h, cur.b, cur.pl, (unsigned)grid, &hi, &lo);
              kbuf_push(kb, hi, lo);
              res_cnt++;
              rec = 1;
            }
          }
          if (log && *log_n + 5 <= log_cap) {
            int *ev = log + *log_n;
            ev[0] = (int)t;
            ev[1] = smp;
            ev[2] = h;
            ev[3] = grid;
            ev[4] = rec;
            *log_n += 5;
          }
          cur.o = add(cur.o, scl(cur.d, sq.t));
          cur.last_pl = -1;
          continue;
        }
        if (tq.tri < 0) break;
        int tri = tq.tri;
        cur.o = add(cur.o, scl(cur.d, tq.t));
        cur.last_pl = s->plane[tri];
        if (cur.ttl > 1) {
          v3 N0 = s->n0[tri];
          v3 N = dot(N0, scl(cur.d, -1.0)) < 0 ? scl(N0, -1.0) : N0;
          int m = s->mat[tri];
          int thb = theta_bin(cur.d, N);
          int pen = s->mat_prop[thb + 19 * m] < 0;
          int blo, bhi;
          if (theta_edge(cur.d, N, &blo, &bhi) &&
              (s->mat_prop[blo + 19 * m] < 0) != (s->mat_prop[bhi + 19 * m] < 0))
            tie = 1;
          cur.ttl--;
          int h = K + 1 - cur.ttl;
          int cs = cur.b | (1 << (h - 1));
          fbranch ch = cur;
          ch.b = cs;
          ch.d = sub(cur.d, scl(N, 2.0 * dot(cur.d, N)));
          if (ch.nref < OR_MAX_HIT) ch.pl[ch.nref] = (unsigned)s->plane[tri];
          ch.nref++;
          slot[cs] = ch;
          if (!pen) break;
        }
      }
    }
  }
  if (tie && g_skip_ties) kb->n = kb_before;
  if (tie && n_tie_threads) (*n_tie_threads)++;
  return nq;
}

long or_discover_ref(void *scene, const float *src, int n_src, const float *dst, int n0_dst, int n1_dst,
                     double dst_r, int max_reflection, int gridsize, int launch_x, int launch_y, int samples,
                     int reserve, int gating, uint64_t *out, long cap, long *stats) {
  or_scene *s = (or_scene *)scene;
  or_spheres *rx = spheres_create(n0_dst * n1_dst, dst, dst_r);
  int n1g = (n1_dst - 1) / gridsize + 1, n0g = (n0_dst - 1) / gridsize + 1, n_grid = n0g * n1g;
  long nthr = (long)launch_x * launch_y;
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  kbuf *bufs = calloc(nt, sizeof(kbuf));
  long nq = 0, ntie = 0;
#pragma omp parallel reduction(+ : nq, ntie)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    kbuf *kb = &bufs[tid];
    int *min_hc = malloc(sizeof(int) * n_grid);
    fbranch slot[1 << OR_MAX_HIT];
#pragma omp for schedule(dynamic, 16)
    for (long t = 0; t < nthr; t++)
      nq += ref_thread(s, rx, src, n_src, n1_dst, gridsize, n_grid, max_reflection, launch_x, launch_y, samples,
                       reserve, gating, t, min_hc, slot, kb, NULL, 0, NULL, &ntie);
    free(min_hc);
  }
  long total = 0;
  for (int t = 0; t < nt; t++) total += bufs[t].n;
  key128 *all = malloc(sizeof(key128) * (total + 1));
  long off = 0;
  for (int t = 0; t < nt; t++) {
    memcpy(all + off, bufs[t].k, sizeof(key128) * bufs[t].n);
    off += bufs[t].n;
    free(bufs[t].k);
  }
  free(bufs);
  long m = key_unique(all, total);
  if (stats) {
    stats[0] = nthr * samples;
    stats[1] = nq;
    stats[2] = m;
    stats[3] = total;
    stats[4] = ntie;
  }
  long ret = m <= cap ? m : -1;
  if (ret >= 0)
    for (long i = 0; i < m; i++) { out[2 * i] = all[i].hi; out[2 * i + 1] = all[i].lo; }
  free(all);
  spheres_destroy(rx);
  return ret;
}

/* The Rx events of launch threads [t_begin, t_end) of or_discover_ref, in
 * the order each thread meets them: 5 ints per event (t, sample, h, tile,
 * recorded).  Returns the number of ints written (at most log_cap). */
long or_discover_ref_log(void *scene, const float *src, int n_src, const float *dst, int n0_dst, int n1_dst,
                         double dst_r, int max_reflection, int gridsize, int launch_x, int launch_y, int samples,
                         int reserve, int gating, long t_begin, long t_end, int *log, long log_cap) {
  or_scene *s = (or_scene *)scene;
  or_spheres *rx = spheres_create(n0_dst * n1_dst, dst, dst_r);
  int n1g = (n1_dst - 1) / gridsize + 1, n0g = (n0_dst - 1) / gridsize + 1, n_grid = n0g * n1g;
  int *min_hc = malloc(sizeof(int) * n_grid);
  fbranch slot[1 << OR_MAX_HIT];
  kbuf kb = {0};
  long n = 0;
  for (long t = t_begin; t < t_end; t++)
    ref_thread(s, rx, src, n_src, n1_dst, gridsize, n_grid, max_reflection, launch_x, launch_y, samples, reserve,
               gating, t, min_hc, slot, &kb, log, log_cap, &n, NULL);
  free(kb.k);
  free(min_hc);
  spheres_destroy(rx);
  return n;
}
