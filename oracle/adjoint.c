/*
 * adjoint.c -- FP64 restatement of the differentiable renderer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/adjoint
 * kernels (csrc/diff.hip), loaded by the test suite.
 *
 * The differentiable image is the expectation of the reference's
 * stochastic marching (deviceCodeRBF.hlsl:67-139, primary ray, ambient 1)
 * over its blue-noise threshold: a sample marches the points
 * t_k = tenter + unit*(r1 + k) (r1 the sample's first LCG draw, as in
 * ParticleRBFRayGen), accumulates L_k = D_1 + ... + D_k over the points
 * some particle covers (D = densitymap(density)), and stops at point k with
 * probability F(L_k) - F(L_{k-1}), F(L) = min(1, (1 - e^-L) / 0.99) =
 * P(-ln(1 - 0.99 xi) < L) for uniform xi.  The sample's colour is
 *     C = sum_k (F(L_k) - F(L_{k-1})) c_k,
 * c_k the attribute-weighted colour at point k, and a pixel is the mean of
 * its samples.  C is a smooth function of particle positions, radii,
 * attributes and transfer-function texels except where a sphere boundary
 * crosses a march point (the truncated RBF jumps there; such jumps are not
 * differentiated).
 *
 * The adjoint replays each sample twice (path replay, nothing stored per
 * step): the first pass sums dC/dL_k = F'(L_k)(c_k - c_{k+1}) over all k,
 * the second recovers dC/dD_j = sum_{k>=j} dC/dL_k as total minus prefix and
 * pushes it, with dC/dc_j = F(L_j) - F(L_{j-1}), through the densitymap,
 * the colour mean and each covering particle's RBF weight, radius,
 * colormap and radiusmap lookups.
 *
 * Estimator 1, ratio tracking: the expectation of the reference's delta
 * tracking (deviceCodeRBF.hlsl:141-203, primary ray, ambient 1) over its
 * accept draws.  Tentative points follow the tracker's free flights,
 * t_i = t_{i-1} - ln(1 - xi_i) unit from t = 0 (one LCG draw each, after
 * the sample's three set-up draws), and point i is accepted with
 * probability D_i, so
 *     C = sum_i T_{i-1} D_i c_i,   T_i = T_{i-1} (1 - D_i),
 * the sum ending after the first point with T_i < 2^-20 (the remaining
 * terms are below that fraction of the colour).  The adjoint here stores
 * the sample's points and runs the reverse recursion
 *     R_J = 0,  R_{j-1} = D_j c_j + (1 - D_j) R_j,
 *     dC/dD_j = T_{j-1} (c_j - R_j),  dC/dc_j = T_{j-1} D_j.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/aarbf_detmath.h"

#define ORC_EXPORT __attribute__((visibility("default")))

typedef struct orc_diff {
  const double *pos;   /* 3N */
  const double *attr;  /* N */
  const double *radii; /* N */
  uint32_t n;
  const double *cm, *dm, *rm; /* RGBA texels (dm, rm: channel 0 used) */
  uint32_t cm_n, dm_n, rm_n;
  float cam_pos[3], cam_d00[3], cam_du[3], cam_dv[3];
  int32_t width, height;
  uint32_t spp, frameID, max_depth;
  float unit, rbfRadius;
  float aabb[6];
  uint32_t estimator; /* 0 expected marching, 1 ratio tracking */
} orc_diff;

#define ORC_RT_TMIN (1.0 / 1048576.0)

/* linear texel filter with clamp-to-edge (aa_tf_coords in double): value,
 * slope d/du, and the two texels with their weights */
static double tfd(const double *t, uint32_t n, int c, double u, double *slope, uint32_t *j0, uint32_t *j1,
                  double *f) {
  double x = u * (double)n - 0.5;
  int clamped = 0;
  if (!(x > -1.0)) x = -1.0;
  if (x > (double)n) x = (double)n;
  double fl = floor(x);
  int i0 = (int)fl;
  *f = x - fl;
  int a = i0 < 0 ? 0 : (i0 > (int)n - 1 ? (int)n - 1 : i0);
  int b = i0 + 1 < 0 ? 0 : (i0 + 1 > (int)n - 1 ? (int)n - 1 : i0 + 1);
  if (a == b) clamped = 1;
  *j0 = (uint32_t)a;
  *j1 = (uint32_t)b;
  double ta = t[4 * a + c], tb = t[4 * b + c];
  *slope = clamped ? 0.0 : (double)n * (tb - ta);
  return ta + *f * (tb - ta);
}

static double Fcdf(double L, double *dF) {
  double v = (1.0 - exp(-L)) / 0.99;
  if (v >= 1.0) {
    *dF = 0.0;
    return 1.0;
  }
  *dF = exp(-L) / 0.99;
  return v;
}

typedef struct {
  double alpha, a3, c[3], c22[3], r, rm;
} dpart;

static void part_values(const orc_diff *p, uint32_t i, dpart *q) {
  double s, f;
  uint32_t j0, j1;
  for (int c = 0; c < 3; ++c) q->c[c] = tfd(p->cm, p->cm_n, c, p->attr[i], &s, &j0, &j1, &f);
  q->alpha = tfd(p->cm, p->cm_n, 3, p->attr[i], &s, &j0, &j1, &f);
  q->a3 = q->alpha * q->alpha * q->alpha;
  for (int c = 0; c < 3; ++c) q->c22[c] = q->c[c] > 0.0 ? pow(q->c[c], 2.2) : 0.0;
  q->rm = tfd(p->rm, p->rm_n, 0, p->attr[i], &s, &j0, &j1, &f);
  q->r = q->alpha == 0.0 ? 0.0 : p->radii[i] * q->rm;
}

/* one march point: density, colour sums and count at x */
typedef struct {
  uint32_t count;
  double rho, S[3];
} dpay;

static void query(const orc_diff *p, const dpart *pp, const double x[3], dpay *pl) {
  pl->count = 0;
  pl->rho = pl->S[0] = pl->S[1] = pl->S[2] = 0.0;
  for (uint32_t i = 0; i < p->n; ++i) {
    if (!(pp[i].r > 0.0)) continue;
    double dx = p->pos[3 * i] - x[0], dy = p->pos[3 * i + 1] - x[1], dz = p->pos[3 * i + 2] - x[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(d < pp[i].r)) continue;
    double q = 3.0 * d / pp[i].r;
    double w = exp(-0.5 * q * q) * pp[i].a3;
    pl->count += 1;
    pl->rho += w;
    for (int c = 0; c < 3; ++c) pl->S[c] += pp[i].c22[c] * w;
  }
}

static void colour(const dpay *pl, double c[3]) {
  for (int k = 0; k < 3; ++k) c[k] = pow(pl->S[k] / pl->rho, 1.0 / 2.2);
}

typedef struct {
  float o[3], d[3];
  float t, tmax;
  int live;
} dray;

static void sample_ray(const orc_diff *p, int px, int py, uint32_t s, dray *r) {
  uint32_t fid = (p->frameID + s) & 0xffffu;
  uint32_t rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)p->width);
  float r1 = aa_lcg_next(&rng);
  float sx = ((float)px + 0.5f) / (float)p->width, sy = ((float)py + 0.5f) / (float)p->height;
  float lb[3], rt[3], d[3];
  for (int k = 0; k < 3; ++k) {
    rt[k] = p->aabb[3 + k] + 2.0f * p->rbfRadius;
    lb[k] = p->aabb[k] - 2.0f * p->rbfRadius;
    d[k] = (p->cam_d00[k] + sx * p->cam_du[k]) + sy * p->cam_dv[k];
  }
  float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  for (int k = 0; k < 3; ++k) d[k] = d[k] / len;
  float inv[3], t1[3], t2[3];
  for (int k = 0; k < 3; ++k) {
    inv[k] = 1.0f / d[k];
    t2[k] = (rt[k] - p->cam_pos[k]) * inv[k];
    t1[k] = (lb[k] - p->cam_pos[k]) * inv[k];
  }
  float te = fmaxf(fmaxf(fminf(t1[0], t2[0]), fminf(t1[1], t2[1])), fminf(t1[2], t2[2]));
  float tx = fminf(fminf(fmaxf(t1[0], t2[0]), fmaxf(t1[1], t2[1])), fmaxf(t1[2], t2[2]));
  te = fmaxf(te, 0.0f);
  tx = fminf(tx, 10000.0f);
  for (int k = 0; k < 3; ++k) {
    r->o[k] = p->cam_pos[k];
    r->d[k] = d[k];
  }
  r->live = te < tx;
  r->t = te + p->unit * r1;
  r->tmax = tx;
}

/* visits the march points that some particle covers: fn(k-th point, x, payload) */
typedef int (*point_fn)(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, void *u);

/* estimator 0: the march points; estimator 1: the tracker's tentative
 * points.  fn returns 0 to end the sample. */
static void march(const orc_diff *p, const dpart *pp, int px, int py, uint32_t s, point_fn fn, void *u) {
  dray r;
  sample_ray(p, px, py, s, &r);
  if (!r.live) return;
  float t = r.t;
  uint32_t rng = 0;
  if (p->estimator == 1) {
    uint32_t fid = (p->frameID + s) & 0xffffu;
    rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)p->width);
    for (int k = 0; k < 3; ++k) aa_lcg_next(&rng); /* r0, r1, r2 of the sample's set-up */
    t = 0.0f;
  }
  uint32_t maxd = p->max_depth ? p->max_depth : 10000u;
  for (uint32_t i = 0; i < maxd; ++i) {
    if (p->estimator == 1) {
      float xi = aa_lcg_next(&rng);
      t = aa_flight(t, 1.0f - xi, p->unit);
    } else {
      t = t + p->unit;
    }
    if (t >= r.tmax) return;
    float xf[3] = {r.o[0] + t * r.d[0], r.o[1] + t * r.d[1], r.o[2] + t * r.d[2]};
    double x[3] = {xf[0], xf[1], xf[2]};
    dpay pl;
    query(p, pp, x, &pl);
    if (pl.count > 0 && pl.rho > 0.0)
      if (!fn(p, pp, x, &pl, u)) return;
  }
}

typedef struct {
  double L, C[3], T;
} fwd_state;

static int fwd_point(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, void *u) {
  (void)pp;
  (void)x;
  fwd_state *st = (fwd_state *)u;
  double s, f, dF;
  uint32_t j0, j1;
  double D = tfd(p->dm, p->dm_n, 0, pl->rho, &s, &j0, &j1, &f);
  double c[3];
  colour(pl, c);
  if (p->estimator == 1) {
    for (int k = 0; k < 3; ++k) st->C[k] += st->T * D * c[k];
    st->T *= 1.0 - D;
    return st->T >= ORC_RT_TMIN;
  }
  double F0 = Fcdf(st->L, &dF);
  st->L += D;
  double F1 = Fcdf(st->L, &dF);
  for (int k = 0; k < 3; ++k) st->C[k] += (F1 - F0) * c[k];
  return 1;
}

static dpart *parts(const orc_diff *p) {
  dpart *pp = (dpart *)malloc(sizeof(dpart) * (p->n ? p->n : 1));
  if (pp)
    for (uint32_t i = 0; i < p->n; ++i) part_values(p, i, pp + i);
  return pp;
}

static void pixel_forward(const orc_diff *p, const dpart *pp, int px, int py, double out[3]) {
  double acc[3] = {0, 0, 0};
  for (uint32_t s = 0; s < p->spp; ++s) {
    fwd_state st = {0.0, {0, 0, 0}, 1.0};
    march(p, pp, px, py, s, fwd_point, &st);
    for (int k = 0; k < 3; ++k) acc[k] += st.C[k];
  }
  for (int k = 0; k < 3; ++k) out[k] = acc[k] / p->spp;
}

/* image: W*H*3 doubles (mean over spp samples) */
ORC_EXPORT int orc_diff_forward(const orc_diff *p, double *image) {
  dpart *pp = parts(p);
  if (!pp) return -1;
  for (int py = 0; py < p->height; ++py)
    for (int px = 0; px < p->width; ++px) pixel_forward(p, pp, px, py, image + 3 * ((size_t)py * p->width + px));
  free(pp);
  return 0;
}

/* The same for the listed (x, y) pixels of the W x H image only: out is
 * npix * 3 doubles. */
ORC_EXPORT int orc_diff_forward_pixels(const orc_diff *p, const int32_t *pixels, uint64_t npix, double *out) {
  dpart *pp = parts(p);
  if (!pp) return -1;
  for (uint64_t i = 0; i < npix; ++i) pixel_forward(p, pp, pixels[2 * i], pixels[2 * i + 1], out + 3 * i);
  free(pp);
  return 0;
}

/* ---- adjoint ---- */
typedef struct {
  const double *g; /* dLoss/dC of this sample (rgb, already / spp) */
  /* pass 1 */
  double L, prev_c[3], total;
  int have_prev;
  /* pass 2 */
  double prefix, Lp, cprev[3], aLprev;
  double *g_pos, *g_rad, *g_attr, *g_cm, *g_dm, *g_rm;
} adj_state;

static int pass1(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, void *u) {
  (void)pp;
  (void)x;
  adj_state *st = (adj_state *)u;
  double s, f, dF;
  uint32_t j0, j1;
  double D = tfd(p->dm, p->dm_n, 0, pl->rho, &s, &j0, &j1, &f);
  double c[3];
  colour(pl, c);
  if (st->have_prev) { /* dC/dL_{k-1} = F'(L_{k-1}) (c_{k-1} - c_k) */
    Fcdf(st->L, &dF);
    double gd = 0;
    for (int k = 0; k < 3; ++k) gd += st->g[k] * (st->prev_c[k] - c[k]);
    st->total += dF * gd;
  }
  st->L += D;
  memcpy(st->prev_c, c, sizeof(c));
  st->have_prev = 1;
  return 1;
}

/* pushes the adjoints of one point's density (aD = dC/dD) and colour
 * (pj = dC/dc, the same for the three channels) through the densitymap, the
 * colour mean and every covering particle */
static void point_adjoint(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, const double *g,
                          double aD, double pj, double *g_pos, double *g_rad, double *g_attr, double *g_cm,
                          double *g_dm, double *g_rm) {
  double sD, fD;
  uint32_t j0, j1;
  tfd(p->dm, p->dm_n, 0, pl->rho, &sD, &j0, &j1, &fD);
  g_dm[j0] += aD * (1.0 - fD);
  g_dm[j1] += aD * fD;
  double a_rho = aD * sD, a_S[3];
  /* colour mean c = (S / rho)^(1/2.2) */
  for (int k = 0; k < 3; ++k) {
    double gc = g[k] * pj;
    double m = pl->S[k] / pl->rho;
    double dcdm = m > 0.0 ? (1.0 / 2.2) * pow(m, 1.0 / 2.2 - 1.0) : 0.0;
    a_S[k] = gc * dcdm / pl->rho;
    a_rho -= gc * dcdm * m / pl->rho;
  }
  /* covering particles */
  for (uint32_t i = 0; i < p->n; ++i) {
    const dpart *q = pp + i;
    if (!(q->r > 0.0)) continue;
    double dx = p->pos[3 * i] - x[0], dy = p->pos[3 * i + 1] - x[1], dz = p->pos[3 * i + 2] - x[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(d < q->r)) continue;
    double qq = 3.0 * d / q->r;
    double rbf = exp(-0.5 * qq * qq), w = rbf * q->a3;
    double a_w = a_rho;
    for (int k = 0; k < 3; ++k) a_w += a_S[k] * q->c22[k];
    double a_rbf = a_w * q->a3, a_a3 = a_w * rbf;
    double a_c22[3];
    for (int k = 0; k < 3; ++k) a_c22[k] = a_S[k] * w;
    /* rbf = exp(-4.5 d^2 / r^2) */
    double drbf_dd = -9.0 * d / (q->r * q->r) * rbf, drbf_dr = 9.0 * d * d / (q->r * q->r * q->r) * rbf;
    if (d > 0.0) {
      double k = a_rbf * drbf_dd / d;
      g_pos[3 * i] += k * dx;
      g_pos[3 * i + 1] += k * dy;
      g_pos[3 * i + 2] += k * dz;
    }
    double a_r = a_rbf * drbf_dr;
    /* r = radius * radiusmap(attr) */
    double s, f;
    uint32_t k0, k1;
    tfd(p->rm, p->rm_n, 0, p->attr[i], &s, &k0, &k1, &f);
    g_rad[i] += a_r * q->rm;
    double a_rm = a_r * p->radii[i];
    g_rm[k0] += a_rm * (1.0 - f);
    g_rm[k1] += a_rm * f;
    g_attr[i] += a_rm * s;
    /* a3 = alpha^3, c22 = c^2.2, from colormap(attr) */
    double a_alpha = a_a3 * 3.0 * q->alpha * q->alpha;
    tfd(p->cm, p->cm_n, 3, p->attr[i], &s, &k0, &k1, &f);
    g_cm[4 * k0 + 3] += a_alpha * (1.0 - f);
    g_cm[4 * k1 + 3] += a_alpha * f;
    g_attr[i] += a_alpha * s;
    for (int c2 = 0; c2 < 3; ++c2) {
      double a_c = q->c[c2] > 0.0 ? a_c22[c2] * 2.2 * pow(q->c[c2], 1.2) : 0.0;
      tfd(p->cm, p->cm_n, c2, p->attr[i], &s, &k0, &k1, &f);
      g_cm[4 * k0 + c2] += a_c * (1.0 - f);
      g_cm[4 * k1 + c2] += a_c * f;
      g_attr[i] += a_c * s;
    }
  }
}

static int pass2(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, void *u) {
  adj_state *st = (adj_state *)u;
  double sD, fD, dF;
  uint32_t j0, j1;
  double D = tfd(p->dm, p->dm_n, 0, pl->rho, &sD, &j0, &j1, &fD);
  double c[3];
  colour(pl, c);
  /* the term dC/dL_{j-1} belongs to the prefix once c_j is known */
  if (st->have_prev) {
    Fcdf(st->Lp, &dF);
    double gd = 0;
    for (int k = 0; k < 3; ++k) gd += st->g[k] * (st->cprev[k] - c[k]);
    st->prefix += dF * gd;
  }
  double aD = st->total - st->prefix; /* dC/dD_j */
  double F0 = Fcdf(st->Lp, &dF);
  st->Lp += D;
  double F1 = Fcdf(st->Lp, &dF);
  point_adjoint(p, pp, x, pl, st->g, aD, F1 - F0, st->g_pos, st->g_rad, st->g_attr, st->g_cm, st->g_dm, st->g_rm);
  memcpy(st->cprev, c, sizeof(c));
  st->have_prev = 1;
  return 1;
}

/* ratio tracking: the sample's points, kept for the reverse recursion */
typedef struct {
  double x[3], D, c[3];
  dpay pl;
} rt_point;

typedef struct {
  rt_point *pts;
  size_t n, cap;
  double T;
} rt_state;

static int rt_keep(const orc_diff *p, const dpart *pp, const double x[3], const dpay *pl, void *u) {
  (void)pp;
  rt_state *st = (rt_state *)u;
  if (st->n == st->cap) {
    size_t cap = st->cap ? 2 * st->cap : 256;
    rt_point *q = (rt_point *)realloc(st->pts, cap * sizeof(rt_point));
    if (!q) return 0;
    st->pts = q;
    st->cap = cap;
  }
  rt_point *e = st->pts + st->n++;
  double s, f;
  uint32_t j0, j1;
  memcpy(e->x, x, sizeof(e->x));
  e->D = tfd(p->dm, p->dm_n, 0, pl->rho, &s, &j0, &j1, &f);
  colour(pl, e->c);
  e->pl = *pl;
  st->T *= 1.0 - e->D;
  return st->T >= ORC_RT_TMIN;
}

static void rt_adjoint(const orc_diff *p, const dpart *pp, const rt_state *st, const double *g, double *g_pos,
                       double *g_rad, double *g_attr, double *g_cm, double *g_dm, double *g_rm) {
  double *R = (double *)malloc(sizeof(double) * 3 * (st->n + 1));
  if (!R) return;
  R[3 * st->n] = R[3 * st->n + 1] = R[3 * st->n + 2] = 0.0;
  for (size_t j = st->n; j-- > 0;) /* R[j] = remaining colour after point j (index j: point j+1) */
    for (int k = 0; k < 3; ++k)
      R[3 * j + k] = j + 1 < st->n ? st->pts[j + 1].D * st->pts[j + 1].c[k] + (1.0 - st->pts[j + 1].D) * R[3 * (j + 1) + k]
                                   : 0.0;
  double T = 1.0;
  for (size_t j = 0; j < st->n; ++j) {
    const rt_point *e = st->pts + j;
    double aD = 0.0;
    for (int k = 0; k < 3; ++k) aD += g[k] * T * (e->c[k] - R[3 * j + k]);
    point_adjoint(p, pp, e->x, &e->pl, g, aD, T * e->D, g_pos, g_rad, g_attr, g_cm, g_dm, g_rm);
    T *= 1.0 - e->D;
  }
  free(R);
}

/* adds the adjoint of one pixel (g: dLoss/dpixel, rgb) */
static void pixel_adjoint(const orc_diff *p, const dpart *pp, int px, int py, const double *gpix, double *g_pos,
                          double *g_rad, double *g_attr, double *g_cm, double *g_dm, double *g_rm) {
  double g[3];
  for (int k = 0; k < 3; ++k) g[k] = gpix[k] / p->spp;
  for (uint32_t s = 0; s < p->spp; ++s) {
    if (p->estimator == 1) {
      rt_state rs = {NULL, 0, 0, 1.0};
      march(p, pp, px, py, s, rt_keep, &rs);
      rt_adjoint(p, pp, &rs, g, g_pos, g_rad, g_attr, g_cm, g_dm, g_rm);
      free(rs.pts);
      continue;
    }
    adj_state st;
    memset(&st, 0, sizeof(st));
    st.g = g;
    march(p, pp, px, py, s, pass1, &st);
    if (st.have_prev) { /* the last point: c_{K+1} = 0 */
      double dF;
      Fcdf(st.L, &dF);
      double gd = 0;
      for (int k = 0; k < 3; ++k) gd += g[k] * st.prev_c[k];
      st.total += dF * gd;
    }
    st.have_prev = 0;
    st.g_pos = g_pos;
    st.g_rad = g_rad;
    st.g_attr = g_attr;
    st.g_cm = g_cm;
    st.g_dm = g_dm;
    st.g_rm = g_rm;
    march(p, pp, px, py, s, pass2, &st);
  }
}

/* gimage: dLoss/dimage, W*H*3.  Gradients are accumulated (callers zero
 * them): g_pos 3N, g_rad N, g_attr N, g_cm 4*cm_n, g_dm dm_n, g_rm rm_n. */
ORC_EXPORT int orc_diff_adjoint(const orc_diff *p, const double *gimage, double *g_pos, double *g_rad, double *g_attr,
                                double *g_cm, double *g_dm, double *g_rm) {
  dpart *pp = parts(p);
  if (!pp) return -1;
  for (int py = 0; py < p->height; ++py)
    for (int px = 0; px < p->width; ++px)
      pixel_adjoint(p, pp, px, py, gimage + 3 * ((size_t)py * p->width + px), g_pos, g_rad, g_attr, g_cm, g_dm,
                    g_rm);
  free(pp);
  return 0;
}

/* The adjoint of the listed (x, y) pixels only; gpix: npix * 3 doubles. */
ORC_EXPORT int orc_diff_adjoint_pixels(const orc_diff *p, const int32_t *pixels, uint64_t npix, const double *gpix,
                                       double *g_pos, double *g_rad, double *g_attr, double *g_cm, double *g_dm,
                                       double *g_rm) {
  dpart *pp = parts(p);
  if (!pp) return -1;
  for (uint64_t i = 0; i < npix; ++i)
    pixel_adjoint(p, pp, pixels[2 * i], pixels[2 * i + 1], gpix + 3 * i, g_pos, g_rad, g_attr, g_cm, g_dm, g_rm);
  free(pp);
  return 0;
}
