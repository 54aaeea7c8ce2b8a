/*
 * sop_oracle.c -- CPU oracle for the SOP overdamped force / list / integrate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for orc_port_*, as the timed CPU baseline).  The product (libsop) never
 * links, loads or calls it.
 *
 * It restates, in plain C, the algorithm of the reference (samchogroup/SOP):
 *
 *   - FP32 list predicates, bit-exact with the reference's single-precision
 *     arithmetic: minimum image  d -= boxl*rint(d/boxl)  (energy.cu:248-250,
 *     sop.cu:739-741, rintf = round-half-even), d2 = dx*dx+dy*dy+dz*dz
 *     (sop.cu:744), cutoff  rc = (float)(3.2 * (double)r0)  with the double
 *     literal rounded once to float (sop.cu:746,815; 2.5 at sop.cu:861,910),
 *     rc2 = rc*rc in float, keep iff d2 < rc2 (sop.cu:749,818).  No FMA
 *     contraction: this file is compiled with -ffp-contract=off and the device
 *     side follows the same rule.
 *   - The master non-native list is every pair with |i-j| > 2 inside a chain
 *     (or on different chains) that is not native; a pair is native iff it is
 *     in the native table (io.cpp:190-226, native iff r_ij < rcut_nat).
 *   - Neighbor list = master pairs with d2 < rc_nl^2 (sop.cu:1354-1448);
 *     pair list = neighbor-list pairs with d2 < rc_pl^2 (sop.cu:1710-1823).
 *   - Energies and forces in FP64 from the FP32 coordinates:
 *       FENE   E = -(k R0^2/2) ln(1 - dev^2/R0^2)       energy.cu:194-225,228-273
 *       i,i+2  E = eps_ss sigma_ss^6 / r^6              energy.cu:278-307,310-362
 *       native E = eps (r0/r)^12 - 2 eps (r0/r)^6       energy.cu:381-410,445-509
 *       non-native E = eps (c12 (s/r)^12 + c6 (s/r)^6) energy.cu:412-438,512-584
 *     with the reference's force guards d2 > 1e-7 r0^2 (energy.cu:481) and
 *     d2 > 1e-7 sigma (energy.cu:555-558).
 *   - Overdamped step  x += a5 (F + var xi),  a5 = h/zeta, var = sqrt(2 T zeta/h)
 *     (sop.cu:487-520, energy.cu:156-167), with xi drawn from Philox4x32-10
 *     (Salmon et al., SC'11) and the Box-Muller transform; the cuRAND stream of
 *     the reference is closed code and is not reproduced.
 *   - Observables: Q (d/r0 < 1.25, sop.cu:538-556) and Rg^2 as the O(N^2) pair
 *     sum / N^2 (sop.cu:564-580).
 *
 * Schedule restated here (the reference's overdamped_ctrl, sop.cu:287-417,
 * with its global inlup = 0): both lists are built before the first step
 * (sop.cu:318-331); before the step whose counter is s (completed steps) the
 * neighbor list is rebuilt if s % n_nl == 0, then the pair list if
 * s % n_pl == 0; then forces are evaluated at the current coordinates and the
 * beads move.
 *
 * Parity pins for every function live in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

typedef struct {
  int32_t n;
  const int32_t *chain;   /* chain id per bead */
  const int32_t *type;    /* bead type 1 or 2 */
  const float *bond_r0;   /* r0 of bond (i,i+1), used iff chain[i]==chain[i+1] */
  int32_t n_nat;
  const int32_t *nat_i, *nat_j; /* native pairs, i<j, 0-based */
  const float *nat_r0;
  float k_bnd, R0, eps_ss, sigma_ss;
  float eps_att[9], eps_rep[9], sigma_rep[9]; /* [ti*3+tj], types 1..2 */
  float rep_c12, rep_c6;
  double nl_factor; float nl_abs;
  double pl_factor; float pl_abs;
  float boxl;
} orc_sys;

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random   */
/* numbers: as easy as 1, 2, 3").  Multipliers and Weyl constants are   */
/* the published ones.                                                  */
/* ------------------------------------------------------------------ */
ORC_EXPORT void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2],
                           uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 24-bit uniform in (0,1): odd multiples of 2^-24, exactly representable
 * in float, never 0 or 1. */
static double u24(uint32_t r) { return (double)((r >> 8) | 1u) * (1.0 / 16777216.0); }

/* Three standard normals for (seed, bead, step): counter = (bead, step_lo,
 * step_hi, 0), key = (seed_lo, seed_hi); Box-Muller on (u0,u1) and (u2,u3). */
ORC_EXPORT void orc_normals(uint64_t seed, uint32_t bead, uint64_t step, double out[3]) {
  uint32_t ctr[4] = {bead, (uint32_t)step, (uint32_t)(step >> 32), 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  orc_philox(ctr, key, r);
  const double two_pi = 6.283185307179586476925286766559;
  double ra = sqrt(-2.0 * log(u24(r[0])));
  double rb = sqrt(-2.0 * log(u24(r[2])));
  out[0] = ra * cos(two_pi * u24(r[1]));
  out[1] = ra * sin(two_pi * u24(r[1]));
  out[2] = rb * cos(two_pi * u24(r[3]));
}

/* ------------------------------------------------------------------ */
/* FP32 predicate pieces (bit-exact restatement)                        */
/* ------------------------------------------------------------------ */
ORC_EXPORT float orc_min_image_f(float d, float boxl) {
  if (boxl > 0.0f) {
    float q = d / boxl;
    float r = rintf(q);
    float p = boxl * r;
    d = d - p;
  }
  return d;
}

ORC_EXPORT float orc_cutoff(double factor, float absval, float base) {
  if (factor > 0.0) return (float)(factor * (double)base);
  return absval;
}

static inline float dist2_f(const float *x, int i, int j, float boxl) {
  float dx = x[3 * j] - x[3 * i];
  float dy = x[3 * j + 1] - x[3 * i + 1];
  float dz = x[3 * j + 2] - x[3 * i + 2];
  dx = orc_min_image_f(dx, boxl);
  dy = orc_min_image_f(dy, boxl);
  dz = orc_min_image_f(dz, boxl);
  float a = dx * dx;
  float b = dy * dy;
  float c = dz * dz;
  return (a + b) + c;
}

ORC_EXPORT float orc_dist2_f(const float *x, int32_t i, int32_t j, float boxl) {
  return dist2_f(x, i, j, boxl);
}

static inline int excluded(const orc_sys *s, int i, int j) {
  int d = i > j ? i - j : j - i;
  return s->chain[i] == s->chain[j] && d <= 2;
}

/* native lookup: per-bead sorted partner table built on demand */
typedef struct {
  int32_t *off; /* n+1 */
  int32_t *j;   /* 2*n_nat */
} nat_index;

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

static int nat_index_build(const orc_sys *s, nat_index *ix) {
  int n = s->n;
  ix->off = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  ix->j = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * s->n_nat + 1));
  if (!ix->off || !ix->j) return -1;
  for (int k = 0; k < s->n_nat; ++k) {
    ix->off[s->nat_i[k] + 1]++;
    ix->off[s->nat_j[k] + 1]++;
  }
  for (int i = 0; i < n; ++i) ix->off[i + 1] += ix->off[i];
  int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  if (!cur) return -1;
  memcpy(cur, ix->off, sizeof(int32_t) * (size_t)n);
  for (int k = 0; k < s->n_nat; ++k) {
    ix->j[cur[s->nat_i[k]]++] = s->nat_j[k];
    ix->j[cur[s->nat_j[k]]++] = s->nat_i[k];
  }
  free(cur);
  for (int i = 0; i < n; ++i)
    qsort(ix->j + ix->off[i], (size_t)(ix->off[i + 1] - ix->off[i]), sizeof(int32_t), cmp_i32);
  return 0;
}

static void nat_index_free(nat_index *ix) {
  free(ix->off);
  free(ix->j);
}

static int is_native(const nat_index *ix, int i, int j) {
  int lo = ix->off[i], hi = ix->off[i + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ix->j[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return lo < ix->off[i + 1] && ix->j[lo] == j;
}

/* Neighbor list by brute force over the implicit master list.  att: indices k
 * into the native table; rep: half pairs (i<j) in row-major order. Returns 0,
 * or -1 if a capacity is too small (counts are still reported). */
static int64_t nl_natives(const orc_sys *s, const float *x, int32_t *att_k, int64_t att_cap) {
  int64_t na = 0;
  for (int k = 0; k < s->n_nat; ++k) {
    int i = s->nat_i[k], j = s->nat_j[k];
    float rc = orc_cutoff(s->nl_factor, s->nl_abs, s->nat_r0[k]);
    float rc2 = rc * rc;
    if (dist2_f(x, i, j, s->boxl) < rc2) {
      if (na < att_cap) att_k[na] = k;
      na++;
    }
  }
  return na;
}

ORC_EXPORT int orc_nl(const orc_sys *s, const float *x, int32_t *att_k, int64_t att_cap,
                      int64_t *n_att, int32_t *rep_i, int32_t *rep_j, int64_t rep_cap,
                      int64_t *n_rep) {
  int64_t na = nl_natives(s, x, att_k, att_cap), nr = 0;
  nat_index ix;
  if (nat_index_build(s, &ix)) return -2;
  float rc2t[9];
  for (int a = 0; a < 9; ++a) {
    float rc = orc_cutoff(s->nl_factor, s->nl_abs, s->sigma_rep[a]);
    rc2t[a] = rc * rc;
  }
  for (int i = 0; i < s->n; ++i) {
    for (int j = i + 1; j < s->n; ++j) {
      if (excluded(s, i, j) || is_native(&ix, i, j)) continue;
      float d2 = dist2_f(x, i, j, s->boxl);
      if (d2 < rc2t[s->type[i] * 3 + s->type[j]]) {
        if (nr < rep_cap) { rep_i[nr] = i; rep_j[nr] = j; }
        nr++;
      }
    }
  }
  nat_index_free(&ix);
  *n_att = na;
  *n_rep = nr;
  return (na > att_cap || nr > rep_cap) ? -1 : 0;
}

/* The neighbor list of orc_nl for open boundaries (boxl == 0), with the
 * candidate pairs taken from a uniform grid instead of all n(n-1)/2 pairs, for
 * systems too large for the brute-force loop.  The output arrays are identical
 * to orc_nl's: the same FP32 predicate on the same pairs, rows in increasing
 * i, partners in increasing j.  The grid drops only pairs in cells that are
 * not adjacent along some axis.  Cells are computed in double from the
 * minimum corner with an edge of at least rc_max + 1 A, so such a pair has a
 * coordinate difference above rc_max in exact arithmetic; the FP32 difference
 * is then >= rc (rounding is monotone and rc is a float), its square >= rc*rc
 * as rounded, and adding non-negative terms keeps d2 >= rc2: the predicate
 * d2 < rc2 rejects it in orc_nl as well.  Returns -4 for a periodic box,
 * otherwise as orc_nl. */
ORC_EXPORT int orc_nl_grid(const orc_sys *s, const float *x, int32_t *att_k, int64_t att_cap,
                           int64_t *n_att, int32_t *rep_i, int32_t *rep_j, int64_t rep_cap,
                           int64_t *n_rep) {
  if (s->boxl > 0.0f) return -4;
  const int n = s->n;
  int64_t na = nl_natives(s, x, att_k, att_cap), nr = 0;
  float rc2t[9], rcmax = 0.0f;
  for (int a = 0; a < 9; ++a) {
    float rc = orc_cutoff(s->nl_factor, s->nl_abs, s->sigma_rep[a]);
    rc2t[a] = rc * rc;
    if (rc > rcmax) rcmax = rc;
  }
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      double v = (double)x[3 * i + d];
      if (i == 0 || v < lo[d]) lo[d] = v;
      if (i == 0 || v > hi[d]) hi[d] = v;
    }
  double edge = (double)rcmax + 1.0;
  int64_t dims[3], ncell;
  for (;;) {  /* wider cells (still valid) when the grid would outnumber the beads */
    ncell = 1;
    for (int d = 0; d < 3; ++d) {
      dims[d] = (int64_t)floor((hi[d] - lo[d]) / edge) + 1;
      ncell *= dims[d];
    }
    if (ncell <= 8 * (int64_t)n + 64) break;
    edge *= 1.25;
  }
  int32_t *cell = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  int32_t *start = (int32_t *)calloc((size_t)ncell + 1, sizeof(int32_t));
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  int64_t bcap = 1024;
  int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)bcap);
  nat_index ix;
  if (!cell || !start || !order || !buf || nat_index_build(s, &ix)) {
    free(cell); free(start); free(order); free(buf);
    return -2;
  }
  for (int i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int d = 2; d >= 0; --d) {
      int64_t q = (int64_t)floor(((double)x[3 * i + d] - lo[d]) / edge);
      if (q < 0) q = 0;
      if (q >= dims[d]) q = dims[d] - 1;
      c = c * dims[d] + q;
    }
    cell[i] = (int32_t)c;
    start[c + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncell + 1);
    if (!cur) { free(cell); free(start); free(order); free(buf); nat_index_free(&ix); return -2; }
    memcpy(cur, start, sizeof(int32_t) * (size_t)ncell);
    for (int i = 0; i < n; ++i) order[cur[cell[i]]++] = i;
    free(cur);
  }
  int status = 0;
  for (int i = 0; i < n && !status; ++i) {
    int64_t c = cell[i], q[3];
    q[0] = c % dims[0];
    q[1] = (c / dims[0]) % dims[1];
    q[2] = c / (dims[0] * dims[1]);
    int64_t m = 0;
    for (int64_t z = q[2] - 1; z <= q[2] + 1; ++z) {
      if (z < 0 || z >= dims[2]) continue;
      for (int64_t y = q[1] - 1; y <= q[1] + 1; ++y) {
        if (y < 0 || y >= dims[1]) continue;
        for (int64_t w = q[0] - 1; w <= q[0] + 1; ++w) {
          if (w < 0 || w >= dims[0]) continue;
          int64_t cc = (z * dims[1] + y) * dims[0] + w;
          for (int32_t t = start[cc]; t < start[cc + 1]; ++t) {
            int j = order[t];
            if (j <= i || !(dist2_f(x, i, j, s->boxl) < rc2t[s->type[i] * 3 + s->type[j]])) continue;
            if (!excluded(s, i, j) && !is_native(&ix, i, j)) {
              if (m == bcap) {
                int32_t *nb = (int32_t *)realloc(buf, sizeof(int32_t) * (size_t)(2 * bcap));
                if (!nb) { status = -2; break; }
                buf = nb;
                bcap *= 2;
              }
              buf[m++] = j;
            }
          }
          if (status) break;
        }
        if (status) break;
      }
      if (status) break;
    }
    if (status) break;
    qsort(buf, (size_t)m, sizeof(int32_t), cmp_i32);
    for (int64_t k = 0; k < m; ++k) {
      if (nr < rep_cap) { rep_i[nr] = i; rep_j[nr] = buf[k]; }
      nr++;
    }
  }
  free(cell); free(start); free(order); free(buf);
  nat_index_free(&ix);
  if (status) return status;
  *n_att = na;
  *n_rep = nr;
  return (na > att_cap || nr > rep_cap) ? -1 : 0;
}

/* Pair list: filter a neighbor list with the pair-list cutoff. */
ORC_EXPORT int orc_pl(const orc_sys *s, const float *x, const int32_t *nl_att_k, int64_t n_nl_att,
                      const int32_t *nl_rep_i, const int32_t *nl_rep_j, int64_t n_nl_rep,
                      int32_t *att_k, int64_t *n_att, int32_t *rep_i, int32_t *rep_j,
                      int64_t *n_rep) {
  int64_t na = 0, nr = 0;
  for (int64_t m = 0; m < n_nl_att; ++m) {
    int k = nl_att_k[m];
    float rc = orc_cutoff(s->pl_factor, s->pl_abs, s->nat_r0[k]);
    float rc2 = rc * rc;
    if (dist2_f(x, s->nat_i[k], s->nat_j[k], s->boxl) < rc2) att_k[na++] = k;
  }
  float rc2t[9];
  for (int a = 0; a < 9; ++a) {
    float rc = orc_cutoff(s->pl_factor, s->pl_abs, s->sigma_rep[a]);
    rc2t[a] = rc * rc;
  }
  for (int64_t m = 0; m < n_nl_rep; ++m) {
    int i = nl_rep_i[m], j = nl_rep_j[m];
    if (dist2_f(x, i, j, s->boxl) < rc2t[s->type[i] * 3 + s->type[j]]) {
      rep_i[nr] = i;
      rep_j[nr] = j;
      nr++;
    }
  }
  *n_att = na;
  *n_rep = nr;
  return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 energies and forces                                             */
/* ------------------------------------------------------------------ */
static inline double mi_d(double d, double boxl) {
  if (boxl > 0.0) d -= boxl * rint(d / boxl);
  return d;
}

/* e[0]=bond e[1]=i,i+2 e[2]=native e[3]=non-native e[4]=total e[5]=virial.
 * Forces accumulate into f (3n, zeroed here).  Returns 0, or -3 if a bond
 * is stretched to |dev| >= R0 (the reference has no guard: energy.cu:220). */
ORC_EXPORT int orc_forces(const orc_sys *s, const double *x, const int32_t *pl_att_k,
                          int64_t n_pl_att, const int32_t *pl_rep_i, const int32_t *pl_rep_j,
                          int64_t n_pl_rep, double *f, double *e) {
  const double boxl = s->boxl;
  double eb = 0, ea = 0, enat = 0, erep = 0, vir = 0;
  int status = 0;
  memset(f, 0, sizeof(double) * 3 * (size_t)s->n);
  const double k = s->k_bnd, R0 = s->R0, R0sq = R0 * R0;
  for (int i = 0; i + 1 < s->n; ++i) {
    if (s->chain[i] != s->chain[i + 1]) continue;
    int j = i + 1;
    double dx = mi_d(x[3 * j] - x[3 * i], boxl);
    double dy = mi_d(x[3 * j + 1] - x[3 * i + 1], boxl);
    double dz = mi_d(x[3 * j + 2] - x[3 * i + 2], boxl);
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    double dev = d - (double)s->bond_r0[i];
    double g = 1.0 - dev * dev / R0sq;
    if (g <= 0.0) status = -3;
    eb += -0.5 * k * R0sq * log(g);
    double c = -k * dev / d / g; /* force on j = c * (xj - xi) */
    f[3 * j] += c * dx; f[3 * j + 1] += c * dy; f[3 * j + 2] += c * dz;
    f[3 * i] -= c * dx; f[3 * i + 1] -= c * dy; f[3 * i + 2] -= c * dz;
    vir += c * (dx * dx + dy * dy + dz * dz);
  }
  const double s6 = pow((double)s->sigma_ss, 6.0);
  for (int i = 0; i + 2 < s->n; ++i) {
    if (s->chain[i] != s->chain[i + 2]) continue;
    int j = i + 2;
    double dx = mi_d(x[3 * j] - x[3 * i], boxl);
    double dy = mi_d(x[3 * j + 1] - x[3 * i + 1], boxl);
    double dz = mi_d(x[3 * j + 2] - x[3 * i + 2], boxl);
    double d2 = dx * dx + dy * dy + dz * dz;
    double inv6 = 1.0 / (d2 * d2 * d2);
    ea += s->eps_ss * s6 * inv6;
    double c = 6.0 * s->eps_ss * s6 * inv6 / d2;
    f[3 * j] += c * dx; f[3 * j + 1] += c * dy; f[3 * j + 2] += c * dz;
    f[3 * i] -= c * dx; f[3 * i + 1] -= c * dy; f[3 * i + 2] -= c * dz;
    vir += c * d2;
  }
  for (int64_t m = 0; m < n_pl_att; ++m) {
    int kk = pl_att_k[m];
    int i = s->nat_i[kk], j = s->nat_j[kk];
    double dx = mi_d(x[3 * j] - x[3 * i], boxl);
    double dy = mi_d(x[3 * j + 1] - x[3 * i + 1], boxl);
    double dz = mi_d(x[3 * j + 2] - x[3 * i + 2], boxl);
    double d2 = dx * dx + dy * dy + dz * dz;
    double r0 = s->nat_r0[kk], r02 = r0 * r0;
    double eps = s->eps_att[s->type[i] * 3 + s->type[j]];
    double q6 = (r02 * r02 * r02) / (d2 * d2 * d2);
    double q12 = q6 * q6;
    enat += eps * (q12 - 2.0 * q6);
    if (d2 > 1.0e-7 * r02) {
      double c = 12.0 * eps * (q12 - q6) / d2;
      f[3 * j] += c * dx; f[3 * j + 1] += c * dy; f[3 * j + 2] += c * dz;
      f[3 * i] -= c * dx; f[3 * i + 1] -= c * dy; f[3 * i + 2] -= c * dz;
      vir += c * d2;
    }
  }
  for (int64_t m = 0; m < n_pl_rep; ++m) {
    int i = pl_rep_i[m], j = pl_rep_j[m];
    double dx = mi_d(x[3 * j] - x[3 * i], boxl);
    double dy = mi_d(x[3 * j + 1] - x[3 * i + 1], boxl);
    double dz = mi_d(x[3 * j + 2] - x[3 * i + 2], boxl);
    double d2 = dx * dx + dy * dy + dz * dz;
    int t = s->type[i] * 3 + s->type[j];
    double sg = s->sigma_rep[t], sg2 = sg * sg;
    double eps = s->eps_rep[t];
    double q6 = (sg2 * sg2 * sg2) / (d2 * d2 * d2);
    double q12 = q6 * q6;
    erep += eps * (s->rep_c12 * q12 + s->rep_c6 * q6);
    if (d2 > 1.0e-7 * sg) {
      double c = 6.0 * eps * (2.0 * s->rep_c12 * q12 + s->rep_c6 * q6) / d2;
      f[3 * j] += c * dx; f[3 * j + 1] += c * dy; f[3 * j + 2] += c * dz;
      f[3 * i] -= c * dx; f[3 * i + 1] -= c * dy; f[3 * i + 2] -= c * dz;
      vir += c * d2;
    }
  }
  e[0] = eb; e[1] = ea; e[2] = enat; e[3] = erep;
  e[4] = eb + ea + enat + erep;
  e[5] = vir;
  return status;
}

/* Q and native contact count with the reference's FP32 test d/r0 < 1.25
 * (sop.cu:538-556); Rg^2 as the FP64 pair sum over i<j divided by N^2
 * (sop.cu:564-580). */
ORC_EXPORT int orc_observables(const orc_sys *s, const float *xf, const double *xd, double *Q,
                               int32_t *contacts, double *rgsq) {
  int32_t c = 0;
  for (int k = 0; k < s->n_nat; ++k) {
    float d2 = dist2_f(xf, s->nat_j[k], s->nat_i[k], s->boxl);
    float d = sqrtf(d2);
    if (d / s->nat_r0[k] < 1.25f) c++;
  }
  *contacts = c;
  *Q = s->n_nat > 0 ? (double)c / (double)s->n_nat : 0.0;
  double acc = 0.0;
  for (int i = 0; i < s->n - 1; ++i)
    for (int j = i + 1; j < s->n; ++j) {
      double dx = mi_d(xd[3 * i] - xd[3 * j], s->boxl);
      double dy = mi_d(xd[3 * i + 1] - xd[3 * j + 1], s->boxl);
      double dz = mi_d(xd[3 * i + 2] - xd[3 * j + 2], s->boxl);
      acc += dx * dx + dy * dy + dz * dz;
    }
  *rgsq = acc / ((double)s->n * (double)s->n);
  return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 overdamped trajectory                                           */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t *nla_k, *nlr_i, *nlr_j, *pla_k, *plr_i, *plr_j;
  int64_t nla, nlr, pla, plr, cap_rep;
} orc_lists;

/* cap_rep: pair capacity of the repulsive lists (grown by nl_build). */
static int lists_alloc(const orc_sys *s, orc_lists *L, int64_t cap_rep) {
  L->cap_rep = cap_rep;
  L->nla_k = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n_nat + 1));
  L->pla_k = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n_nat + 1));
  L->nlr_i = (int32_t *)malloc(sizeof(int32_t) * (size_t)L->cap_rep);
  L->nlr_j = (int32_t *)malloc(sizeof(int32_t) * (size_t)L->cap_rep);
  L->plr_i = (int32_t *)malloc(sizeof(int32_t) * (size_t)L->cap_rep);
  L->plr_j = (int32_t *)malloc(sizeof(int32_t) * (size_t)L->cap_rep);
  L->nla = L->nlr = L->pla = L->plr = 0;
  return (L->nla_k && L->pla_k && L->nlr_i && L->nlr_j && L->plr_i && L->plr_j) ? 0 : -2;
}

static void lists_free(orc_lists *L) {
  free(L->nla_k); free(L->pla_k); free(L->nlr_i); free(L->nlr_j); free(L->plr_i); free(L->plr_j);
}

/* Brute-force lists get capacity for every pair up front; grid lists start
 * at 192 pairs per bead and grow to the count the last build reported. */
static int64_t lists_cap(const orc_sys *s, int grid) {
  int64_t n = s->n;
  return grid ? 192 * n + 1024 : n * (n - 1) / 2 + 1;
}

static int nl_build(const orc_sys *s, const float *xf, orc_lists *L, int grid) {
  for (;;) {
    int st = grid ? orc_nl_grid(s, xf, L->nla_k, s->n_nat + 1, &L->nla, L->nlr_i, L->nlr_j, L->cap_rep, &L->nlr)
                  : orc_nl(s, xf, L->nla_k, s->n_nat + 1, &L->nla, L->nlr_i, L->nlr_j, L->cap_rep, &L->nlr);
    if (st != -1 || L->nlr <= L->cap_rep) return st;
    int64_t cap = L->nlr + L->nlr / 4 + 1024;
    int32_t *a = (int32_t *)realloc(L->nlr_i, sizeof(int32_t) * (size_t)cap);
    if (a) L->nlr_i = a;
    int32_t *b = (int32_t *)realloc(L->nlr_j, sizeof(int32_t) * (size_t)cap);
    if (b) L->nlr_j = b;
    int32_t *c = (int32_t *)realloc(L->plr_i, sizeof(int32_t) * (size_t)cap);
    if (c) L->plr_i = c;
    int32_t *d = (int32_t *)realloc(L->plr_j, sizeof(int32_t) * (size_t)cap);
    if (d) L->plr_j = d;
    if (!a || !b || !c || !d) return -2;
    L->cap_rep = cap;
  }
}

/* samples (nup > 0): after each step that completes a multiple of nup steps
 * (step counter (step0 + it + 1) % nup == 0), 4 doubles: the FP64 total
 * energy over the current pair list at the new coordinates (what
 * energy_eval reports between steps, energy.cu:69-78), Rg^2, Q and the
 * native contact count (calculate_observables, sop.cu:528-611). */
static int run_od(const orc_sys *s, const float *x0, int64_t step0, int64_t nsteps, uint64_t seed, float a5,
                  float var, int32_t n_nl, int32_t n_pl, double *x_out, int grid, int32_t nup,
                  double *samples) {
  int n = s->n;
  double *x = x_out;
  float *xf = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  orc_lists L;
  int st = lists_alloc(s, &L, lists_cap(s, grid));
  if (!xf || !f || st) { free(xf); free(f); lists_free(&L); return -2; }
  for (int i = 0; i < 3 * n; ++i) x[i] = (double)x0[i];
  const double A5 = a5, VAR = var;
  for (int64_t it = 0; it < nsteps; ++it) {
    int64_t step = step0 + it;
    for (int i = 0; i < 3 * n; ++i) xf[i] = (float)x[i];
    if (it == 0 || step % n_nl == 0) {
      st = nl_build(s, xf, &L, grid);
      if (st) break;
    }
    if (it == 0 || step % n_pl == 0)
      orc_pl(s, xf, L.nla_k, L.nla, L.nlr_i, L.nlr_j, L.nlr, L.pla_k, &L.pla, L.plr_i, L.plr_j, &L.plr);
    double e[6];
    st = orc_forces(s, x, L.pla_k, L.pla, L.plr_i, L.plr_j, L.plr, f, e);
    if (st) break;
    for (int i = 0; i < n; ++i) {
      double xi[3];
      orc_normals(seed, (uint32_t)i, (uint64_t)step, xi);
      for (int d = 0; d < 3; ++d) x[3 * i + d] += A5 * (f[3 * i + d] + VAR * xi[d]);
    }
    if (nup > 0 && samples && (step + 1) % nup == 0) {
      for (int i = 0; i < 3 * n; ++i) xf[i] = (float)x[i];
      st = orc_forces(s, x, L.pla_k, L.pla, L.plr_i, L.plr_j, L.plr, f, e);
      if (st) break;
      double q, rg2;
      int32_t c;
      orc_observables(s, xf, x, &q, &c, &rg2);
      samples[0] = e[4];
      samples[1] = rg2;
      samples[2] = q;
      samples[3] = (double)c;
      samples += 4;
    }
  }
  free(xf); free(f); lists_free(&L);
  return st;
}

/* Runs nsteps from x0 (float) with step counter starting at step0; writes the
 * final FP64 coordinates into x_out (3n).  Lists are built with the FP32
 * predicate on the coordinates rounded to float.  Returns 0 or <0. */
ORC_EXPORT int orc_run(const orc_sys *s, const float *x0, int64_t step0, int64_t nsteps,
                       uint64_t seed, float a5, float var, int32_t n_nl, int32_t n_pl,
                       double *x_out) {
  return run_od(s, x0, step0, nsteps, seed, a5, var, n_nl, n_pl, x_out, 0, 0, NULL);
}

/* orc_run with the neighbor list of orc_nl_grid (open boundaries only). */
ORC_EXPORT int orc_run_grid(const orc_sys *s, const float *x0, int64_t step0, int64_t nsteps,
                            uint64_t seed, float a5, float var, int32_t n_nl, int32_t n_pl,
                            double *x_out) {
  if (s->boxl > 0.0f) return -4;
  return run_od(s, x0, step0, nsteps, seed, a5, var, n_nl, n_pl, x_out, 1, 0, NULL);
}

/* orc_run that also records, every nup steps, the total energy, Rg^2, Q
 * and the native contact count (4 doubles per sample, floor(nsteps / nup)
 * samples when step0 is a multiple of nup): the long-run observables of the
 * reference's update.out (sop.cu:394-396). */
ORC_EXPORT int orc_run_sample(const orc_sys *s, const float *x0, int64_t step0, int64_t nsteps,
                              uint64_t seed, float a5, float var, int32_t n_nl, int32_t n_pl,
                              int32_t nup, double *x_out, double *samples) {
  if (nup < 1) return -4;
  return run_od(s, x0, step0, nsteps, seed, a5, var, n_nl, n_pl, x_out, 0, nup, samples);
}

/* Underdamped Langevin dynamics (sop.cu:421-484, coefficients
 * param.cpp:104-108) in FP64.  The reference computes incr = a1 v + a2 F,
 * moves, evaluates F at the new coordinates, then v = a3 incr + a4 F (no
 * velocity update when T < 1e-5, sop.cu:475).  Restated per step s with the
 * lists rebuilt before the force evaluation as in orc_run:
 *   F_s = F(x_s) + var * xi(seed, bead, s)
 *   v_s = a3 u + a4 F_s    if u holds the previous increment, else v_s = u
 *   incr = a1 v_s + a2 F_s;  x_{s+1} = x_s + incr
 *   u = incr (velocities updated) or v_s (frozen).
 * v0: velocities entering the first step (NULL = 0).  v_out: the reference's
 * velocities after the last step (a3 u + a4 F_end), with F_end evaluated at
 * the final coordinates, step step0 + nsteps, on the lists of the last step. */
ORC_EXPORT int orc_run_ud(const orc_sys *s, const float *x0, const float *v0, int64_t step0, int64_t nsteps,
                          uint64_t seed, const float *a, float var, int32_t vupd, int32_t n_nl, int32_t n_pl,
                          double *x_out, double *v_out) {
  int n = s->n;
  double *x = x_out;
  float *xf = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  orc_lists L;
  int st = lists_alloc(s, &L, lists_cap(s, 0));
  if (!xf || !f || !u || st) { free(xf); free(f); free(u); lists_free(&L); return -2; }
  for (int i = 0; i < 3 * n; ++i) {
    x[i] = (double)x0[i];
    u[i] = v0 ? (double)v0[i] : 0.0;
  }
  int have_incr = 0;
  const double A1 = a[0], A2 = a[1], A3 = a[2], A4 = a[3], VAR = var;
  double e[6];
  for (int64_t it = 0; it < nsteps; ++it) {
    int64_t step = step0 + it;
    for (int i = 0; i < 3 * n; ++i) xf[i] = (float)x[i];
    if (it == 0 || step % n_nl == 0) {
      st = orc_nl(s, xf, L.nla_k, s->n_nat + 1, &L.nla, L.nlr_i, L.nlr_j, L.cap_rep, &L.nlr);
      if (st) break;
    }
    if (it == 0 || step % n_pl == 0)
      orc_pl(s, xf, L.nla_k, L.nla, L.nlr_i, L.nlr_j, L.nlr, L.pla_k, &L.pla, L.plr_i, L.plr_j, &L.plr);
    st = orc_forces(s, x, L.pla_k, L.pla, L.plr_i, L.plr_j, L.plr, f, e);
    if (st) break;
    for (int i = 0; i < n; ++i) {
      double xi[3];
      orc_normals(seed, (uint32_t)i, (uint64_t)step, xi);
      for (int d = 0; d < 3; ++d) {
        const int k = 3 * i + d;
        const double F = f[k] + VAR * xi[d];
        const double v = have_incr ? A3 * u[k] + A4 * F : u[k];
        const double incr = A1 * v + A2 * F;
        x[k] += incr;
        u[k] = vupd ? incr : v;
      }
    }
    have_incr = vupd;
  }
  if (!st && v_out) {
    if (have_incr) {
      if (nsteps == 0) {
        for (int i = 0; i < 3 * n; ++i) xf[i] = (float)x[i];
        orc_nl(s, xf, L.nla_k, s->n_nat + 1, &L.nla, L.nlr_i, L.nlr_j, L.cap_rep, &L.nlr);
        orc_pl(s, xf, L.nla_k, L.nla, L.nlr_i, L.nlr_j, L.nlr, L.pla_k, &L.pla, L.plr_i, L.plr_j, &L.plr);
      }
      st = orc_forces(s, x, L.pla_k, L.pla, L.plr_i, L.plr_j, L.plr, f, e);
      for (int i = 0; i < n && !st; ++i) {
        double xi[3];
        orc_normals(seed, (uint32_t)i, (uint64_t)(step0 + nsteps), xi);
        for (int d = 0; d < 3; ++d) {
          const int k = 3 * i + d;
          v_out[k] = A3 * u[k] + A4 * (f[k] + VAR * xi[d]);
        }
      }
    } else {
      for (int i = 0; i < 3 * n; ++i) v_out[i] = u[i];
    }
  }
  free(xf); free(f); free(u); lists_free(&L);
  return st;
}

/* ------------------------------------------------------------------ */
/* FP32 port of the reference's host method (CPU baseline).             */
/* Explicit O(N^2) master non-native list (as loaded from the vdw file, */
/* io.cpp:190-226), neighbor filter over the master list every n_nl     */
/* (sop.cu:1354-1448), pair filter every n_pl (sop.cu:1710-1823), the   */
/* four force terms of energy.cu with half lists and serial +=          */
/* accumulation, FP32 random force, overdamped update (sop.cu:487-520). */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_sys *s;
  int32_t *mi, *mj; int64_t nm;            /* master non-native list */
  int32_t *nli, *nlj; int64_t nnl;
  int32_t *pli, *plj; int64_t npl;
  int32_t *nla; int64_t nnla;
  int32_t *pla; int64_t npla;
  float *f;
} orc_port;

ORC_EXPORT void *orc_port_create(const orc_sys *s) {
  orc_port *p = (orc_port *)calloc(1, sizeof(orc_port));
  if (!p) return NULL;
  p->s = s;
  int64_t n = s->n, cap = n * (n - 1) / 2 + 1;
  p->mi = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->mj = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->nli = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->nlj = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->pli = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->plj = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  p->nla = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n_nat + 1));
  p->pla = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n_nat + 1));
  p->f = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  nat_index ix;
  if (!p->mi || !p->mj || !p->nli || !p->nlj || !p->pli || !p->plj || !p->nla || !p->pla ||
      !p->f || nat_index_build(s, &ix)) {
    return NULL;
  }
  for (int i = 0; i < s->n; ++i)
    for (int j = i + 1; j < s->n; ++j)
      if (!excluded(s, i, j) && !is_native(&ix, i, j)) {
        p->mi[p->nm] = i;
        p->mj[p->nm] = j;
        p->nm++;
      }
  nat_index_free(&ix);
  return p;
}

ORC_EXPORT void orc_port_destroy(void *h) {
  orc_port *p = (orc_port *)h;
  if (!p) return;
  free(p->mi); free(p->mj); free(p->nli); free(p->nlj); free(p->pli); free(p->plj);
  free(p->nla); free(p->pla); free(p->f);
  free(p);
}

static void port_forces(orc_port *p, const float *x) {
  const orc_sys *s = p->s;
  const float boxl = s->boxl;
  float *f = p->f;
  memset(f, 0, sizeof(float) * 3 * (size_t)s->n);
  const float k = s->k_bnd, R0sq = s->R0 * s->R0;
  for (int i = 0; i + 1 < s->n; ++i) {
    if (s->chain[i] != s->chain[i + 1]) continue;
    int j = i + 1;
    float dx = orc_min_image_f(x[3 * j] - x[3 * i], boxl);
    float dy = orc_min_image_f(x[3 * j + 1] - x[3 * i + 1], boxl);
    float dz = orc_min_image_f(x[3 * j + 2] - x[3 * i + 2], boxl);
    float d = sqrtf(dx * dx + dy * dy + dz * dz);
    float dev = d - s->bond_r0[i];
    float t = -k * dev / d / (1.0f - dev * dev / R0sq);
    f[3 * i] -= t * dx; f[3 * i + 1] -= t * dy; f[3 * i + 2] -= t * dz;
    f[3 * j] += t * dx; f[3 * j + 1] += t * dy; f[3 * j + 2] += t * dz;
  }
  const float fss = 6.0f * s->eps_ss * powf(s->sigma_ss, 6.0f);
  for (int i = 0; i + 2 < s->n; ++i) {
    if (s->chain[i] != s->chain[i + 2]) continue;
    int j = i + 2;
    float dx = orc_min_image_f(x[3 * j] - x[3 * i], boxl);
    float dy = orc_min_image_f(x[3 * j + 1] - x[3 * i + 1], boxl);
    float dz = orc_min_image_f(x[3 * j + 2] - x[3 * i + 2], boxl);
    float d = sqrtf(dx * dx + dy * dy + dz * dz);
    float co = fss / powf(d, 8.0f);
    f[3 * i] -= co * dx; f[3 * i + 1] -= co * dy; f[3 * i + 2] -= co * dz;
    f[3 * j] += co * dx; f[3 * j + 1] += co * dy; f[3 * j + 2] += co * dz;
  }
  for (int64_t m = 0; m < p->npla; ++m) {
    int kk = p->pla[m];
    int i = s->nat_i[kk], j = s->nat_j[kk];
    float dx = orc_min_image_f(x[3 * j] - x[3 * i], boxl);
    float dy = orc_min_image_f(x[3 * j + 1] - x[3 * i + 1], boxl);
    float dz = orc_min_image_f(x[3 * j + 2] - x[3 * i + 2], boxl);
    float d2 = dx * dx + dy * dy + dz * dz;
    float r02 = s->nat_r0[kk] * s->nat_r0[kk];
    if (d2 > 1.0e-7f * r02) {
      float d6 = d2 * d2 * d2, d12 = d6 * d6;
      float r06 = r02 * r02 * r02, r012 = r06 * r06;
      float co = -12.0f * s->eps_att[s->type[i] * 3 + s->type[j]] / d2 * (r012 / d12 - r06 / d6);
      f[3 * i] += co * dx; f[3 * i + 1] += co * dy; f[3 * i + 2] += co * dz;
      f[3 * j] -= co * dx; f[3 * j + 1] -= co * dy; f[3 * j + 2] -= co * dz;
    }
  }
  for (int64_t m = 0; m < p->npl; ++m) {
    int i = p->pli[m], j = p->plj[m];
    float dx = orc_min_image_f(x[3 * j] - x[3 * i], boxl);
    float dy = orc_min_image_f(x[3 * j + 1] - x[3 * i + 1], boxl);
    float dz = orc_min_image_f(x[3 * j + 2] - x[3 * i + 2], boxl);
    float d2 = dx * dx + dy * dy + dz * dz;
    int t = s->type[i] * 3 + s->type[j];
    float sg = s->sigma_rep[t];
    if (d2 > sg * 1.0e-7f) {
      float s2 = sg * sg, s6 = s2 * s2 * s2, s12 = s6 * s6;
      float d6 = d2 * d2 * d2, d12 = d6 * d6;
      float co = -6.0f * s->eps_rep[t] / d2 * (2.0f * s->rep_c12 * s12 / d12 + s->rep_c6 * s6 / d6);
      f[3 * i] += co * dx; f[3 * i + 1] += co * dy; f[3 * i + 2] += co * dz;
      f[3 * j] -= co * dx; f[3 * j + 1] -= co * dy; f[3 * j + 2] -= co * dz;
    }
  }
}

/* Advance x (float, 3n, in place) by nsteps; returns 0. */
ORC_EXPORT int orc_port_run(void *h, float *x, int64_t step0, int64_t nsteps, uint64_t seed,
                            float a5, float var, int32_t n_nl, int32_t n_pl) {
  orc_port *p = (orc_port *)h;
  const orc_sys *s = p->s;
  const float boxl = s->boxl;
  for (int64_t it = 0; it < nsteps; ++it) {
    int64_t step = step0 + it;
    if (it == 0 || step % n_nl == 0) {
      p->nnla = 0;
      for (int k = 0; k < s->n_nat; ++k) {
        float rc = orc_cutoff(s->nl_factor, s->nl_abs, s->nat_r0[k]);
        if (dist2_f(x, s->nat_i[k], s->nat_j[k], boxl) < rc * rc) p->nla[p->nnla++] = k;
      }
      p->nnl = 0;
      for (int64_t m = 0; m < p->nm; ++m) {
        int i = p->mi[m], j = p->mj[m];
        float rc = orc_cutoff(s->nl_factor, s->nl_abs, s->sigma_rep[s->type[i] * 3 + s->type[j]]);
        if (dist2_f(x, i, j, boxl) < rc * rc) {
          p->nli[p->nnl] = i;
          p->nlj[p->nnl] = j;
          p->nnl++;
        }
      }
    }
    if (it == 0 || step % n_pl == 0) {
      p->npla = 0;
      for (int64_t m = 0; m < p->nnla; ++m) {
        int k = p->nla[m];
        float rc = orc_cutoff(s->pl_factor, s->pl_abs, s->nat_r0[k]);
        if (dist2_f(x, s->nat_i[k], s->nat_j[k], boxl) < rc * rc) p->pla[p->npla++] = k;
      }
      p->npl = 0;
      for (int64_t m = 0; m < p->nnl; ++m) {
        int i = p->nli[m], j = p->nlj[m];
        float rc = orc_cutoff(s->pl_factor, s->pl_abs, s->sigma_rep[s->type[i] * 3 + s->type[j]]);
        if (dist2_f(x, i, j, boxl) < rc * rc) {
          p->pli[p->npl] = i;
          p->plj[p->npl] = j;
          p->npl++;
        }
      }
    }
    port_forces(p, x);
    for (int i = 0; i < s->n; ++i) {
      double xi[3];
      orc_normals(seed, (uint32_t)i, (uint64_t)step, xi);
      for (int d = 0; d < 3; ++d) {
        float ft = p->f[3 * i + d] + (float)xi[d] * var;
        x[3 * i + d] += a5 * ft;
      }
    }
  }
  return 0;
}

ORC_EXPORT int64_t orc_port_master_size(void *h) { return ((orc_port *)h)->nm; }
