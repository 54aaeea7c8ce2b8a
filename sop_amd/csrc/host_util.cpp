// host_util.cpp -- host-only parts of libsop: parameter sets, the seeded
// compact-chain generator, native maps, and readers/writers of the
// reference's input and trajectory files.  Nothing here touches a device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sop.h"

extern "C" {

sop_status sop_params_reference(sop_params *p) {
  if (!p) return SOP_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->k_bnd = 20.0f;    // param.cpp:14
  p->R0 = 2.0f;        // param.cpp:15
  p->eps_ss = 1.0f;    // param.cpp:37
  p->sigma_ss = 3.5f;  // param.cpp:35
  // One well-depth table for energy and force: the force table of
  // energy.cu:3-7 (the energy table global.cpp:51-55 is its transpose).
  const float ea[3][3] = {{0, 0, 0}, {0, 1.0f, 0.8f}, {0, 0.8f, 0.7f}};
  const float sr[3][3] = {{0, 0, 0}, {0, 3.8f, 5.4f}, {0, 5.4f, 7.0f}};  // global.cpp:75-79
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      p->eps_att[i][j] = ea[i][j];
      p->eps_rep[i][j] = (i && j) ? 1.0f : 0.0f;  // global.cpp:57-61
      p->sigma_rep[i][j] = sr[i][j];
    }
  p->rep_c12 = 1.0f;  // energy.cu:436-437
  p->rep_c6 = 1.0f;
  p->nl_factor = 3.2;  // sop.cu:746,815
  p->nl_abs = 0.0f;
  p->pl_factor = 2.5;  // sop.cu:861,910
  p->pl_abs = 0.0f;
  p->n_nl = 50;        // param.cpp:103
  p->n_pl = 1;         // sop.cu:375-387 (pair list every step)
  p->T = 0.6f;         // param.cpp:92
  p->zeta = 0.05f;     // param.cpp:100
  p->h = 0.02f;        // param.cpp:252 (dynamics overdamped)
  p->boxl = 500.0f;    // param.cpp:97
  p->seed = 1234;      // sop.cu:138
  return SOP_OK;
}

sop_status sop_params_config(sop_params *p) {
  if (!p) return SOP_ERR_INVALID;
  sop_params_reference(p);
  for (int i = 1; i < 3; ++i)
    for (int j = 1; j < 3; ++j) {
      p->eps_att[i][j] = 1.0f;  // eps_h
      p->eps_rep[i][j] = 1.0f;  // eps_l
    }
  p->sigma_ss = 5.4f;
  p->nl_factor = 0.0;
  p->nl_abs = 32.0f;
  p->pl_factor = 0.0;
  p->pl_abs = 16.0f;
  p->n_nl = 50;
  p->n_pl = 10;
  p->T = 0.6f;
  p->zeta = 50.0f;
  p->h = 0.05f;
  p->boxl = 0.0f;
  p->seed = 1;
  return SOP_OK;
}

sop_status sop_params_derived(const sop_params *p, float *a5, float *var) {
  if (!p) return SOP_ERR_INVALID;
  const float h = p->h, zeta = p->zeta, T = p->T;
  if (a5) *a5 = h / zeta;
  if (var) *var = (float)std::sqrt(2.0 * (double)T * (double)zeta / (double)h);
  return SOP_OK;
}

sop_status sop_params_underdamped(const sop_params *p, float *a) {
  if (!p || !a) return SOP_ERR_INVALID;
  // Products of two floats round to float; every expression that meets a
  // double literal continues in double and is stored to float at the end.
  const float h = p->h, zeta = p->zeta;
  const float halfh = (float)((double)h / 2.0);
  const float zh = zeta * halfh;
  const float hz = h * zeta;
  const float hz2 = hz * hz;
  a[0] = (float)((double)h * (1.0 - (double)zh));
  a[1] = h * halfh;
  a[2] = (float)((1.0 - (double)hz / 2.0 + (double)hz2 / 4.0) / (double)h);
  a[3] = (float)((double)halfh * (1.0 - (double)hz / 2.0));
  return SOP_OK;
}

void sop_gen_opts_default(sop_gen_opts *o) {
  o->bond = 3.8f;
  o->min_dist = 4.0f;
  o->volume_per_bead = 140.0f;
  o->relax_steps = 0;
}

} // extern "C"

// ------------------------------------------------------------ generator

namespace {

// xoshiro256** (Blackman & Vigna, 2018) seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Grid {
  double lo, edge;
  int dim;
  std::vector<int32_t> head, next;
  void init(double radius, double e, int n) {
    edge = e;
    lo = -radius - 2.0 * e;
    dim = std::max(1, (int)std::ceil((2.0 * radius + 4.0 * e) / e));
    head.assign((size_t)dim * dim * dim, -1);
    next.assign((size_t)n, -1);
  }
  int cell1(double v) const { return std::min(dim - 1, std::max(0, (int)std::floor((v - lo) / edge))); }
  int cell(const double *x) const { return (cell1(x[2]) * dim + cell1(x[1])) * dim + cell1(x[0]); }
  void insert(int k, const double *x) {
    const int c = cell(x);
    next[k] = head[c];
    head[c] = k;
  }
  void remove_last(int k, const double *x) {  // LIFO removal
    const int c = cell(x);
    if (head[c] == k) head[c] = next[k];
  }
};

} // namespace

extern "C" sop_status sop_generate_chain(uint64_t seed, int32_t n, const sop_gen_opts *o, float *xyz,
                                         float *bond_r0) {
  if (n < 1 || !xyz) return SOP_ERR_INVALID;
  sop_gen_opts opt;
  sop_gen_opts_default(&opt);
  if (o) opt = *o;
  const double b = opt.bond, md = opt.min_dist, md2 = md * md;
  double radius = std::cbrt(3.0 * n * opt.volume_per_bead / (4.0 * M_PI)) + b;
  std::vector<double> x(3 * (size_t)n);
  for (int attempt_scale = 0; attempt_scale < 40; ++attempt_scale) {
    Rng rng(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull + (uint64_t)attempt_scale);
    Grid g;
    g.init(radius, md, n);
    x[0] = x[1] = x[2] = 0.0;
    g.insert(0, &x[0]);
    int i = 1;
    int64_t budget = (int64_t)n * 4000;
    bool failed = false;
    while (i < n) {
      bool placed = false;
      for (int t = 0; t < 64 && !placed; ++t) {
        if (--budget < 0) { failed = true; break; }
        const double u = 2.0 * rng.uniform() - 1.0, ph = 2.0 * M_PI * rng.uniform();
        const double sr = std::sqrt(std::max(0.0, 1.0 - u * u));
        double c[3] = {x[3 * (i - 1)] + b * sr * std::cos(ph), x[3 * (i - 1) + 1] + b * sr * std::sin(ph),
                       x[3 * (i - 1) + 2] + b * u};
        if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > radius * radius) continue;
        bool clash = false;
        const int cx = g.cell1(c[0]), cy = g.cell1(c[1]), cz = g.cell1(c[2]);
        for (int dz = -1; dz <= 1 && !clash; ++dz)
          for (int dy = -1; dy <= 1 && !clash; ++dy)
            for (int dx = -1; dx <= 1 && !clash; ++dx) {
              const int zz = cz + dz, yy = cy + dy, xx = cx + dx;
              if (zz < 0 || yy < 0 || xx < 0 || zz >= g.dim || yy >= g.dim || xx >= g.dim) continue;
              for (int k = g.head[(zz * g.dim + yy) * g.dim + xx]; k >= 0; k = g.next[k]) {
                if (k == i - 1) continue;  // bonded neighbor
                const double ex = c[0] - x[3 * k], ey = c[1] - x[3 * k + 1], ez = c[2] - x[3 * k + 2];
                if (ex * ex + ey * ey + ez * ez < md2) { clash = true; break; }
              }
            }
        if (clash) continue;
        x[3 * i] = c[0]; x[3 * i + 1] = c[1]; x[3 * i + 2] = c[2];
        g.insert(i, &x[3 * i]);
        placed = true;
      }
      if (failed) break;
      if (placed) {
        ++i;
      } else {
        // dead end: back off a few beads
        const int back = std::min(i - 1, 1 + (int)(rng.next() % 8));
        for (int k = 0; k < back; ++k) {
          --i;
          g.remove_last(i, &x[3 * i]);
        }
      }
    }
    if (!failed) {
      for (int k = 0; k < 3 * n; ++k) xyz[k] = (float)x[k];
      if (bond_r0) {
        for (int k = 0; k + 1 < n; ++k) {
          const double ex = (double)xyz[3 * k + 3] - xyz[3 * k], ey = (double)xyz[3 * k + 4] - xyz[3 * k + 1],
                       ez = (double)xyz[3 * k + 5] - xyz[3 * k + 2];
          bond_r0[k] = (float)std::sqrt(ex * ex + ey * ey + ez * ez);
        }
        bond_r0[n - 1] = (float)b;
      }
      return SOP_OK;
    }
    radius *= 1.03;  // too tight: loosen the confinement and retry
  }
  return SOP_ERR_STATE;
}

extern "C" sop_status sop_native_contacts(int32_t n, const float *xyz, const int32_t *chain, float cutoff,
                                          int32_t intra_only, int32_t *nat_i, int32_t *nat_j, float *nat_r0,
                                          int64_t cap, int64_t *n_nat) {
  if (n < 1 || !xyz || !n_nat || !(cutoff > 0.0f)) return SOP_ERR_INVALID;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], (double)xyz[3 * i + d]);
      hi[d] = std::max(hi[d], (double)xyz[3 * i + d]);
    }
  const double e = cutoff;
  int dims[3];
  for (int d = 0; d < 3; ++d) dims[d] = std::max(1, std::min(1024, (int)((hi[d] - lo[d]) / e) + 1));
  const int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
  std::vector<int32_t> head((size_t)ncell, -1), nxt((size_t)n, -1);
  auto cidx = [&](int i, int *c) {
    for (int d = 0; d < 3; ++d) c[d] = std::min(dims[d] - 1, std::max(0, (int)((xyz[3 * i + d] - lo[d]) / e)));
  };
  for (int i = n - 1; i >= 0; --i) {
    int c[3];
    cidx(i, c);
    const int64_t q = ((int64_t)c[2] * dims[1] + c[1]) * dims[0] + c[0];
    nxt[i] = head[q];
    head[q] = i;
  }
  int64_t m = 0;
  const double cut2 = (double)cutoff * cutoff;
  std::vector<std::pair<int32_t, float>> row;
  for (int i = 0; i < n; ++i) {
    int c[3];
    cidx(i, c);
    row.clear();
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int zz = c[2] + dz, yy = c[1] + dy, xx = c[0] + dx;
          if (zz < 0 || yy < 0 || xx < 0 || zz >= dims[2] || yy >= dims[1] || xx >= dims[0]) continue;
          for (int j = head[((int64_t)zz * dims[1] + yy) * dims[0] + xx]; j >= 0; j = nxt[j]) {
            if (j <= i) continue;
            const bool same = !chain || chain[i] == chain[j];
            if (same && j - i <= 2) continue;
            if (!same && intra_only) continue;
            const double ex = (double)xyz[3 * j] - xyz[3 * i], ey = (double)xyz[3 * j + 1] - xyz[3 * i + 1],
                         ez = (double)xyz[3 * j + 2] - xyz[3 * i + 2];
            const double d2 = ex * ex + ey * ey + ez * ez;
            if (d2 < cut2) row.push_back({j, (float)std::sqrt(d2)});
          }
        }
    std::sort(row.begin(), row.end());
    for (auto &pr : row) {
      if (m < cap) {
        if (nat_i) nat_i[m] = i;
        if (nat_j) nat_j[m] = pr.first;
        if (nat_r0) nat_r0[m] = pr.second;
      }
      ++m;
    }
  }
  *n_nat = m;
  return SOP_OK;
}

// -------------------------------------------------- reference file formats

struct sop_topology {
  std::vector<float> xyz, bond_r0, nat_r0;
  std::vector<int32_t> chain, type, nat_i, nat_j;
  sop_params params{};
};

namespace {

bool next_line(std::ifstream &in, std::string &s) {
  while (std::getline(in, s)) {
    if (s.find_first_not_of(" \t\r") != std::string::npos) return true;
  }
  return false;
}

// Header lines are "<word> <count>" (io.cpp:131-134,154-157,179-185,236-239).
bool header_count(std::ifstream &in, int which, long &v) {
  std::string s;
  if (!next_line(in, s)) return false;
  std::istringstream is(s);
  std::string tok;
  std::vector<std::string> t;
  while (is >> tok) t.push_back(tok);
  if ((int)t.size() <= which) return false;
  v = std::atol(t[which].c_str());
  return true;
}

} // namespace

extern "C" sop_status sop_topology_load(const char *bonds, const char *angles, const char *vdw, const char *init,
                                        sop_topology **out) {
  if (!bonds || !vdw || !init || !out) return SOP_ERR_INVALID;
  *out = nullptr;
  auto *t = new sop_topology();
  sop_params_reference(&t->params);
  std::string s;
  // init: "X nbead" then "idx x y z" (io.cpp:231-257)
  {
    std::ifstream in(init);
    long n = 0;
    if (!in || !header_count(in, 1, n) || n <= 0) { delete t; return SOP_ERR_IO; }
    t->xyz.resize(3 * (size_t)n);
    for (long i = 0; i < n; ++i) {
      if (!next_line(in, s)) { delete t; return SOP_ERR_IO; }
      std::istringstream is(s);
      long idx;
      double a, b, c;
      if (!(is >> idx >> a >> b >> c)) { delete t; return SOP_ERR_IO; }
      t->xyz[3 * i] = (float)a;
      t->xyz[3 * i + 1] = (float)b;
      t->xyz[3 * i + 2] = (float)c;
    }
  }
  const int n = (int)(t->xyz.size() / 3);
  // bonds: "X nbnd" then "i j r0", 1-based (io.cpp:126-148); chains follow bonds (i, i+1)
  t->bond_r0.assign((size_t)n, 3.8f);
  std::vector<char> bonded((size_t)n, 0);
  {
    std::ifstream in(bonds);
    long nb = 0;
    if (!in || !header_count(in, 1, nb) || nb < 0) { delete t; return SOP_ERR_IO; }
    for (long k = 0; k < nb; ++k) {
      if (!next_line(in, s)) { delete t; return SOP_ERR_IO; }
      std::istringstream is(s);
      long i, j;
      double r0;
      if (!(is >> i >> j >> r0)) { delete t; return SOP_ERR_IO; }
      if (i > j) std::swap(i, j);
      if (i < 1 || j > n || j != i + 1) { delete t; return SOP_ERR_INVALID; }  // chain topology only
      bonded[i - 1] = 1;
      t->bond_r0[i - 1] = (float)r0;
    }
  }
  t->chain.assign((size_t)n, 0);
  for (int i = 1; i < n; ++i) t->chain[i] = t->chain[i - 1] + (bonded[i - 1] ? 0 : 1);
  // angles: "X nang" then "i j k a" (io.cpp:149-173); must be (i, i+1, i+2) in one chain
  if (angles) {
    std::ifstream in(angles);
    long na = 0;
    if (!in || !header_count(in, 1, na) || na < 0) { delete t; return SOP_ERR_IO; }
    for (long k = 0; k < na; ++k) {
      if (!next_line(in, s)) { delete t; return SOP_ERR_IO; }
      std::istringstream is(s);
      long i, j, kk;
      if (!(is >> i >> j >> kk)) { delete t; return SOP_ERR_IO; }
      if (i > kk) std::swap(i, kk);
      if (i < 1 || kk > n || kk != i + 2 || t->chain[i - 1] != t->chain[kk - 1]) { delete t; return SOP_ERR_INVALID; }
    }
  }
  // vdw: "X ncon_att X ncon_rep" then "i j r_ij itype jtype"; native iff
  // r_ij < rcut_nat[itype][jtype] (io.cpp:174-230, global.cpp:88-92)
  {
    const float rcut_nat[3][3] = {{0, 0, 0}, {0, 8.0f, 11.0f}, {0, 11.0f, 14.0f}};
    std::ifstream in(vdw);
    std::string hs;
    if (!in || !next_line(in, hs)) { delete t; return SOP_ERR_IO; }
    std::istringstream hsi(hs);
    std::string w0, w2;
    long natt = 0, nrep = 0;
    if (!(hsi >> w0 >> natt >> w2 >> nrep)) { delete t; return SOP_ERR_IO; }
    t->type.assign((size_t)n, 0);
    const long ntot = natt + nrep;
    long got_att = 0;
    for (long k = 0; k < ntot; ++k) {
      if (!next_line(in, s)) { delete t; return SOP_ERR_IO; }
      const char *c = s.c_str();
      char *end;
      long i = std::strtol(c, &end, 10);
      long j = std::strtol(end, &end, 10);
      float r = std::strtof(end, &end);
      long ti = std::strtol(end, &end, 10);
      long tj = std::strtol(end, &end, 10);
      if (i > j) { std::swap(i, j); std::swap(ti, tj); }
      if (i < 1 || j > n || i == j || ti < 1 || ti > 2 || tj < 1 || tj > 2) { delete t; return SOP_ERR_INVALID; }
      if ((t->type[i - 1] && t->type[i - 1] != ti) || (t->type[j - 1] && t->type[j - 1] != tj)) {
        delete t;
        return SOP_ERR_INVALID;  // a bead with two types
      }
      t->type[i - 1] = (int32_t)ti;
      t->type[j - 1] = (int32_t)tj;
      if (t->chain[i - 1] == t->chain[j - 1] && j - i <= 2) { delete t; return SOP_ERR_INVALID; }
      if (r < rcut_nat[ti][tj]) {
        t->nat_i.push_back((int32_t)(i - 1));
        t->nat_j.push_back((int32_t)(j - 1));
        t->nat_r0.push_back(r);
        ++got_att;
      }
    }
    for (auto &v : t->type)
      if (!v) v = 1;
    // the engine's non-native set is every non-excluded, non-native pair:
    // the file must enumerate exactly that master list
    int64_t master = 0;
    std::vector<int64_t> per_chain;
    for (int i = 0; i < n;) {
      int j = i;
      while (j < n && t->chain[j] == t->chain[i]) ++j;
      per_chain.push_back(j - i);
      i = j;
    }
    const int64_t all = (int64_t)n * (n - 1) / 2;
    int64_t excl = 0;
    for (int64_t L : per_chain) excl += std::max<int64_t>(0, L - 1) + std::max<int64_t>(0, L - 2);
    master = all - excl;
    if (ntot != master) { delete t; return SOP_ERR_INVALID; }
    (void)got_att;
  }
  *out = t;
  return SOP_OK;
}

extern "C" sop_status sop_topology_desc(const sop_topology *t, sop_system_desc *d) {
  if (!t || !d) return SOP_ERR_INVALID;
  std::memset(d, 0, sizeof(*d));
  d->n_beads = (int32_t)(t->xyz.size() / 3);
  d->xyz = t->xyz.data();
  d->chain = t->chain.data();
  d->type = t->type.data();
  d->bond_r0 = t->bond_r0.data();
  d->n_native = (int32_t)t->nat_i.size();
  d->nat_i = t->nat_i.data();
  d->nat_j = t->nat_j.data();
  d->nat_r0 = t->nat_r0.data();
  d->params = t->params;
  return SOP_OK;
}

extern "C" void sop_topology_free(sop_topology *t) { delete t; }

extern "C" sop_status sop_write_traj(const char *path, int32_t n, const float *xyz) {
  FILE *f = std::fopen(path, "a");
  if (!f) return SOP_ERR_IO;
  for (int i = 0; i < n; ++i) std::fprintf(f, "%f %f %f\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fclose(f);
  return SOP_OK;
}

extern "C" sop_status sop_save_coords(const char *path, int32_t n, const float *xyz) {
  FILE *f = std::fopen(path, "w");
  if (!f) return SOP_ERR_IO;
  for (int i = 0; i < n; ++i) std::fprintf(f, "%d %f %f %f\n", i + 1, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fclose(f);
  return SOP_OK;
}

extern "C" sop_status sop_load_coords(const char *path, int32_t n, float *xyz) {
  std::ifstream in(path);
  if (!in) return SOP_ERR_IO;
  std::string s;
  for (int i = 0; i < n; ++i) {
    if (!next_line(in, s)) return SOP_ERR_IO;
    std::istringstream is(s);
    long idx;
    double a, b, c;
    if (!(is >> idx >> a >> b >> c)) return SOP_ERR_IO;
    xyz[3 * i] = (float)a;
    xyz[3 * i + 1] = (float)b;
    xyz[3 * i + 2] = (float)c;
  }
  return SOP_OK;
}
