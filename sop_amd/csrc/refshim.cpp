// refshim.cpp -- the reference's energy.h / sop.h / GPUvars.h entry points on
// top of libsop (include/sop_refshim.h).  The driver's globals stay the state
// it sees: coordinates go to the engine before each evaluation, forces,
// energies and lists come back into the reference-named arrays.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sop.h"
#include "../../include/sop_refshim.h"

namespace {

sop_engine *g_eng = nullptr;
int32_t g_sys = -1;
uint64_t g_seed = 1234;    // setup_rng(1234, 0) in sop.cu:138
int64_t g_counter = 0;     // Philox counter of the next force_eval
std::unordered_map<uint64_t, float> g_nat_r0;  // native (i<j) -> r0
std::vector<int32_t> g_type;                   // bead types 1/2
std::vector<sop_ref_float3> g_incr;            // underdamped increments

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "sop_refshim: %s\n", msg.c_str());
  std::exit(-1);
}

void check(sop_status st, const char *what) {
  if (st == SOP_OK) return;
  if (!g_eng) die(std::string(what) + ": status " + std::to_string(st) + (st == SOP_ERR_NO_DEVICE ? " (no HIP device)" : ""));
  die(std::string(what) + ": " + sop_last_error(g_eng));
}

void need_engine(const char *what) {
  if (!g_eng || g_sys < 0) die(std::string(what) + " called before alloc_GPU_arrays");
}

uint64_t key(int32_t i, int32_t j) { return ((uint64_t)(uint32_t)i << 32) | (uint32_t)j; }

// COMBINE / GET_IDX / GET_TYPE of global.h:51-67 (14-bit 1-based index, 2-bit type)
int get_idx(unsigned short c) { return c & 0x3FFF; }
int get_type(unsigned short c) { return c >> 14; }
unsigned short combine(int type, int idx) { return (unsigned short)((type << 14) | idx); }

// misc.cpp:4-11: round half away from zero
float rnd(float x) { return (float)(x > 0 ? std::floor(x + 0.5) : std::ceil(x - 0.5)); }

void push_coords() { check(sop_set_coords(g_eng, g_sys, &unc_pos[0].x), "sop_set_coords"); }

void set_type(int32_t i, int t) {
  if (i < 0 || i >= nbead) die("pair index out of range");
  if (t != 1 && t != 2) die("bead types must be 1 or 2");
  if (g_type[i] != 0 && g_type[i] != t) die("bead " + std::to_string(i + 1) + " has two types");
  g_type[i] = t;
}

// Copies a half list into the driver's COMBINE'd array (when representable).
void copy_list(int which, int *count, sop_ref_ushort2 *dst, int cap, float *r0dst) {
  int64_t n = 0;
  check(sop_get_list(g_eng, g_sys, which, nullptr, 0, &n), "sop_get_list");
  *count = (int)n;
  if (!dst || n == 0 || n > cap || nbead > 0x3FFF) return;
  std::vector<int32_t> ij(2 * (size_t)n);
  check(sop_get_list(g_eng, g_sys, which, ij.data(), n, &n), "sop_get_list");
  for (int64_t k = 0; k < n; ++k) {
    const int32_t i = ij[2 * k], j = ij[2 * k + 1];
    dst[k].x = combine(g_type[i], i + 1);
    dst[k].y = combine(g_type[j], j + 1);
    if (r0dst) r0dst[k] = g_nat_r0[key(i, j)];
  }
}

void energies(double *e) {
  need_engine("energy_eval");
  push_coords();
  check(sop_energy_eval(g_eng, g_sys, e), "sop_energy_eval");
}

} // namespace

extern "C" {

void alloc_GPU_arrays(void) {
  if (g_eng) sop_refshim_shutdown();
  const int n = nbead;
  if (n < 1 || !unc_pos) die("no beads loaded");
  // chains from the bond list: every bond joins consecutive beads (i, i+1)
  std::vector<int32_t> bonded(n, 0), chain(n, 0);
  std::vector<float> r0(n, 3.8f);
  for (int k = 0; k < nbnd; ++k) {
    int i = ibead_bnd[k] - 1, j = jbead_bnd[k] - 1;
    if (i > j) std::swap(i, j);
    if (i < 0 || j >= n || j != i + 1) die("bond " + std::to_string(k + 1) + " does not join beads i, i+1");
    if (bonded[i]) die("duplicate bond");
    bonded[i] = 1;
    r0[i] = pdb_dist[k];
  }
  for (int i = 1; i < n; ++i) chain[i] = chain[i - 1] + (bonded[i - 1] ? 0 : 1);
  // the i,i+2 soft-sphere term acts on every (i, i+1, i+2) of a chain
  int triples = 0;
  for (int i = 0; i + 2 < n; ++i) triples += bonded[i] && bonded[i + 1];
  for (int k = 0; k < nang; ++k) {
    const int i = ibead_ang[k] - 1, k3 = kbead_ang[k] - 1;
    const int lo = std::min(i, k3);
    if (lo < 0 || std::max(i, k3) != lo + 2 || !bonded[lo] || !bonded[lo + 1])
      die("angle " + std::to_string(k + 1) + " is not (i, i+1, i+2) inside a chain");
  }
  if (nang != triples) die("angle list must hold every (i, i+1, i+2) of each chain");
  // types and natives from the pair lists
  g_type.assign(n, 0);
  std::vector<int32_t> ni, nj;
  std::vector<float> nr;
  g_nat_r0.clear();
  for (int k = 0; k < ncon_att; ++k) {
    int i = get_idx(idx_bead_lj_nat[k].x) - 1, j = get_idx(idx_bead_lj_nat[k].y) - 1;
    set_type(i, get_type(idx_bead_lj_nat[k].x));
    set_type(j, get_type(idx_bead_lj_nat[k].y));
    if (i > j) std::swap(i, j);
    ni.push_back(i);
    nj.push_back(j);
    nr.push_back(lj_nat_pdb_dist[k]);
    g_nat_r0[key(i, j)] = lj_nat_pdb_dist[k];
  }
  for (int k = 0; k < ncon_rep; ++k) {
    set_type(get_idx(idx_bead_lj_non_nat[k].x) - 1, get_type(idx_bead_lj_non_nat[k].x));
    set_type(get_idx(idx_bead_lj_non_nat[k].y) - 1, get_type(idx_bead_lj_non_nat[k].y));
  }
  for (auto &t : g_type)
    if (t == 0) t = 1;
  // the non-native list must be the whole master list (every other pair)
  int64_t excluded = 0;
  for (int i = 0; i < n; ++i) excluded += (i + 1 < n && chain[i + 1] == chain[i]) + (i + 2 < n && chain[i + 2] == chain[i]);
  const int64_t master = (int64_t)n * (n - 1) / 2 - excluded - ncon_att;
  if (ncon_rep != master)
    die("the non-native list holds " + std::to_string(ncon_rep) + " pairs, the master list " + std::to_string(master));
  sop_system_desc d{};
  d.n_beads = n;
  d.xyz = &unc_pos[0].x;
  d.chain = chain.data();
  d.type = g_type.data();
  d.bond_r0 = r0.data();
  d.n_native = ncon_att;
  d.nat_i = ni.data();
  d.nat_j = nj.data();
  d.nat_r0 = nr.data();
  sop_params &P = d.params;
  sop_params_reference(&P);
  P.k_bnd = k_bnd;
  P.R0 = R0;
  P.eps_ss = epsilon_ss;
  P.sigma_ss = sigma_ss;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      P.eps_att[a][b] = force_coeff_att[a][b];
      P.eps_rep[a][b] = coeff_rep[a][b];
      P.sigma_rep[a][b] = sigma_rep[a][b];
    }
  P.T = T;
  P.zeta = zeta;
  P.h = h;
  P.boxl = boxl;
  P.n_nl = nnlup > 0 ? nnlup : 50;
  P.n_pl = 1;
  P.seed = g_seed;
  P.dynamics = sim_type == 1 ? SOP_DYN_UNDERDAMPED : SOP_DYN_OVERDAMPED;
  d.vel = (sim_type == 1 && vel) ? &vel[0].x : nullptr;
  check(sop_engine_create(deviceno, nullptr, &g_eng), "sop_engine_create");
  check(sop_system_add(g_eng, &d, &g_sys), "sop_system_add");
  g_counter = 0;
  g_incr.assign(n, sop_ref_float3{0, 0, 0});
}

void alloc_cudpp(void) {}

void setup_rng(unsigned long long seed, unsigned long long offset) {
  g_seed = seed;
  g_counter = (int64_t)offset;
}

void set_potential(void) {
  int iterm = 0;
  for (int i = 1; i <= 10; ++i) {
    if (!pot_term_on[i]) continue;
    if (i == 1) pot_term[++iterm] = &fene_energy;
    if (i == 2) pot_term[++iterm] = &soft_sphere_angular_energy;
    if (i == 5) pot_term[++iterm] = &vdw_energy;
  }
}

void clear_forces(void) {
  for (int i = 0; i < nbead; ++i) force[i] = sop_ref_float3{0, 0, 0};
}

void energy_eval(void) {
  double e[SOP_E_COUNT];
  energies(e);
  if (pot_term_on[1]) e_bnd = (float)e[SOP_E_BOND];
  if (pot_term_on[2]) e_ang_ss = (float)e[SOP_E_ANGLE];
  if (pot_term_on[5]) {
    e_vdw_rr_att = (float)e[SOP_E_NATIVE];
    e_vdw_rr_rep = (float)e[SOP_E_REP];
    e_vdw_rr = (float)(e[SOP_E_NATIVE] + e[SOP_E_REP]);
  }
  rna_etot = e_bnd + e_ang_ss + e_vdw_rr;
  system_etot = rna_etot + e_vdw_rc + e_vdw_cc;
}

void fene_energy(void) {
  double e[SOP_E_COUNT];
  energies(e);
  e_bnd = (float)e[SOP_E_BOND];
}

void soft_sphere_angular_energy(void) {
  double e[SOP_E_COUNT];
  energies(e);
  e_ang_ss = (float)e[SOP_E_ANGLE];
}

void vdw_energy(void) {
  double e[SOP_E_COUNT];
  energies(e);
  e_vdw_rr_att = (float)e[SOP_E_NATIVE];
  e_vdw_rr_rep = (float)e[SOP_E_REP];
  e_vdw_rr = (float)(e[SOP_E_NATIVE] + e[SOP_E_REP]);
}

void force_eval(void) {
  need_engine("force_eval");
  push_coords();
  check(sop_set_rng(g_eng, g_sys, g_seed, g_counter), "sop_set_rng");
  ++g_counter;
  check(sop_force_eval(g_eng, g_sys, 1), "sop_force_eval");
  check(sop_get_forces(g_eng, g_sys, &force[0].x), "sop_get_forces");
}

void random_force(void) {}

void update_neighbor_list(void) {
  need_engine("update_neighbor_list");
  push_coords();
  check(sop_update_neighbor_list(g_eng, g_sys), "sop_update_neighbor_list");
  copy_list(SOP_LIST_NL_NATIVE, &nnl_att, idx_neighbor_list_att, ncon_att, nl_lj_nat_pdb_dist);
  copy_list(SOP_LIST_NL_REP, &nnl_rep, idx_neighbor_list_rep, ncon_rep, nullptr);
}

void update_pair_list(void) {
  need_engine("update_pair_list");
  push_coords();
  check(sop_update_pair_list(g_eng, g_sys), "sop_update_pair_list");
  copy_list(SOP_LIST_PL_NATIVE, &nil_att, idx_pair_list_att, ncon_att, pl_lj_nat_pdb_dist);
  copy_list(SOP_LIST_PL_REP, &nil_rep, idx_pair_list_rep, ncon_rep, nullptr);
}

// The cell list of the reference yields the pair list directly; the hybrid
// list yields the neighbor list (sop.cu:1827-2608).
void update_cell_list(void) {
  update_neighbor_list();
  update_pair_list();
}

void update_hybrid_list(void) { update_neighbor_list(); }

void overdamped_iteration(sop_ref_float3 *incr) {
  for (int i = 0; i < nbead; ++i) {
    incr[i].x = a5 * force[i].x;
    incr[i].y = a5 * force[i].y;
    incr[i].z = a5 * force[i].z;
    unc_pos[i].x += incr[i].x;
    unc_pos[i].y += incr[i].y;
    unc_pos[i].z += incr[i].z;
    pos[i].x += incr[i].x;
    pos[i].y += incr[i].y;
    pos[i].z += incr[i].z;
    pos[i].x -= boxl * rnd(pos[i].x / boxl);
    pos[i].y -= boxl * rnd(pos[i].y / boxl);
    pos[i].z -= boxl * rnd(pos[i].z / boxl);
  }
  force_eval();
}

void underdamped_iteration(void) {
  need_engine("underdamped_iteration");
  for (int i = 0; i < nbead; ++i) {
    sop_ref_float3 &d = g_incr[i];
    d.x = a1 * vel[i].x + a2 * force[i].x;
    d.y = a1 * vel[i].y + a2 * force[i].y;
    d.z = a1 * vel[i].z + a2 * force[i].z;
    pos[i].x += d.x;
    pos[i].y += d.y;
    pos[i].z += d.z;
    pos[i].x -= boxl * rintf(pos[i].x / boxl);
    pos[i].y -= boxl * rintf(pos[i].y / boxl);
    pos[i].z -= boxl * rintf(pos[i].z / boxl);
    unc_pos[i].x += d.x;
    unc_pos[i].y += d.y;
    unc_pos[i].z += d.z;
  }
  force_eval();
  if (T < 1.0e-5f) return;  // steepest descent: velocities stay
  for (int i = 0; i < nbead; ++i) {
    vel[i].x = a3 * g_incr[i].x + a4 * force[i].x;
    vel[i].y = a3 * g_incr[i].y + a4 * force[i].y;
    vel[i].z = a3 * g_incr[i].z + a4 * force[i].z;
  }
}

void sop_refshim_shutdown(void) {
  if (g_eng) sop_engine_destroy(g_eng);
  g_eng = nullptr;
  g_sys = -1;
}

} // extern "C"
