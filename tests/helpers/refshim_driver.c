/* A driver written the way the reference's sop.cu is: it owns the globals
 * (as global.cpp does), reads the reference's input files, computes the
 * integrator coefficients as param.cpp does and runs the control loop of
 * overdamped_ctrl / underdamped_ctrl (sop.cu:104-417) through the entry
 * points of libsop_refshim.  Test infrastructure only.
 *
 *   refshim_driver <dir> <sim_type 1|2> <nstep> <zeta> <h> <T> <nnlup> <seed> <out>
 *
 * <dir> holds bonds.dat, angles.dat, vdw.dat, init.dat.  <out> receives the
 * final unwrapped coordinates, energies, list counts and velocities. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sop_refshim.h"

int deviceno = 0;
float k_bnd = 20.0f, R0 = 2.0f, R0sq = 4.0f;
int nbnd, nang;
unsigned short *ibead_bnd, *jbead_bnd, *ibead_ang, *jbead_ang, *kbead_ang;
float *pdb_dist;
float sigma_ss = 3.5f, epsilon_ss = 1.0f;
int ncon_att, ncon_rep, nnl_att, nnl_rep, nil_att, nil_rep;
sop_ref_ushort2 *idx_bead_lj_nat, *idx_bead_lj_non_nat;
sop_ref_ushort2 *idx_neighbor_list_att, *idx_neighbor_list_rep, *idx_pair_list_att, *idx_pair_list_rep;
float *lj_nat_pdb_dist, *nl_lj_nat_pdb_dist, *pl_lj_nat_pdb_dist;
float force_coeff_att[3][3] = {{0, 0, 0}, {0, 1.0f, 0.8f}, {0, 0.8f, 0.7f}};
float coeff_rep[3][3] = {{0, 0, 0}, {0, 1.0f, 1.0f}, {0, 1.0f, 1.0f}};
float sigma_rep[3][3] = {{0, 0, 0}, {0, 3.8f, 5.4f}, {0, 5.4f, 7.0f}};
float rcut_nat[3][3] = {{0, 0, 0}, {0, 8.0f, 11.0f}, {0, 11.0f, 14.0f}};
int nbead;
sop_ref_float3 *pos, *unc_pos, *vel, *force;
int sim_type = 1;
float T = 0.6f, zeta = 0.05f, h = 2.5e-3f, boxl = 500.0f;
float halfh, a1, a2, a3, a4, a5;
int nnlup = 50, inlup = 0;
int pot_term_on[11] = {0, 1, 1, 0, 0, 1, 0, 0, 0, 0, 0};
sop_ref_term_fn pot_term[11];
float e_bnd, e_ang_ss, e_vdw_rr, e_vdw_rr_att, e_vdw_rr_rep, e_vdw_rc, e_vdw_cc, rna_etot, system_etot;

static FILE *open_in(const char *dir, const char *name) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  return f;
}

static void load(const char *dir) {
  char w[64], w2[64];
  FILE *f = open_in(dir, "bonds.dat");
  if (fscanf(f, "%63s %d", w, &nbnd) != 2) exit(2);
  ibead_bnd = malloc(2 * nbnd); jbead_bnd = malloc(2 * nbnd); pdb_dist = malloc(4 * nbnd);
  for (int k = 0; k < nbnd; ++k) {
    int i, j;
    if (fscanf(f, "%d %d %f", &i, &j, &pdb_dist[k]) != 3) exit(2);
    ibead_bnd[k] = (unsigned short)i; jbead_bnd[k] = (unsigned short)j;
  }
  fclose(f);
  f = open_in(dir, "angles.dat");
  if (fscanf(f, "%63s %d", w, &nang) != 2) exit(2);
  ibead_ang = malloc(2 * nang); jbead_ang = malloc(2 * nang); kbead_ang = malloc(2 * nang);
  for (int k = 0; k < nang; ++k) {
    int i, j, l;
    float a;
    if (fscanf(f, "%d %d %d %f", &i, &j, &l, &a) != 4) exit(2);
    ibead_ang[k] = (unsigned short)i; jbead_ang[k] = (unsigned short)j; kbead_ang[k] = (unsigned short)l;
  }
  fclose(f);
  f = open_in(dir, "vdw.dat");
  int natt, nrep;
  if (fscanf(f, "%63s %d %63s %d", w, &natt, w2, &nrep) != 4) exit(2);
  ncon_att = natt; ncon_rep = nrep;
  idx_bead_lj_nat = malloc(4 * (natt + 1)); lj_nat_pdb_dist = malloc(4 * (natt + 1));
  idx_bead_lj_non_nat = malloc(4 * (nrep + 1));
  idx_neighbor_list_att = malloc(4 * (natt + 1)); nl_lj_nat_pdb_dist = malloc(4 * (natt + 1));
  idx_pair_list_att = malloc(4 * (natt + 1)); pl_lj_nat_pdb_dist = malloc(4 * (natt + 1));
  idx_neighbor_list_rep = malloc(4 * (nrep + 1)); idx_pair_list_rep = malloc(4 * (nrep + 1));
  int ia = 0, ir = 0;
  for (int k = 0; k < natt + nrep; ++k) {
    int i, j, ti, tj;
    float r;
    if (fscanf(f, "%d %d %f %d %d", &i, &j, &r, &ti, &tj) != 5) exit(2);
    sop_ref_ushort2 c = {(unsigned short)((ti << 14) | i), (unsigned short)((tj << 14) | j)};
    if (r < rcut_nat[ti][tj]) { idx_bead_lj_nat[ia] = c; lj_nat_pdb_dist[ia++] = r; }
    else idx_bead_lj_non_nat[ir++] = c;
  }
  fclose(f);
  f = open_in(dir, "init.dat");
  if (fscanf(f, "%63s %d", w, &nbead) != 2) exit(2);
  pos = calloc(nbead, sizeof *pos); unc_pos = calloc(nbead, sizeof *unc_pos);
  vel = calloc(nbead, sizeof *vel); force = calloc(nbead, sizeof *force);
  for (int k = 0; k < nbead; ++k) {
    int i;
    if (fscanf(f, "%d %f %f %f", &i, &unc_pos[k].x, &unc_pos[k].y, &unc_pos[k].z) != 4) exit(2);
    pos[k] = unc_pos[k];
  }
  fclose(f);
}

int main(int argc, char **argv) {
  if (argc != 10) { fprintf(stderr, "usage: see source\n"); return 2; }
  load(argv[1]);
  sim_type = atoi(argv[2]);
  const double nstep = atof(argv[3]);
  zeta = (float)atof(argv[4]);
  h = (float)atof(argv[5]);
  T = (float)atof(argv[6]);
  nnlup = atoi(argv[7]);
  const unsigned long long seed = strtoull(argv[8], NULL, 10);
  /* param.cpp:100-110 */
  halfh = (float)(h / 2.0);
  a1 = (float)(h * (1.0 - zeta * halfh));
  a2 = h * halfh;
  a3 = (float)((1.0 - h * zeta / 2.0 + (h * zeta) * (h * zeta) / 4.0) / h);
  a4 = (float)(halfh * (1.0 - h * zeta / 2.0));
  a5 = h / zeta;

  alloc_GPU_arrays();
  alloc_cudpp();
  setup_rng(seed, 0);
  sop_ref_float3 *incr = calloc(nbead, sizeof *incr);
  for (int i = 0; i < nbead; ++i) { vel[i].x = vel[i].y = vel[i].z = 0; force[i] = vel[i]; }
  update_neighbor_list();
  update_pair_list();
  set_potential();
  energy_eval();
  force_eval();
  int lup = sim_type == 1 ? 1 : inlup;  /* underdamped: local counter from 1; overdamped: global from 0 */
  for (double istep = 1.0; istep <= nstep; istep += 1.0) {
    if (lup % nnlup == 0) {
      update_neighbor_list();
      lup = 0;
    }
    lup++;
    update_pair_list();
    if (sim_type == 1) underdamped_iteration();
    else overdamped_iteration(incr);
  }
  energy_eval();
  FILE *o = fopen(argv[9], "w");
  if (!o) return 2;
  for (int i = 0; i < nbead; ++i) fprintf(o, "%.9g %.9g %.9g\n", unc_pos[i].x, unc_pos[i].y, unc_pos[i].z);
  fprintf(o, "%.9g %.9g %.9g %.9g %.9g\n", e_bnd, e_ang_ss, e_vdw_rr_att, e_vdw_rr_rep, rna_etot);
  fprintf(o, "%d %d %d %d\n", nnl_att, nnl_rep, nil_att, nil_rep);
  for (int i = 0; i < nbead; ++i) fprintf(o, "%.9g %.9g %.9g\n", vel[i].x, vel[i].y, vel[i].z);
  for (int k = 0; k < nil_rep && k < 5; ++k) fprintf(o, "%u %u\n", idx_pair_list_rep[k].x, idx_pair_list_rep[k].y);
  fclose(o);
  sop_refshim_shutdown();
  return 0;
}
