/*
 * sop_refshim.h -- the reference's entry points on top of libsop.
 *
 * libsop_refshim.so exports the names a sop.cu-style driver calls, with the
 * reference's meaning, and reads / writes the reference's global variables,
 * which the driver defines (as global.cpp does).  Host arrays are the state
 * the driver sees; every call moves what it needs to and from one libsop
 * engine.  Errors print a message and exit(-1), as the reference does.
 *
 * Reference interface (samchogroup/SOP):
 *   energy.h:13-24   set_potential, clear_forces, energy_eval, force_eval,
 *                    random_force, fene_energy, soft_sphere_angular_energy,
 *                    vdw_energy
 *   sop.h:27-28      underdamped_iteration, overdamped_iteration
 *   sop.h:32-35      update_neighbor_list, update_cell_list,
 *                    update_hybrid_list, update_pair_list
 *   GPUvars.h:150-153 alloc_GPU_arrays, alloc_cudpp, setup_rng
 *   globals          global.h:83-233 (the subset below)
 *
 * The driver's float3 / FLOAT3 arrays must be 12-byte {x, y, z} triples and
 * its ushort2 arrays 4-byte {x, y} pairs, as in the reference (FLOAT = float).
 */
#ifndef SOP_REFSHIM_H_
#define SOP_REFSHIM_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sop_ref_float3 { float x, y, z; } sop_ref_float3;
typedef struct sop_ref_ushort2 { unsigned short x, y; } sop_ref_ushort2;

/* --- globals defined by the driver (global.h:83-233) --------------------- */
extern int deviceno;
/* bonded (1-based bead indices, io.cpp:126-173) */
extern float k_bnd, R0, R0sq;
extern int nbnd;
extern unsigned short *ibead_bnd, *jbead_bnd;
extern float *pdb_dist;
extern int nang;
extern unsigned short *ibead_ang, *jbead_ang, *kbead_ang;
extern float sigma_ss, epsilon_ss;
/* native / non-native pairs: COMBINE(type, 1-based index) (global.h:51-67) */
extern int ncon_att, ncon_rep;
extern sop_ref_ushort2 *idx_bead_lj_nat, *idx_bead_lj_non_nat;
extern float *lj_nat_pdb_dist;
extern int nnl_att, nnl_rep, nil_att, nil_rep;
extern sop_ref_ushort2 *idx_neighbor_list_att, *idx_neighbor_list_rep;
extern sop_ref_ushort2 *idx_pair_list_att, *idx_pair_list_rep;
extern float *nl_lj_nat_pdb_dist, *pl_lj_nat_pdb_dist;
extern float force_coeff_att[][3], coeff_rep[][3], sigma_rep[][3];
/* coordinates */
extern int nbead;
extern sop_ref_float3 *pos, *unc_pos, *vel, *force;
/* run parameters */
extern int sim_type; /* 1 underdamped, 2 overdamped */
extern float T, zeta, h, boxl;
extern float a1, a2, a3, a4, a5; /* integrator coefficients (param.cpp:100-110) */
extern int nnlup;
extern int pot_term_on[];
typedef void (*sop_ref_term_fn)(void);
extern sop_ref_term_fn pot_term[];
/* outputs */
extern float e_bnd, e_ang_ss, e_vdw_rr, e_vdw_rr_att, e_vdw_rr_rep;
extern float e_vdw_rc, e_vdw_cc, rna_etot, system_etot;

/* --- entry points --------------------------------------------------------- */
/* GPUvars.h:150-153: start the engine on `deviceno` and load the system held
 * in the globals (chains from the bond list, types from the pair lists, the
 * native list, the parameters); alloc_cudpp has nothing to do; setup_rng
 * sets the Philox key of the random force (offset: first counter). */
void alloc_GPU_arrays(void);
void alloc_cudpp(void);
void setup_rng(unsigned long long seed, unsigned long long offset);

/* energy.h:13-24 */
void set_potential(void);
void clear_forces(void);
void energy_eval(void);
/* forces at unc_pos over the current pair list plus the random force of the
 * next Philox counter, into `force` */
void force_eval(void);
/* the random force is drawn inside force_eval (as energy.cu:118 does); this
 * entry point has nothing left to do */
void random_force(void);
void fene_energy(void);
void soft_sphere_angular_energy(void);
void vdw_energy(void);

/* sop.h:32-35: lists at unc_pos; counts into nnl_*, nil_*; the lists
 * themselves into idx_neighbor_list_* / idx_pair_list_* (1-based, COMBINE'd)
 * when the driver allocated them with room for ncon_att / ncon_rep pairs */
void update_neighbor_list(void);
void update_pair_list(void);
void update_cell_list(void);
void update_hybrid_list(void);

/* sop.h:27-28: one step of the driver's integrator on the host arrays,
 * followed by force_eval (sop.cu:421-520) */
void overdamped_iteration(sop_ref_float3 *incr);
void underdamped_iteration(void);

/* Not in the reference: release the engine (the reference never frees). */
void sop_refshim_shutdown(void);

#ifdef __cplusplus
}
#endif

#endif /* SOP_REFSHIM_H_ */
