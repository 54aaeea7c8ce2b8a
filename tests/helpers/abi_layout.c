/* Prints the size and field offsets of the C ABI structs, for the ctypes
 * layout test. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/sop.h"

#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))

int main(void) {
  printf("sop_params %zu\n", sizeof(sop_params));
  F(sop_params, nl_factor); F(sop_params, pl_factor); F(sop_params, n_nl); F(sop_params, T);
  F(sop_params, boxl); F(sop_params, seed); F(sop_params, dynamics);
  printf("sop_system_desc %zu\n", sizeof(sop_system_desc));
  F(sop_system_desc, n_native); F(sop_system_desc, params); F(sop_system_desc, step0); F(sop_system_desc, vel);
  printf("sop_engine_opts %zu\n", sizeof(sop_engine_opts));
  F(sop_engine_opts, lds_beads_max); F(sop_engine_opts, pool_bytes); F(sop_engine_opts, step_tiles);
  printf("sop_serve_opts %zu\n", sizeof(sop_serve_opts));
  printf("sop_job_result %zu\n", sizeof(sop_job_result));
  F(sop_job_result, steps); F(sop_job_result, t_submit); F(sop_job_result, energies);
  printf("sop_list_stats %zu\n", sizeof(sop_list_stats));
  printf("sop_gen_opts %zu\n", sizeof(sop_gen_opts));
  printf("sop_tune_result %zu\n", sizeof(sop_tune_result));
  F(sop_tune_result, pl_shell); F(sop_tune_result, block_ms_start); F(sop_tune_result, trials);
  printf("sop_serve_stats %zu\n", sizeof(sop_serve_stats));
  F(sop_serve_stats, gpu_block_s); F(sop_serve_stats, mean_running); F(sop_serve_stats, bead_steps);
  return 0;
}
