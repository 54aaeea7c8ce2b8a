// serve.cpp -- continuous-batching serve loop (see sop.h, "serving").
#include "engine.h"

extern "C" {
sop_status sop_serve_start(sop_engine *e, const sop_serve_opts *o) { (void)o; return e ? SOP_ERR_STATE : SOP_ERR_INVALID; }
sop_status sop_serve_submit(sop_engine *e, const sop_system_desc *d, int64_t nsteps, int64_t *job) {
  (void)d; (void)nsteps; (void)job; return e ? SOP_ERR_STATE : SOP_ERR_INVALID; }
sop_status sop_serve_poll(sop_engine *e, int64_t job, sop_job_result *r, float *xyz) {
  (void)job; (void)r; (void)xyz; return e ? SOP_ERR_STATE : SOP_ERR_INVALID; }
sop_status sop_serve_stop(sop_engine *e) { return e ? SOP_OK : SOP_ERR_INVALID; }
}
