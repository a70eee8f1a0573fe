/*
 * mps_oracle.h -- C interface of the CPU oracle for the explicit MPS step.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP
 * implementation and the CPU baseline that bench.py times beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product library never links or calls it.
 *
 * The functions mirror the product's C ABI (include/mps.h) so that a test
 * drives both the same way.  Semantics and argument layouts are identical.
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#include "../include/mps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_handle orc_handle;

int orc_create(const mps_params* params, const mps_scene* scene, orc_handle** out);
int orc_destroy(orc_handle* h);
int orc_set_threads(int nthreads);
int orc_step(orc_handle* h, int nsteps);
int orc_phase(orc_handle* h, int phase);
int orc_get_constants(orc_handle* h, mps_constants* out);
int orc_get_state(orc_handle* h, int8_t* typ, double* pos, double* vel, double* acc,
                  double* prs, double* pav, double* nd);
int orc_set_state(orc_handle* h, const int8_t* typ, const double* pos, const double* vel,
                  const double* prs, const double* pav);
int orc_get_edge(orc_handle* h, int8_t* edge);
int orc_get_buckets(orc_handle* h, int which, int* cell_start, int* ids, int* count);
int orc_get_mirrors(orc_handle* h, int* ids, double* pos, double* prs, int* count);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
