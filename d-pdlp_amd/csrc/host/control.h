/*
 * The reference's per-LP scalar decisions, computed on the host from the raw
 * device reduction slots: residuals and objectives (utils.cu:639-719),
 * infeasibility rays (utils.cu:721-813), fixed-point error
 * (pdlp_core_op.cu:658-696), termination (utils.cu:214-263), the restart test
 * (utils.cu:265-293) and the PID primal weight (pdlp_core_op.cu:593-619).
 * Shared by the single-LP driver and the batched driver.
 */
#pragma once

#include "cupdlpx_types.h"
#include "dpdlp_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

double dpdlp_blue(const double *three_sums);
void ctl_residual(hstate_t *h, const lpk_t *k, const double *S, norm_type_t norm);
void ctl_infeasibility(hstate_t *h, const lpk_t *k, const double *S);
void ctl_pending_fpe(hstate_t *h, const double *S, int polish);
void ctl_check_termination(hstate_t *h, const termination_criteria_t *tc);
void ctl_check_polish_termination(hstate_t *h, const termination_criteria_t *tc, int primal);
int ctl_should_restart(hstate_t *h, const restart_parameters_t *rp, int freq);
void ctl_pid_update(hstate_t *h, const restart_parameters_t *rp, double pd, double dd);
/* Mixed precision: 1 while FP32 iterations should continue.  FP32 ends once
 * the max relative KKT error is below switch_tol, or once the best value
 * seen in FP32 has not improved by 10% for W = max(200, T/4) iterations
 * while it is within 100x of the target eps (the FP32 rounding floor of the
 * residuals is near the target), or for 2W iterations wherever it is (a
 * floor far above the target: FP32 then never gets there).  With anchored
 * shadows (FP32 products of the distance from the last restart point) the
 * rounding has no floor at the iterates' scale, and plateaus of the method
 * itself must not end FP32: W = max(800, T/2) and the far stall 3W/2. */
int ctl_mixed_keep_fp32(hstate_t *h, double kkt, double switch_tol, double eps, int anchored);
double ctl_blue(const double *three_sums);
double ctl_fpe_value(const double *G, double omega, double step);
int ctl_print_frequency(int iter);
void ctl_init_state(hstate_t *h, const lpk_t *k, const pdhg_parameters_t *params, double step);

#ifdef __cplusplus
}
#endif
