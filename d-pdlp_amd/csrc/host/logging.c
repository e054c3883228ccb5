/*
 * Console output in the reference's formats so logs stay comparable:
 * termination names (src/utils.cu:191-212), settings banner (:352-421),
 * iteration table (pdlp_core_op.cu:376-388, utils.cu:478-490), final summary
 * (utils.cu:427-476) and the polishing logs (utils.cu:1041-1108).
 */
#include "dpdlp_host.h"

#include <math.h>
#include <stdio.h>

const char *dpdlp_termination_string(termination_reason_t r) {
  switch (r) {
  case TERMINATION_REASON_OPTIMAL: return "OPTIMAL";
  case TERMINATION_REASON_PRIMAL_INFEASIBLE: return "PRIMAL_INFEASIBLE";
  case TERMINATION_REASON_DUAL_INFEASIBLE: return "DUAL_INFEASIBLE";
  case TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED: return "INFEASIBLE_OR_UNBOUNDED";
  case TERMINATION_REASON_TIME_LIMIT: return "TIME_LIMIT";
  case TERMINATION_REASON_ITERATION_LIMIT: return "ITERATION_LIMIT";
  case TERMINATION_REASON_UNSPECIFIED: return "UNSPECIFIED";
  case TERMINATION_REASON_FEAS_POLISH_SUCCESS: return "FEAS_POLISH_SUCCESS";
  default: return "UNKNOWN";
  }
}

/* utils.cu:492-501 */
int dpdlp_print_frequency(int iter) {
  int step = 10;
  long long threshold = 1000;
  while (iter >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}

static const char *RULE =
    "---------------------------------------------------------------------------------------";

void dpdlp_print_initial_info(const pdhg_parameters_t *p, const lp_problem_t *prob) {
  if (!p->verbose) return;
  printf("%s\n", RULE);
  printf("  PDHG (PDLP) linear-programming solver for one AMD MI355X (gfx950)\n");
  printf("  restarted reflected-Halpern PDHG, hand-written HIP kernels\n");
  printf("%s\n", RULE);
  printf("problem: %d rows, %d columns, %d nonzeros\n", prob->num_constraints,
         prob->num_variables, prob->constraint_matrix_num_nonzeros);
  printf("settings:\n");
  printf("  iter_limit         : %d\n", p->termination_criteria.iteration_limit);
  printf("  time_limit         : %.2f sec\n", p->termination_criteria.time_sec_limit);
  printf("  eps_opt            : %.1e\n", p->termination_criteria.eps_optimal_relative);
  printf("  eps_feas           : %.1e\n", p->termination_criteria.eps_feasible_relative);
  printf("  eps_infeas_detect  : %.1e\n", p->termination_criteria.eps_infeasible);
  if (p->optimality_norm != NORM_TYPE_L2) printf("  optimality_norm    : L_inf\n");
  if (p->l_inf_ruiz_iterations != 10)
    printf("  %-18s : %d\n", "l_inf_ruiz_iter", p->l_inf_ruiz_iterations);
  if (fabs(p->pock_chambolle_alpha - 1.0) > 1e-9)
    printf("  %-18s : %.1e\n", "pock_chambolle_alpha", p->pock_chambolle_alpha);
  if (!p->has_pock_chambolle_alpha) printf("  %-18s : %s\n", "has_pock_chambolle_alpha", "off");
  if (!p->bound_objective_rescaling) printf("  %-18s : %s\n", "bound_obj_rescaling", "off");
  if (p->sv_max_iter != 5000) printf("  %-18s : %d\n", "sv_max_iter", p->sv_max_iter);
  if (fabs(p->sv_tol - 1e-4) > 1e-9) printf("  %-18s : %.1e\n", "sv_tol", p->sv_tol);
  if (p->termination_evaluation_frequency != 200)
    printf("  %-18s : %d\n", "evaluation_freq", p->termination_evaluation_frequency);
  if (p->feasibility_polishing) printf("  %-18s : %s\n", "feasibility_polishing", "on");
  if (fabs(p->termination_criteria.eps_feas_polish_relative - 1e-6) > 1e-9)
    printf("  %-18s : %.1e\n", "eps_feas_polish_relative",
           p->termination_criteria.eps_feas_polish_relative);
  if (p->presolve) printf("  presolve           : unavailable, solving without it\n");
}

void dpdlp_print_table_header(void) {
  printf("%s\n", RULE);
  printf("%s | %s | %s | %s \n", "   runtime    ", "    objective     ",
         "  absolute residuals   ", "  relative residuals   ");
  printf("%s %s | %s %s | %s %s %s | %s %s %s \n", "  iter", "  time ", " pr obj ", "  du obj ",
         " pr res", " du res", "  gap  ", " pr res", " du res", "  gap  ");
  printf("%s\n", RULE);
}

void dpdlp_print_iteration(int total, double t, double pobj, double dobj, double abs_p,
                           double abs_d, double gap, double rel_p, double rel_d,
                           double rel_gap) {
  if (total % dpdlp_print_frequency(total) != 0) return;
  printf("%6d %.1e | %8.1e  %8.1e | %.1e %.1e %.1e | %.1e %.1e %.1e \n", total, t, pobj, dobj,
         abs_p, abs_d, gap, rel_p, rel_d, rel_gap);
}

void dpdlp_print_final(const cupdlpx_result_t *r, const pdhg_parameters_t *p) {
  if (p->verbose) printf("%s\n", RULE);
  printf("Solution Summary\n");
  printf("  Status             : %s\n", dpdlp_termination_string(r->termination_reason));
  printf("  Solve time         : %.3g sec\n", r->cumulative_time_sec);
  printf("  Iterations         : %d\n", r->total_count);
  printf("  Primal objective   : %.10g\n", r->primal_objective_value);
  printf("  Dual objective     : %.10g\n", r->dual_objective_value);
  printf("  Objective gap      : %.3e\n", r->relative_objective_gap);
  printf("  Primal infeas      : %.3e\n", r->relative_primal_residual);
  printf("  Dual infeas        : %.3e\n", r->relative_dual_residual);
  fflush(stdout);
}

void dpdlp_print_polish_header(int primal, const pdhg_parameters_t *p) {
  if (!p->verbose) return;
  printf("%s\n", RULE);
  printf("Starting %s Feasibility Polishing Phase with relative tolerance %.2e\n",
         primal ? "Primal" : "Dual", p->termination_criteria.eps_feas_polish_relative);
  printf("%s\n", RULE);
  if (primal)
    printf("%s %s |  %s  | %s | %s \n", "  iter", "  time ", "pr obj", " abs pr res ",
           " rel pr res ");
  else
    printf("%s %s |  %s  | %s | %s \n", "  iter", "  time ", "du obj", " abs du res ",
           " rel du res ");
  printf("%s\n", RULE);
}

void dpdlp_print_polish_iteration(int primal, int total, double t, double obj, double abs_r,
                                  double rel_r) {
  (void)primal;
  if (total % dpdlp_print_frequency(total) != 0) return;
  printf("%6d %.1e | %8.1e |    %.1e   |   %.1e   \n", total, t, obj, abs_r, rel_r);
}

void dpdlp_print_polish_final(termination_reason_t pr, int piters, double ptime,
                              termination_reason_t dr, int diters, double dtime,
                              double prel, double drel, double gap) {
  printf("Feasibility Polishing Summary\n");
  printf("  Primal Status        : %s\n", dpdlp_termination_string(pr));
  printf("  Primal Iterations    : %d\n", piters);
  printf("  Primal Time Usage    : %.3g sec\n", ptime);
  printf("  Dual Status          : %s\n", dpdlp_termination_string(dr));
  printf("  Dual Iterations      : %d\n", diters);
  printf("  Dual Time Usage      : %.3g sec\n", dtime);
  printf("  Primal Residual      : %.3e\n", prel);
  printf("  Dual Residual        : %.3e\n", drel);
  printf("  Primal Dual Gap      : %.3e\n", gap);
}
