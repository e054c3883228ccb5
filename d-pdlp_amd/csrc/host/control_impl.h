/*
 * Bodies of the reference's per-LP scalar decisions, written once and
 * compiled twice: as the host controller (control.c: CTL_FN empty,
 * CTL(x) = ctl_x) and inside the persistent kernel (persist.hip: CTL_FN
 * __device__ inline, CTL(x) = dctl_x), so the device makes the decisions with
 * the same arithmetic as the host.  S is an LP's array of raw reduction slots
 * (DPDLP_SL_* / DPDLP_SC_* of dpdlp_dev.h).
 */
#ifndef CTL_FN
#error "define CTL_FN and CTL(x) before including control_impl.h"
#endif

/* blue_combine of dev_common.h: Blue's three scaled sums of squares to a norm */
#define CTL_SSML 0x1p+537
#define CTL_SBIG 0x1p-538
CTL_FN double CTL(blue)(const double *a) {
  double asml = a[0], amed = a[1], abig = a[2];
  double scl, sumsq;
  if (abig > 0.0) {
    if (amed > 0.0 || isnan(amed)) abig += (amed * CTL_SBIG) * CTL_SBIG;
    scl = 1.0 / CTL_SBIG;
    sumsq = abig;
  } else if (asml > 0.0) {
    if (amed > 0.0 || isnan(amed)) {
      double ymed = sqrt(amed), ysml = sqrt(asml) / CTL_SSML;
      double ymin = ysml > ymed ? ymed : ysml, ymax = ysml > ymed ? ysml : ymed;
      double r = ymin / ymax;
      scl = 1.0;
      sumsq = ymax * ymax * (1.0 + r * r);
    } else {
      scl = 1.0 / CTL_SSML;
      sumsq = asml;
    }
  } else {
    scl = 1.0;
    sumsq = amed;
  }
  return scl * sqrt(sumsq);
}

/* utils.cu:665-718 from the raw slots */
CTL_FN void CTL(residual)(hstate_t *h, const lpk_t *c, const double *S, norm_type_t norm) {
  h->abs_p = (norm == NORM_TYPE_L_INF ? S[DPDLP_SL_RESP + 7] : CTL(blue)(S + DPDLP_SL_RESP)) / c->beta;
  h->abs_d = (norm == NORM_TYPE_L_INF ? S[DPDLP_SL_RESD + 8] : CTL(blue)(S + DPDLP_SL_RESD)) / c->gamma;
  h->pobj = S[DPDLP_SL_RESD + 3] / (c->beta * c->gamma) + c->c0;
  h->dobj = (S[DPDLP_SL_RESD + 4] + S[DPDLP_SL_RESP + 3]) / (c->beta * c->gamma) + c->c0;
  h->rel_p = h->abs_p / (1.0 + c->bound_norm);
  h->rel_d = h->abs_d / (1.0 + c->obj_norm);
  h->gap = fabs(h->pobj - h->dobj);
  h->rel_gap = h->gap / (1.0 + fabs(h->pobj) + fabs(h->dobj));
}

/* utils.cu:760-812 from the raw slots */
CTL_FN void CTL(infeasibility)(hstate_t *h, const lpk_t *c, const double *S0) {
  const double *S = S0 + DPDLP_SL_INF;
  double d_inf = S[1];
  h->p_ray_lin_obj = S[5] / (c->beta * c->gamma);
  h->d_ray_obj = (S[2] + S[4]) / (c->beta * c->gamma);
  h->max_p_ray_inf = S[3];
  double ds = S[6];
  h->max_d_ray_inf = ds;
  double sf = fmax(d_inf, ds);
  if (sf > 0.0) {
    h->max_d_ray_inf /= sf;
    h->d_ray_obj /= sf;
  } else {
    h->max_d_ray_inf = 0.0;
    h->d_ray_obj = 0.0;
  }
}

/* pdlp_core_op.cu:687-695 */
CTL_FN double CTL(fpe_value)(const double *G, double omega, double step) {
  double pn = CTL(blue)(G + 0), dn = CTL(blue)(G + 3);
  double movement = pn * pn * omega + dn * dn / omega;
  double interaction = 2 * step * G[6];
  return sqrt(movement + interaction);
}

CTL_FN void CTL(pending_fpe)(hstate_t *h, const double *S, int polish) {
  if (h->pend_rst) {
    const double *G = S + DPDLP_SC_FPE_RESTART;
    double v;
    if (polish == 1) {
      double pn = CTL(blue)(G);
      v = pn * pn * h->omega;
    } else if (polish == 2) {
      double dn = CTL(blue)(G + 3);
      v = dn * dn / h->omega;
    } else {
      v = CTL(fpe_value)(G, h->omega, h->step);
    }
    h->fpe = v;
    h->fpe0 = v;
    h->pend_rst = 0;
  }
  if (h->pend_maj) {
    const double *G = S + DPDLP_SC_FPE_MAJOR;
    if (polish == 1) {
      double pn = CTL(blue)(G);
      h->fpe = pn * pn * h->omega;
    } else if (polish == 2) {
      double dn = CTL(blue)(G + 3);
      h->fpe = dn * dn / h->omega;
    } else {
      h->fpe = CTL(fpe_value)(G, h->omega, h->step);
    }
    h->pend_maj = 0;
  }
}

/* utils.cu:214-263 */
CTL_FN void CTL(check_termination)(hstate_t *h, const termination_criteria_t *tc) {
  if (h->rel_d < tc->eps_feasible_relative && h->rel_p < tc->eps_feasible_relative &&
      h->rel_gap < tc->eps_optimal_relative) {
    h->reason = TERMINATION_REASON_OPTIMAL;
    return;
  }
  if (h->d_ray_obj > 0.0 && h->max_d_ray_inf / h->d_ray_obj <= tc->eps_infeasible) {
    h->reason = TERMINATION_REASON_PRIMAL_INFEASIBLE;
    return;
  }
  if (h->p_ray_lin_obj < 0.0 && h->max_p_ray_inf / (-h->p_ray_lin_obj) <= tc->eps_infeasible) {
    h->reason = TERMINATION_REASON_DUAL_INFEASIBLE;
    return;
  }
  if (h->total >= tc->iteration_limit) {
    h->reason = TERMINATION_REASON_ITERATION_LIMIT;
    return;
  }
  if (h->t_cum >= tc->time_sec_limit) h->reason = TERMINATION_REASON_TIME_LIMIT;
}

/* utils.cu:1015-1039 */
CTL_FN void CTL(check_polish_termination)(hstate_t *h, const termination_criteria_t *tc,
                                          int primal) {
  if (primal ? (h->rel_p <= tc->eps_feas_polish_relative)
             : (h->rel_d <= tc->eps_feas_polish_relative)) {
    h->reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
    return;
  }
  if (h->total >= tc->iteration_limit) {
    h->reason = TERMINATION_REASON_ITERATION_LIMIT;
    return;
  }
  if (h->t_cum >= tc->time_sec_limit) h->reason = TERMINATION_REASON_TIME_LIMIT;
}

/* utils.cu:265-293 */
CTL_FN int CTL(should_restart)(hstate_t *h, const restart_parameters_t *rp, int freq) {
  int r = 0;
  if (h->total == freq) {
    r = 1;
  } else if (h->total > freq) {
    if (h->fpe <= rp->sufficient_reduction_for_restart * h->fpe0) r = 1;
    if (h->fpe <= rp->necessary_reduction_for_restart * h->fpe0 && h->fpe > h->fpe_last) r = 1;
    if (h->inner >= rp->artificial_restart_threshold * h->total) r = 1;
  }
  h->fpe_last = h->fpe;
  return r;
}

/* PID primal weight of pdlp_core_op.cu:593-619 */
CTL_FN void CTL(pid_update)(hstate_t *h, const restart_parameters_t *rp, double pd, double dd) {
  double ratio = h->rel_d / h->rel_p;
  if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
    double err = log(dd) - log(pd) - log(h->omega);
    h->pw_err_sum *= rp->i_smooth;
    h->pw_err_sum += err;
    double derr = err - h->pw_last_err;
    h->omega *= exp(rp->k_p * err + rp->k_i * h->pw_err_sum + rp->k_d * derr);
    h->pw_last_err = err;
  } else {
    h->omega = h->best_omega;
    h->pw_err_sum = 0.0;
    h->pw_last_err = 0.0;
  }
  double g = fabs(log10(h->rel_d / h->rel_p));
  if (g < h->best_gap) {
    h->best_gap = g;
    h->best_omega = h->omega;
  }
}

CTL_FN int CTL(mixed_keep_fp32)(hstate_t *h, double kkt, double switch_tol, double eps,
                                int anchored) {
  if (kkt < switch_tol) return 0;
  if (kkt <= 0.9 * h->m32_best) {
    h->m32_best = kkt;
    h->m32_best_T = h->total;
    return 1;
  }
  /* anchored shadows have no rounding floor at the iterates' scale: only
   * a much longer stall is taken for one */
  int window = anchored ? (h->total / 2 > 800 ? h->total / 2 : 800)
                        : (h->total / 4 > 200 ? h->total / 4 : 200);
  int far = anchored ? window + window / 2 : 2 * window;
  int since = h->total - h->m32_best_T;
  return !((h->m32_best < 100.0 * eps && since >= window) || since >= far);
}

/* get_print_frequency (utils.cu:492-501): the evaluation cadence */
CTL_FN int CTL(print_frequency)(int iter) {
  int step = 10;
  long long threshold = 1000;
  while (iter >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}

/* initial controller state once the step size is known */
CTL_FN void CTL(init_state)(hstate_t *h, const lpk_t *k, const pdhg_parameters_t *params,
                            double step) {
  hstate_t z = {0};
  *h = z;
  h->step = step;
  h->omega = params->bound_objective_rescaling ? 1.0 : (k->obj_norm + 1.0) / (k->bound_norm + 1.0);
  h->best_omega = h->omega;
  h->best_gap = INFINITY;
  h->fpe_last = INFINITY;
  h->reason = TERMINATION_REASON_UNSPECIFIED;
  h->m32_best = INFINITY;
}
