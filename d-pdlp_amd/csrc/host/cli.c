/*
 * cupdlpx-dist command line, single-GPU mode (src/cli.c:32-350 of the
 * reference): same options, same output files and formats:
 *   <dir>/<name>_summary.txt, <name>_primal_solution.txt,
 *   <name>_dual_solution.txt   (one %.10g per line)
 * where <name> is the input's basename up to its first '.'.
 * The MPI / multi-GPU branch of the reference (cli.c:352-640) is not built:
 * this solver runs on one GPU.  Exit status: 1 on bad arguments or an
 * unreadable input, 0 otherwise (also when the solve itself fails, as in the
 * reference).
 */
#define _GNU_SOURCE
#include "cupdlpx.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "mps_parser.h"

#include <getopt.h>
#include <libgen.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char *out_path(const char *dir, const char *name, const char *suffix) {
  size_t len = strlen(dir) + strlen(name) + strlen(suffix) + 2;
  char *p = (char *)malloc(len);
  if (p) snprintf(p, len, "%s/%s%s", dir, name, suffix);
  return p;
}

/* cli.c:41-57: basename up to the first dot */
static char *instance_name(const char *filename) {
  char *copy = strdup(filename);
  if (!copy) return NULL;
  char *base = basename(copy);
  char *dot = strchr(base, '.');
  if (dot) *dot = '\0';
  char *name = strdup(base);
  free(copy);
  return name;
}

/* cli.c:59-79 */
static void save_vector(const double *data, int size, const char *dir, const char *name,
                        const char *suffix) {
  char *path = out_path(dir, name, suffix);
  if (!path || !data) {
    free(path);
    return;
  }
  if (dpdlp_write_vector(path, data, size)) perror("Error opening solution file");
  free(path);
}

/* cli.c:81-154 (the presolve block never prints: no presolver here) */
static void save_summary(const cupdlpx_result_t *r, const char *dir, const char *name) {
  char *path = out_path(dir, name, "_summary.txt");
  if (!path) return;
  FILE *f = fopen(path, "w");
  if (!f) {
    perror("Error opening summary file");
    free(path);
    return;
  }
  fprintf(f, "Termination Reason: %s\n", dpdlp_termination_string(r->termination_reason));
  fprintf(f, "Runtime (sec): %e\n", r->cumulative_time_sec);
  fprintf(f, "Iterations Count: %d\n", r->total_count);
  fprintf(f, "Primal Objective Value: %e\n", r->primal_objective_value);
  fprintf(f, "Dual Objective Value: %e\n", r->dual_objective_value);
  fprintf(f, "Relative Primal Residual: %e\n", r->relative_primal_residual);
  fprintf(f, "Relative Dual Residual: %e\n", r->relative_dual_residual);
  fprintf(f, "Absolute Objective Gap: %e\n", r->objective_gap);
  fprintf(f, "Relative Objective Gap: %e\n", r->relative_objective_gap);
  fprintf(f, "Rows: %d\n", r->num_constraints);
  fprintf(f, "Columns: %d\n", r->num_variables);
  fprintf(f, "Nonzeros: %d\n", r->num_nonzeros);
  if (r->feasibility_polishing_time > 0.0) {
    fprintf(f, "Feasibility Polishing Time (sec): %e\n", r->feasibility_polishing_time);
    fprintf(f, "Feasibility Polishing Iteration Count: %d\n", r->feasibility_iteration);
  }
  fclose(f);
  free(path);
}

static void usage(const char *prog) {
  fprintf(stderr, "Usage: %s [OPTIONS] <mps_file> <output_dir>\n\n", prog);
  fprintf(stderr, "Arguments:\n");
  fprintf(stderr, "  <mps_file>               Path to the input problem in MPS format (.mps or .mps.gz).\n");
  fprintf(stderr, "  <output_dir>             Directory where output files will be saved. It will contain:\n");
  fprintf(stderr, "                             - <basename>_summary.txt\n");
  fprintf(stderr, "                             - <basename>_primal_solution.txt\n");
  fprintf(stderr, "                             - <basename>_dual_solution.txt\n\n");
  fprintf(stderr, "Options:\n");
  fprintf(stderr, "  -h, --help                          Display this help message.\n");
  fprintf(stderr, "  -v, --verbose                       Enable verbose logging (default: false).\n");
  fprintf(stderr, "      --time_limit <seconds>          Time limit in seconds (default: 3600.0).\n");
  fprintf(stderr, "      --iter_limit <iterations>       Iteration limit (default: %d).\n", INT32_MAX);
  fprintf(stderr, "      --eps_opt <tolerance>           Relative optimality tolerance (default: 1e-4).\n");
  fprintf(stderr, "      --eps_feas <tolerance>          Relative feasibility tolerance (default: 1e-4).\n");
  fprintf(stderr, "      --eps_infeas_detect <tolerance> Infeasibility detection tolerance (default: 1e-10).\n");
  fprintf(stderr, "      --l_inf_ruiz_iter <int>         Iterations for L-inf Ruiz rescaling (default: 10).\n");
  fprintf(stderr, "      --no_pock_chambolle             Disable Pock-Chambolle rescaling (default: enabled).\n");
  fprintf(stderr, "      --pock_chambolle_alpha <float>  Value for Pock-Chambolle alpha (default: 1.0).\n");
  fprintf(stderr, "      --no_bound_obj_rescaling        Disable bound objective rescaling (default: enabled).\n");
  fprintf(stderr, "      --eval_freq <int>               Termination evaluation frequency (default: 200).\n");
  fprintf(stderr, "      --sv_max_iter <int>             Max iterations for singular value estimation (default: 5000).\n");
  fprintf(stderr, "      --sv_tol <float>                Tolerance for singular value estimation (default: 1e-4).\n");
  fprintf(stderr, "  -f  --feasibility_polishing         Enable feasibility polishing (default: false).\n");
  fprintf(stderr, "      --eps_feas_polish <tolerance>   Relative feasibility polish tolerance (default: 1e-6).\n");
  fprintf(stderr, "      --opt_norm <norm_type>          Norm for optimality criteria: l2 or linf (default: l2).\n");
  fprintf(stderr, "      --no_presolve                   Disable presolve (default: enabled; no presolver here).\n");
  fprintf(stderr, "      --mixed_precision               FP32 matrix / FP64 state iterations (extension).\n");
  fprintf(stderr, "      --grid_size <r,c>               Process grid; one GPU here, so only 1,1.\n");
  fprintf(stderr, "      --partition_method <m>          uniform or nnz (accepted; no partition on one GPU).\n");
  fprintf(stderr, "      --permute_method <m>            none, random or block (accepted; no effect here).\n");
  fprintf(stderr, "      --permute_block_size <int>      Permutation block size (accepted; no effect here).\n");
}

int main(int argc, char *argv[]) {
  pdhg_parameters_t params;
  set_default_parameters(&params);
  dpdlp_solve_opts_t opts;
  dpdlp_default_opts(&opts);
  static struct option lo[] = {{"help", no_argument, 0, 'h'},
                               {"verbose", no_argument, 0, 'v'},
                               {"time_limit", required_argument, 0, 1001},
                               {"iter_limit", required_argument, 0, 1002},
                               {"eps_opt", required_argument, 0, 1003},
                               {"eps_feas", required_argument, 0, 1004},
                               {"eps_infeas_detect", required_argument, 0, 1005},
                               {"eps_feas_polish", required_argument, 0, 1006},
                               {"feasibility_polishing", no_argument, 0, 'f'},
                               {"l_inf_ruiz_iter", required_argument, 0, 1007},
                               {"pock_chambolle_alpha", required_argument, 0, 1008},
                               {"no_pock_chambolle", no_argument, 0, 1009},
                               {"no_bound_obj_rescaling", no_argument, 0, 1010},
                               {"sv_max_iter", required_argument, 0, 1011},
                               {"sv_tol", required_argument, 0, 1012},
                               {"eval_freq", required_argument, 0, 1013},
                               {"opt_norm", required_argument, 0, 1014},
                               {"no_presolve", no_argument, 0, 1015},
                               {"mixed_precision", no_argument, 0, 1100},
                               {"grid_size", required_argument, 0, 2001},
                               {"partition_method", required_argument, 0, 2002},
                               {"permute_method", required_argument, 0, 2003},
                               {"permute_block_size", required_argument, 0, 2004},
                               {0, 0, 0, 0}};
  int o;
  while ((o = getopt_long(argc, argv, "hvfp", lo, NULL)) != -1) {
    switch (o) {
    case 'h': usage(argv[0]); return 0;
    case 'v': params.verbose = true; break;
    case 1001: params.termination_criteria.time_sec_limit = atof(optarg); break;
    case 1002: params.termination_criteria.iteration_limit = atoi(optarg); break;
    case 1003: params.termination_criteria.eps_optimal_relative = atof(optarg); break;
    case 1004: params.termination_criteria.eps_feasible_relative = atof(optarg); break;
    case 1005: params.termination_criteria.eps_infeasible = atof(optarg); break;
    case 1006: params.termination_criteria.eps_feas_polish_relative = atof(optarg); break;
    case 'f': params.feasibility_polishing = true; break;
    case 1007: params.l_inf_ruiz_iterations = atoi(optarg); break;
    case 1008: params.pock_chambolle_alpha = atof(optarg); break;
    case 1009: params.has_pock_chambolle_alpha = false; break;
    case 1010: params.bound_objective_rescaling = false; break;
    case 1011: params.sv_max_iter = atoi(optarg); break;
    case 1012: params.sv_tol = atof(optarg); break;
    case 1013: params.termination_evaluation_frequency = atoi(optarg); break;
    case 1014:
      if (strcmp(optarg, "l2") == 0) {
        params.optimality_norm = NORM_TYPE_L2;
      } else if (strcmp(optarg, "linf") == 0) {
        params.optimality_norm = NORM_TYPE_L_INF;
      } else {
        fprintf(stderr, "Error: opt_norm must be 'l2' or 'linf'\n");
        return 1;
      }
      break;
    case 1015: params.presolve = false; break;
    case 1100: opts.precision = DPDLP_PREC_MIXED; break;
    /* the reference's process-grid options (cli.c:470-547); here one GPU
     * holds the whole matrix, so the grid must be 1,1 and the partition and
     * permutation settings are recorded in the parameters without effect */
    case 2001: {
      int r = 0, c = 0;
      if (sscanf(optarg, "%d,%d", &r, &c) != 2) {
        fprintf(stderr, "Error: Invalid grid_size format. Use --grid_size r,c (e.g., --grid_size 1,1)\n");
        return 1;
      }
      if (r <= 0 || c <= 0) {
        fprintf(stderr, "Error: Grid dimensions must be positive integers. Got: %s\n", optarg);
        return 1;
      }
      if (r * c != 1) {
        fprintf(stderr, "Error: --grid_size %s asks for %d GPUs; this solver runs on one (1,1)\n",
                optarg, r * c);
        return 1;
      }
      params.grid_size.row_dims = r;
      params.grid_size.col_dims = c;
      params.grid_size.decided = true;
      break;
    }
    case 2002:
      if (strcmp(optarg, "uniform") == 0) {
        params.partition_method = UNIFORM_PARTITION;
      } else if (strcmp(optarg, "nnz") == 0 || strcmp(optarg, "nnz_balance") == 0) {
        params.partition_method = NNZ_BALANCE_PARTITION;
      } else {
        fprintf(stderr, "Error: partition_method must be 'uniform' or 'nnz'\n");
        return 1;
      }
      break;
    case 2003:
      if (strcmp(optarg, "none") == 0 || strcmp(optarg, "no") == 0) {
        params.permute_method = NO_PERMUTATION;
      } else if (strcmp(optarg, "random") == 0 || strcmp(optarg, "full") == 0) {
        params.permute_method = FULL_RANDOM_PERMUTATION;
      } else if (strcmp(optarg, "block") == 0) {
        params.permute_method = BLOCK_RANDOM_PERMUTATION;
      } else {
        fprintf(stderr, "Error: permute_method must be 'none', 'random', or 'block'\n");
        return 1;
      }
      break;
    case 2004:
      params.permute_block_size = atoi(optarg);
      if (params.permute_block_size <= 0) {
        fprintf(stderr, "Error: permute_block_size must be a positive integer.\n");
        return 1;
      }
      break;
    case 'p': break; /* accepted and ignored, as in the reference */
    case '?': return 1;
    }
  }
  if (argc - optind != 2) {
    fprintf(stderr, "Error: You must specify an input file and an output directory.\n\n");
    usage(argv[0]);
    return 1;
  }
  const char *file = argv[optind], *dir = argv[optind + 1];
  char *name = instance_name(file);
  if (!name) return 1;
  lp_problem_t *prob = read_mps_file(file);
  if (!prob) {
    fprintf(stderr, "Failed to read or parse the file.\n");
    free(name);
    return 1;
  }
  cupdlpx_result_t *r = dpdlp_solve_ex(prob, &params, &opts);
  if (!r) {
    fprintf(stderr, "Solver failed.\n");
  } else {
    save_summary(r, dir, name);
    save_vector(r->primal_solution, prob->num_variables, dir, name, "_primal_solution.txt");
    save_vector(r->dual_solution, prob->num_constraints, dir, name, "_dual_solution.txt");
    cupdlpx_result_free(r);
  }
  lp_problem_free(prob);
  free(name);
  return 0;
}
