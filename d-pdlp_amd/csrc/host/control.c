/*
 * Per-LP scalar logic of the reference, from raw reduction slots (see
 * control.h for the reference lines each function follows).
 */
#include "control.h"
#include "dpdlp_dev.h"

#include <math.h>

#define CTL_FN
#define CTL(x) ctl_##x
#include "control_impl.h"

double dpdlp_blue(const double *a) { return ctl_blue(a); }
