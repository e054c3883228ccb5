/*
 * MPS reader (free format, optionally gzip-compressed when the name ends in
 * ".gz").  Replaces include/mps_parser.h:21 of the reference
 * (src/mps_parser.c:365-507).  Returns NULL when the file cannot be opened or
 * parsed (the reference dereferences NULL for a missing file).
 */
#pragma once

#include "cupdlpx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

lp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
