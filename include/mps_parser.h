/*
 * The declarations in this header keep the interface of the cuPDLPx / D-PDLP
 * reference header of the same name, which carries this notice:
 *
 *   Copyright 2025 Haihao Lu
 *
 *   Licensed under the Apache License, Version 2.0 (the "License");
 *   you may not use this file except in compliance with the License.
 *   You may obtain a copy of the License at
 *
 *       http://www.apache.org/licenses/LICENSE-2.0
 *
 *   Unless required by applicable law or agreed to in writing, software
 *   distributed under the License is distributed on an "AS IS" BASIS,
 *   WITHOUT WARRANTIES OR CONDITIONS OF ANY KIND, either express or implied.
 *   See the License for the specific language governing permissions and
 *   limitations under the License.
 */
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
