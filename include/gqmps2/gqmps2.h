// Umbrella include of the gqmps2-compatible API (GraceQ/MPS2 0.2-alpha.0,
// include/gqmps2/gqmps2.h:25-34), backed by the mps2_amd device engine.
#ifndef MPS2_AMD_GQMPS2_GQMPS2_H
#define MPS2_AMD_GQMPS2_GQMPS2_H

#define GQMPS2_VERSION_MAJOR 0
#define GQMPS2_VERSION_MINOR 2
#define GQMPS2_VERSION_PATCH 0

#include "gqmps2/case_params_parser.h"
#include "gqmps2/framework.h"
#include "gqmps2/finite_mps.h"
#include "gqmps2/finite_mps_measu.h"
#include "gqmps2/mpogen.h"
#include "gqmps2/algorithm.h"

#endif
