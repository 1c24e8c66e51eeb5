/*
 * Umbrella C header (SPLA 1.6.1 include/spla/spla.h:31-40). As in SPLA, the
 * ssbtr header is not included here.
 */
#ifndef SPLA_SPLA_H
#define SPLA_SPLA_H

#include "spla/config.h"
#include "spla/context.h"
#include "spla/errors.h"
#include "spla/gemm.h"
#include "spla/matrix_distribution.h"
#include "spla/pgemm_sbs.h"
#include "spla/pgemm_ssb.h"
#include "spla/types.h"

#endif
