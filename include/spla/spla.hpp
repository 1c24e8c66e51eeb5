/*
 * Umbrella C++ header (SPLA 1.6.1 include/spla/spla.hpp:35-42). As in SPLA,
 * the ssbtr header is not included here.
 */
#ifndef SPLA_SPLA_HPP
#define SPLA_SPLA_HPP

#include "spla/config.h"
#include "spla/context.hpp"
#include "spla/exceptions.hpp"
#include "spla/gemm.hpp"
#include "spla/matrix_distribution.hpp"
#include "spla/pgemm_sbs.hpp"
#include "spla/pgemm_ssb.hpp"
#include "spla/types.h"

#endif
