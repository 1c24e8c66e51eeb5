/*
 * spla-mi355x public configuration header.
 *
 * Replaces the CMake-generated include/spla/config.h of SPLA 1.6.1
 * (include/spla/config.h.in:36-42). This build is a single-backend library:
 * the GPU backend is always HIP on gfx950, so SPLA_ROCM is fixed and
 * SPLA_CUDA is never defined.
 */
#ifndef SPLA_CONFIG_H
#define SPLA_CONFIG_H

#define SPLA_ROCM
#define SPLA_CBLAS

#if defined(__GNUC__) || defined(__clang__)
#define SPLA_EXPORT __attribute__((visibility("default")))
#else
#define SPLA_EXPORT
#endif

#endif
