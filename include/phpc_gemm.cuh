/*
 * libphpc — the reference's header name for the local GEMM entry points.
 *
 * The reference's SUMMA translation unit includes "phpc_gemm.cuh" (src/phpc_summa.c:5)
 * for phpc_gemm_cuda and phpc_gemm_cublas (src/phpc_gemm.cuh:4-14). This file keeps that
 * name so such code compiles unchanged against this include directory; the declarations
 * themselves live in phpc_gemm.h. It is plain C: no CUDA or HIP types.
 */
#ifndef PHPC_GEMM_CUH
#define PHPC_GEMM_CUH

#include "phpc_gemm.h"

#endif
