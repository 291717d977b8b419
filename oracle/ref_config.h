/*
 * Forced-include configuration for building the reference GCGE library
 * as the oracle's comparison build (oracle/_ref).  It pre-defines the
 * include guard of src/ops_config.h:13, so that file's defaults
 * (MPI on, OpenMP off) are replaced without editing the reference.
 *
 * One MPI rank is arithmetic-neutral, so MPI is switched off: the only
 * path difference is that the Rayleigh-Ritz step calls dsyevx with
 * RANGE='A' instead of RANGE='I', IL=1, IU=N (src/ops_eig_sol_gcg.c:
 * 1090-1219); both return all eigenpairs of the projected matrix.
 *
 * REF_OMP selects the OpenMP column-splitting variant; its thread count is
 * the OpenMP runtime's (OMP_NUM_THREADS in the environment).
 */
#ifndef _OPS_CONFIG_H_
#define _OPS_CONFIG_H_

#define OPS_USE_INTEL_MKL 0
#define OPS_USE_MPI 0
#ifdef REF_OMP
#define OPS_USE_OMP 1
#define OMP_NUM_THREADS omp_get_max_threads()
#else
#define OPS_USE_OMP 0
#endif
#define PRINT_RANK 0
#define FORTRAN_WRAPPER(x) x##_

#endif
