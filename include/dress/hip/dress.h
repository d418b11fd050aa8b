/*
 * dress/hip/dress.h -- select the HIP backend by include, the way
 * libdress/include/dress/cuda/dress.h:20-33 selects CUDA:
 *
 *   #include "dress/hip/dress.h"
 *   dress_fit(g, 100, 1e-6, &iters, &delta);   // runs dress_fit_hip
 *
 * The guard name DRESS_HIP_REDIRECT_H plays the role DRESS_CUDA_REDIRECT_H
 * plays for include/dress/mpi/dress.h.  Do not include dress/dress.h's
 * redirect headers for two backends in one translation unit.
 */
#ifndef DRESS_HIP_REDIRECT_H
#define DRESS_HIP_REDIRECT_H

#include "dress/dress.h"
#include "dress/hip/dress_hip.h"

#define dress_fit               dress_fit_hip
#define dress_delta_fit         dress_delta_fit_hip
#define dress_delta_fit_strided dress_delta_fit_hip_strided
#define dress_nabla_fit         dress_nabla_fit_hip
#define dress_nabla_fit_strided dress_nabla_fit_hip_strided

#endif /* DRESS_HIP_REDIRECT_H */
