/*
 * gpu/common.cuh for building the reference's GPU driver (mcb/mcb_cycle_deg_two_remove.cpp) with g++
 * against libmcb_hip.so.  Put include/reference_shim ahead of the reference tree on the include path:
 * the driver's `#include <gpu/common.cuh>` then finds this file instead of the CUDA one, and no
 * reference file needs an edit.
 *
 * It provides what the driver takes from the reference's gpu/common.cuh:
 *   the five extern "C" entry points of gpu/common.cuh:18-27 (declared in mcb_hip.h, same names and
 *   signatures) and the gpu_struct class (gpu/gpu_struct.cuh:15-84, here a shim over the mcb_gpu
 *   handle).  The CUDA-typed globals of gpu/common.cuh:13-16 (prop, device_id, dimGrid, dimBlock)
 *   are defined by the CUDA build's gpu/common.cu and never used by the driver, so they are left out.
 */
#ifndef __H_COMMON_CUDA
#define __H_COMMON_CUDA

#include "gpu_struct.cuh"

#endif
