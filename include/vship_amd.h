/* vship_amd.h - entry points of libvship_amd.so beyond the reference's ABI.
 *
 * None of these has a counterpart in the reference's src/VshipAPI.h. They
 * serve callers whose frames already live in GPU memory (decoders on the same
 * device, batch pipelines, benchmarks) and the parity tests that inspect the
 * intermediate results of one call. All of them use the error codes and
 * detailed-message contract of VshipAPI.h.
 */
#ifndef VSHIP_AMD_HEADER
#define VSHIP_AMD_HEADER

#include <stdint.h>

#include "VshipAPI.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Vship_ComputeSSIMU2 on planes already resident in the memory of the
 * handler's GPU (device pointers, line sizes in bytes). No PCIe transfer
 * besides the 8-byte score. Blocks until *score is written. */
EXPORTPREPROCESS Vship_Exception VshipAMD_ComputeSSIMU2Device(Vship_SSIMU2Handler handler, double* score,
                                                              const uint8_t* dsrcp1[3], const uint8_t* dsrcp2[3],
                                                              const int64_t lineSize[3], const int64_t lineSize2[3]);

/* The 108 features (index plane*36 + scale*6 + norm*3 + map) of the last
 * compute on this handler; features of maps the score does not use are 0. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Features(Vship_SSIMU2Handler handler, double features[108]);

/* Copy of one stored positive-XYB plane (image 0 = reference, 1 = distorted;
 * plane 0..2 = X, Y, B; scale 0..4) of the last compute, as a dense
 * width*height float array. out = NULL only returns the size. Planes the
 * score does not use are not stored and return Vship_BadPointer. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2DebugPlane(Vship_SSIMU2Handler handler, int image, int plane,
                                                           int scale, float* out, int64_t* width, int64_t* height);

/* Replays the kernels of the last compute `iters` times on the handler's
 * stream and returns the mean device time of each kernel (HIP events). */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Profile(Vship_SSIMU2Handler handler, int iters, double* ms_frontend,
                                                        double* ms_blur, double* ms_score);

/* Algorithmic HBM bytes and blur-stage FLOPs of one pair on this handler:
 * bytes = 2 * input bytes per image + 16 * A + 48 * nblocks (A = pixels of the
 * computed maps, nblocks = 16x16 blocks over them). */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Work(Vship_SSIMU2Handler handler, double* bytes, double* flops);

/* Device memory helpers for callers without another HIP binding. */
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceMalloc(void** ptr, uint64_t size, int gpu_id);
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceFree(void* ptr);
EXPORTPREPROCESS Vship_Exception VshipAMD_MemcpyHtoD(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception VshipAMD_MemcpyDtoH(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceSynchronize(void);

#ifdef __cplusplus
}
#endif
#endif
