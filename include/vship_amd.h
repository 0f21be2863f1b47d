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
 * besides the blur's 32 bytes per work item of partial sums, from which the
 * calling thread forms the score in FP64. Blocks until *score is written. */
EXPORTPREPROCESS Vship_Exception VshipAMD_ComputeSSIMU2Device(Vship_SSIMU2Handler handler, double* score,
                                                              const uint8_t* dsrcp1[3], const uint8_t* dsrcp2[3],
                                                              const int64_t lineSize[3], const int64_t lineSize2[3]);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the score of one pair of device-resident planes with no
 * handler and no allocation. The same kernels and host completion as
 * VshipAMD_ComputeSSIMU2Device (same score, bit for bit), but the tables, both
 * images' FP32 XYB planes and the partial sums live in a fixed 32 MiB device
 * array that each GPU the library runs on holds until the process ends. Pairs
 * that need more (above about 2.3 Mpixel, e.g. 2560x1440; 1920x1080 and
 * 2048x1080 fit) or more than 1024 blur items return Vship_OutOfVRAM, crop or
 * resize Vship_NonRGBSInput. Runs on gpu_id's null stream and blocks until
 * *score is written; uses about 40 KB of the calling thread's stack. Calls that
 * overlap on one GPU share its array and race: serialise them. */
EXPORTPREPROCESS Vship_Exception VshipAMD_ComputeSSIMU2DeviceNoAlloc(Vship_Colorspace_t src_colorspace,
                                                                     Vship_Colorspace_t dis_colorspace, double* score,
                                                                     const uint8_t* dsrcp1[3],
                                                                     const uint8_t* dsrcp2[3],
                                                                     const int64_t lineSize[3],
                                                                     const int64_t lineSize2[3], int gpu_id);

/* The 108 features (index plane*36 + scale*6 + norm*3 + map) of the last
 * compute on this handler; features of maps the score does not use are 0. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Features(Vship_SSIMU2Handler handler, double features[108]);

/* Copy of one stored positive-XYB plane (image 0 = reference, 1 = distorted;
 * plane 0..2 = X, Y, B; scale 0..4) of the last compute, as a dense
 * width*height float array. out = NULL only returns the size. Planes the
 * score does not use are not stored and return Vship_BadPointer. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2DebugPlane(Vship_SSIMU2Handler handler, int image, int plane,
                                                           int scale, float* out, int64_t* width, int64_t* height);

/* Replays the last compute `iters` times on the handler's stream and returns
 * the mean device time of the front-end and blur/map kernels (HIP events) and
 * the mean host time of the score's FP64 completion from the blur's partial
 * sums (ms_score). */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Profile(Vship_SSIMU2Handler handler, int iters, double* ms_frontend,
                                                        double* ms_blur, double* ms_score);

/* Algorithmic HBM bytes and blur-stage FLOPs of one pair on this handler:
 * bytes = 2 * input bytes per image + 16 * A + 48 * nblocks (A = pixels of the
 * computed maps, nblocks = 16x16 blocks over them). */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2Work(Vship_SSIMU2Handler handler, double* bytes, double* flops);

/* ---- batched path: many pairs of uneven sizes per call ---------------- */

typedef struct VshipAMD_SSIMU2Batch {
    unsigned int id;
} VshipAMD_SSIMU2Batch;

typedef enum {
    VshipAMD_StorageFP32 = 0,
    VshipAMD_StorageFP16 = 1,
} VshipAMD_Storage;

/* Tuning of the batched kernels: blur tile shape (FP32 VALU kernel 0: 64x32,
 * 1: 64x16, 2: 32x64, 3: 32x32, 4: 64x64; matrix-core kernels 5: 64x48
 * tiles, 6: 64x64, 7: 64x128, 8: 64x256 streamed strips) and the grids of
 * the blur and front-end kernels in workgroups per CU (1..64; workgroups loop
 * over the items beyond the grid). generic_front: 0 (default) runs each pair
 * whose two images have 8- or 16-bit YUV 4:2:0 BT.709 layout through the
 * front-end compiled for that layout, 1 runs every pair through the generic
 * front-end (the same arithmetic; for comparisons). */
typedef struct VshipAMD_BatchTuning {
    int blur_config;
    int blur_waves;
    int front_waves;
    int generic_front;
} VshipAMD_BatchTuning;

/* A batch context for up to max_pairs pairs per call on gpu_id. XYB planes
 * are stored in FP32 or FP16 (arithmetic is FP32, norms and the final score
 * FP64 on the GPU). Device arenas grow on the first calls and are then
 * reused; the three kernels and the score read-back are replayed as one HIP
 * graph. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchInit(VshipAMD_SSIMU2Batch* batch, int max_pairs, int storage,
                                                          int gpu_id);
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchFree(VshipAMD_SSIMU2Batch batch);

/* Scores of npairs pairs whose planes are resident in the memory of the
 * batch's GPU. Pair i uses src_cs[i] / dis_cs[i] (any sizes, the two of a
 * pair equal), planes dsrcp1[3i..3i+2] / dsrcp2[3i..3i+2] and line sizes
 * lineSize1[3i..] / lineSize2[3i..]. Blocks until scores[0..npairs) are
 * written. Pairs whose reference (the src_cs fields, the three dsrcp1 pointers and
 * line sizes) repeats an earlier pair's reuse that pair's stored reference
 * pyramid (candidate screening); the results are the same bit for bit. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchCompute(VshipAMD_SSIMU2Batch batch, int npairs,
                                                             const Vship_Colorspace_t* src_cs,
                                                             const Vship_Colorspace_t* dis_cs,
                                                             const uint8_t* const* dsrcp1, const uint8_t* const* dsrcp2,
                                                             const int64_t* lineSize1, const int64_t* lineSize2,
                                                             double* scores);

/* Two-step scoring for candidate screening. ComputeScales is BatchCompute
 * with only the map jobs of the scales in scale_mask blurred (bit s = scale s
 * of the pyramid; 0x1f = scales 0-4, all the score uses). Every plane is
 * still stored; the features of the other scales are 0 and scores[] are
 * those of the selected scales only. CompleteScales then blurs the scales in
 * its scale_mask for `count` pairs of that call (indices into it, each at
 * most once) on the planes it stored and writes their full scores (and
 * features, see BatchFeatures); the two masks must together cover 0x1f and
 * no other compute may run on the batch in between. The completed scores and
 * features equal BatchCompute's bit for bit. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchComputeScales(VshipAMD_SSIMU2Batch batch, int npairs,
                                                                   const Vship_Colorspace_t* src_cs,
                                                                   const Vship_Colorspace_t* dis_cs,
                                                                   const uint8_t* const* dsrcp1,
                                                                   const uint8_t* const* dsrcp2,
                                                                   const int64_t* lineSize1, const int64_t* lineSize2,
                                                                   int scale_mask, double* scores);
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchCompleteScales(VshipAMD_SSIMU2Batch batch, int count,
                                                                    const int* pairs, int scale_mask, double* scores);

/* Times every tuning candidate on the given pairs (reps launches each) and
 * keeps the fastest; returns it and its mean ms per batch. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchAutotune(VshipAMD_SSIMU2Batch batch, int npairs,
                                                              const Vship_Colorspace_t* src_cs,
                                                              const Vship_Colorspace_t* dis_cs,
                                                              const uint8_t* const* dsrcp1,
                                                              const uint8_t* const* dsrcp2, const int64_t* lineSize1,
                                                              const int64_t* lineSize2, int reps,
                                                              VshipAMD_BatchTuning* chosen, double* best_ms);
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchSetTuning(VshipAMD_SSIMU2Batch batch, VshipAMD_BatchTuning tuning);
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchGetTuning(VshipAMD_SSIMU2Batch batch, VshipAMD_BatchTuning* tuning);

/* npairs * 108 features of the last batch compute. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchFeatures(VshipAMD_SSIMU2Batch batch, double* features);

/* Mean device ms per kernel and per whole-graph replay of the last batch. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchProfile(VshipAMD_SSIMU2Batch batch, int iters,
                                                             double* ms_frontend, double* ms_blur, double* ms_score,
                                                             double* ms_graph);

/* Algorithmic HBM bytes and blur-stage FLOPs of the last batch. */
EXPORTPREPROCESS Vship_Exception VshipAMD_SSIMU2BatchWork(VshipAMD_SSIMU2Batch batch, double* bytes, double* flops);

/* ---- Score surrogate (no counterpart in the reference) --------------------
 * A 55 -> 128 -> 128 -> 2 GELU MLP with BF16 weights and FP32 accumulation
 * that predicts the SSIMULACRA2 score and its standard deviation from the 54
 * features of scales 0-2 (index plane * 18 + scale * 6 + k, k = ssim L1,
 * artifact L1, detail L1, ssim L4, artifact L4, detail L4, i.e. the first
 * 18 of each plane's 36 entries in the 108-feature vector) followed by the
 * natural log of the pixel count. Parameters are VshipAMD_MLPParamCount()
 * floats in the order W1[128][55], b1[128], W2[128][128], b2[128],
 * W3[2][128], b3[2]. Sample arrays (X: n * 55 floats row-major, Y, mean,
 * sigma: n floats) are device pointers. */
typedef struct {
    unsigned id;
} VshipAMD_MLP;

EXPORTPREPROCESS int VshipAMD_MLPParamCount(void);
/* max_batch bounds the samples of one gradient/training step. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPInit(VshipAMD_MLP* mlp, int max_batch, int gpu_id);
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPFree(VshipAMD_MLP mlp);
/* Sets the FP32 master weights (and their BF16 copies); resets Adam. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPSetParams(VshipAMD_MLP mlp, const float* params, int count);
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPGetParams(VshipAMD_MLP mlp, float* params, int count);
/* Input columns are normalised as (x - mean) / std (std <= 0: column ignored),
 * the target as (y - ymean) / ystd. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPSetNorm(VshipAMD_MLP mlp, const float* xmean, const float* xstd,
                                                     float ymean, float ystd);
/* Predicted score and standard deviation of n samples. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPPredict(VshipAMD_MLP mlp, const float* dX, int n, float* dmean,
                                                     float* dsigma);
/* Gradient (host, ParamCount floats) and value of the mean Gaussian negative
 * log-likelihood over n <= max_batch samples, without an update. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPGradients(VshipAMD_MLP mlp, const float* dX, const float* dY, int n,
                                                       float* grad, double* loss);
/* Adam training over `epochs` seeded shuffles in steps of `batch` samples with
 * a cosine learning-rate decay from lr to lr / 20; loss_per_epoch (host,
 * epochs floats, may be NULL) receives each epoch's mean training loss. */
EXPORTPREPROCESS Vship_Exception VshipAMD_MLPTrain(VshipAMD_MLP mlp, const float* dX, const float* dY, int n,
                                                   int epochs, int batch, float lr, float weight_decay, uint64_t seed,
                                                   float* loss_per_epoch);

/* Device memory helpers for callers without another HIP binding. */
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceMalloc(void** ptr, uint64_t size, int gpu_id);
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceFree(void* ptr);
EXPORTPREPROCESS Vship_Exception VshipAMD_MemcpyHtoD(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception VshipAMD_MemcpyDtoH(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceSynchronize(void);

/* Copies `bytes` (a multiple of 16) with a plain vector-copy kernel `iters`
 * times on the null stream and returns the mean ms per copy: an HBM
 * bandwidth probe and the reference kernel for calibrating byte counters. */
EXPORTPREPROCESS Vship_Exception VshipAMD_DeviceCopy(void* dst, const void* src, uint64_t bytes, int iters, double* ms);

#ifdef __cplusplus
}
#endif
#endif
