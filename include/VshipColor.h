/* VshipColor.h - input colourspace descriptor of the SSIMULACRA2 C ABI.
 *
 * ABI-identical to the reference's src/VshipColor.h:7-96: same type names,
 * same enumerator values (including the non-contiguous Vship_Sample_t values
 * and Vship_PRIMARIES_INTERNAL = -1) and the same struct member order, so a
 * caller compiled against the reference header links against this library
 * unchanged. Vship_Colorspace_t is passed by value (84 bytes, padded to 88).
 */
#ifndef VSHIPCOLOR_API_HEADER
#define VSHIPCOLOR_API_HEADER

#include <stdint.h>
#include <stdbool.h>

/* Sample storage of one plane element. Integer types above 8 bits are
 * little-endian uint16 masked to their bit depth (reference
 * src/gpuColorToLinear/anyDepthToFloat.hpp:53-54). */
typedef enum Vship_Sample_t {
    Vship_SampleFLOAT = 0,
    Vship_SampleHALF = 1,
    Vship_SampleUINT8 = 2,
    Vship_SampleUINT9 = 3,
    Vship_SampleUINT10 = 5,
    Vship_SampleUINT12 = 7,
    Vship_SampleUINT14 = 9,
    Vship_SampleUINT16 = 11,
} Vship_Sample_t;

typedef enum Vship_Range_t {
    Vship_RangeLimited = 0,
    Vship_RangeFull = 1,
} Vship_Range_t;

/* log2 of the horizontal / vertical chroma decimation. */
typedef struct Vship_ChromaSubsample_t {
    int subw;
    int subh;
} Vship_ChromaSubsample_t;

typedef enum Vship_ChromaLocation_t {
    Vship_ChromaLoc_Left = 0,
    Vship_ChromaLoc_Center = 1,
    Vship_ChromaLoc_TopLeft = 2,
    Vship_ChromaLoc_Top = 3,
} Vship_ChromaLocation_t;

typedef enum Vship_ColorFamily_t {
    Vship_ColorYUV,
    Vship_ColorRGB,
} Vship_ColorFamily_t;

typedef enum Vship_YUVMatrix_t {
    Vship_MATRIX_RGB = 0,
    Vship_MATRIX_BT709 = 1,
    Vship_MATRIX_BT470_BG = 5,
    Vship_MATRIX_ST170_M = 6,
    Vship_MATRIX_BT2020_NCL = 9,
    Vship_MATRIX_BT2020_CL = 10,
    Vship_MATRIX_BT2100_ICTCP = 14,
} Vship_YUVMatrix_t;

typedef enum Vship_TransferFunction_t {
    Vship_TRC_BT709 = 1,
    Vship_TRC_BT470_M = 4,
    Vship_TRC_BT470_BG = 5,
    Vship_TRC_BT601 = 6,
    Vship_TRC_ST240_M = 7,
    Vship_TRC_Linear = 8,
    Vship_TRC_sRGB = 13,
    Vship_TRC_PQ = 16,
    Vship_TRC_ST428 = 17,
    Vship_TRC_HLG = 18,
} Vship_TransferFunction_t;

typedef enum Vship_Primaries_t {
    Vship_PRIMARIES_INTERNAL = -1, /* CIE XYZ */
    Vship_PRIMARIES_BT709 = 1,
    Vship_PRIMARIES_BT470_M = 4,
    Vship_PRIMARIES_BT470_BG = 5,
    Vship_PRIMARIES_ST170_M = 6,
    Vship_PRIMARIES_ST240_M = 7,
    Vship_PRIMARIES_BT2020 = 9,
} Vship_Primaries_t;

typedef struct Vship_CropRectangle_t {
    int top;
    int bottom;
    int left;
    int right;
} Vship_CropRectangle_t;

/* Final image size is (target_width - left - right) x (target_height - top -
 * bottom); target_* = -1 means "no resize". */
typedef struct Vship_Colorspace_t {
    int64_t width;
    int64_t height;
    int64_t target_width;
    int64_t target_height;
    Vship_Sample_t sample;
    Vship_Range_t range;
    Vship_ChromaSubsample_t subsampling;
    Vship_ChromaLocation_t chromaLocation;
    Vship_ColorFamily_t colorFamily;
    Vship_YUVMatrix_t YUVMatrix;
    Vship_TransferFunction_t transferFunction;
    Vship_Primaries_t primaries;
    Vship_CropRectangle_t crop;
} Vship_Colorspace_t;

#endif
