/* VshipAPI.h - the C ABI of libvship_amd, a drop-in for the SSIMULACRA2 path
 * of the reference's libvship.
 *
 * Every declaration below has the same name, signature, enumerator values and
 * struct layout as the reference's src/VshipAPI.h (line numbers cited per
 * entry), so callers built against the reference header (FFVship, Av1an-style
 * tools, the reference's own tests) link against libvship_amd.so unchanged.
 *
 * Scope of this build: the SSIMULACRA2 entry points are implemented on
 * MI355X (gfx950) HIP kernels. The Butteraugli and CVVDP entry points exist so
 * that such callers still link; they return Vship_BadErrorType and record a
 * detail message saying the metric is not part of this library.
 */
#ifndef VSHIP_API_HEADER
#define VSHIP_API_HEADER

#include <stdint.h>
#include <stdbool.h>

#include "VshipColor.h"

#define EXPORTPREPROCESS __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* reference src/VshipAPI.h:23-26 */
typedef enum {
    Vship_HIP = 0,
    Vship_Cuda = 1,
} Vship_Backend;

/* reference src/VshipAPI.h:28-33 */
typedef struct {
    int major;
    int minor;
    int minorMinor;
    Vship_Backend backend;
} Vship_Version;

/* replaces reference src/VshipAPI.h:35 (impl src/VshipLib.cpp:91-100);
 * returns {4, 1, 0, Vship_HIP}. */
EXPORTPREPROCESS Vship_Version Vship_GetVersion(void);

/* reference src/VshipAPI.h:38-65 */
typedef enum {
    Vship_NoError = 0,
    Vship_OutOfVRAM = 1,
    Vship_OutOfRAM = 2,
    Vship_HIPError = 12,
    Vship_BadDisplayModel = 3,
    Vship_DifferingInputType = 4,
    Vship_NonRGBSInput = 5,
    Vship_BadPath = 13,
    Vship_BadJson = 14,
    Vship_DeviceCountError = 6,
    Vship_NoDeviceDetected = 7,
    Vship_BadDeviceArgument = 8,
    Vship_BadDeviceCode = 9,
    Vship_BadHandler = 10,
    Vship_BadPointer = 11,
    Vship_BadErrorType = 15,
} Vship_Exception;

/* replaces reference src/VshipAPI.h:68 (impl src/VshipLib.cpp:102-112) */
EXPORTPREPROCESS Vship_Exception Vship_GetDeviceCount(int* number);

/* replaces reference src/VshipAPI.h:71 (impl src/VshipLib.cpp:169-181);
 * selects the device used by Vship_SSIMU2Init. */
EXPORTPREPROCESS Vship_Exception Vship_SetDevice(int gpu_id);

/* reference src/VshipAPI.h:74-80 */
typedef struct Vship_DeviceInfo {
    char name[256];
    uint64_t VRAMSize;
    int integrated;
    int MultiProcessorCount;
    int WarpSize;
} Vship_DeviceInfo;

/* replaces reference src/VshipAPI.h:82 (impl src/VshipLib.cpp:114-139) */
EXPORTPREPROCESS Vship_Exception Vship_GetDeviceInfo(Vship_DeviceInfo* device_info, int gpu_id);

/* replaces reference src/VshipAPI.h:85 (impl src/VshipLib.cpp:141-149,
 * src/util/gpuhelper.hpp:44-66): device count, id range, one probe kernel. */
EXPORTPREPROCESS Vship_Exception Vship_GPUFullCheck(int gpu_id);

/* replaces reference src/VshipAPI.h:89 (impl src/VshipLib.cpp:151-158).
 * Returns the size of the message including its NUL; len = 0 only queries the
 * size; a shorter buffer receives a truncated, NUL-terminated message. */
EXPORTPREPROCESS int Vship_GetErrorMessage(Vship_Exception exception, char* out_message, int len);

/* replaces reference src/VshipAPI.h:93 (impl src/VshipLib.cpp:160-167);
 * same size contract, about the last error of any call in the process. */
EXPORTPREPROCESS int Vship_GetDetailedLastError(char* out_message, int len);

/* reference src/VshipAPI.h:98-100 */
typedef struct Vship_SSIMU2Handler {
    unsigned int id;
} Vship_SSIMU2Handler;

/* replaces reference src/VshipAPI.h:102 (impl src/VshipLib.cpp:183-190):
 * page-locked host memory for the input planes. */
EXPORTPREPROCESS Vship_Exception Vship_PinnedMalloc(void** ptr, uint64_t size);

/* replaces reference src/VshipAPI.h:104 (impl src/VshipLib.cpp:192-199) */
EXPORTPREPROCESS Vship_Exception Vship_PinnedFree(void* ptr);

/* replaces reference src/VshipAPI.h:107 (impl src/VshipLib.cpp:220-222):
 * Init2 on the device chosen by Vship_SetDevice (default 0). */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2Init(Vship_SSIMU2Handler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace);

/* replaces reference src/VshipAPI.h:109 (impl src/VshipLib.cpp:201-218).
 * Both colourspaces must give the same final width and height, else
 * Vship_DifferingInputType. All device memory the handler needs is allocated
 * here, once; Vship_ComputeSSIMU2 allocates nothing. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2Init2(Vship_SSIMU2Handler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, int gpu_id);

/* replaces reference src/VshipAPI.h:112 (impl src/VshipLib.cpp:224-245).
 * A freed id returns Vship_BadHandler on every later use. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2Free(Vship_SSIMU2Handler handler);

/* replaces reference src/VshipAPI.h:115 (impl src/VshipLib.cpp:247-268,
 * src/ssimu2/main.hpp:119-155). srcp1/srcp2 are the three host planes of the
 * reference and distorted image (borrowed for the call, never written), with
 * their line sizes in bytes. Blocks until *score is written. One handler must
 * not be used by two threads at once; several handlers run in parallel. */
EXPORTPREPROCESS Vship_Exception Vship_ComputeSSIMU2(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3], const uint8_t* srcp2[3], const int64_t lineSize[3], const int64_t lineSize2[3]);

/* replaces reference src/VshipAPI.h:119 (impl src/VshipLib.cpp:270-286);
 * returns 0 for an invalid handler. */
EXPORTPREPROCESS int Vship_SSIMU2GetDetailedLastError(Vship_SSIMU2Handler handler, char* out_message, int len);

/* reference src/VshipAPI.h:121-147: Butteraugli entry points, declared so that
 * callers link. Not implemented by this library (see top of file). */
typedef struct Vship_ButteraugliHandler {
    unsigned int id;
} Vship_ButteraugliHandler;

typedef struct Vship_ButteraugliScore {
    double normQ;
    double norm3;
    double norminf;
} Vship_ButteraugliScore;

EXPORTPREPROCESS Vship_Exception Vship_ButteraugliInit(Vship_ButteraugliHandler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, int Qnorm, float intensity_multiplier);
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliInit2(Vship_ButteraugliHandler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, int Qnorm, float intensity_multiplier, int gpu_id);
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliFree(Vship_ButteraugliHandler handler);
EXPORTPREPROCESS Vship_Exception Vship_ComputeButteraugli(Vship_ButteraugliHandler handler, Vship_ButteraugliScore* score, const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3], const uint8_t* srcp2[3], const int64_t lineSize[3], const int64_t lineSize2[3]);
EXPORTPREPROCESS int Vship_ButteraugliGetDetailedLastError(Vship_ButteraugliHandler handler, char* out_message, int len);

/* reference src/VshipAPI.h:149-190: CVVDP entry points, declared so that
 * callers link. Not implemented by this library (see top of file). */
typedef struct Vship_CVVDPHandler {
    unsigned int id;
} Vship_CVVDPHandler;

EXPORTPREPROCESS Vship_Exception Vship_CVVDPInit(Vship_CVVDPHandler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, float fps, bool resizeToDisplay, const char* model_key_cstr);
EXPORTPREPROCESS Vship_Exception Vship_CVVDPInit2(Vship_CVVDPHandler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, float fps, bool resizeToDisplay, const char* model_key_cstr, const char* model_config_json_cstr);
EXPORTPREPROCESS Vship_Exception Vship_CVVDPInit3(Vship_CVVDPHandler* handler, Vship_Colorspace_t src_colorspace, Vship_Colorspace_t dis_colorspace, float fps, bool resizeToDisplay, const char* model_key_cstr, const char* model_config_json_cstr, int gpu_id);
EXPORTPREPROCESS Vship_Exception Vship_CVVDPFree(Vship_CVVDPHandler handler);
EXPORTPREPROCESS Vship_Exception Vship_CVVDPGetDistmapResolution(Vship_CVVDPHandler handler, int64_t* width, int64_t* height);
EXPORTPREPROCESS Vship_Exception Vship_ResetCVVDP(Vship_CVVDPHandler handler);
EXPORTPREPROCESS Vship_Exception Vship_ResetScoreCVVDP(Vship_CVVDPHandler handler);
EXPORTPREPROCESS Vship_Exception Vship_LoadTemporalCVVDP(Vship_CVVDPHandler handler, const uint8_t* srcp1[3], const uint8_t* srcp2[3], const int64_t lineSize[3], const int64_t lineSize2[3]);
EXPORTPREPROCESS Vship_Exception Vship_ComputeCVVDP(Vship_CVVDPHandler handler, double* score, const uint8_t* dstp, int64_t dststride, const uint8_t* srcp1[3], const uint8_t* srcp2[3], const int64_t lineSize[3], const int64_t lineSize2[3]);
EXPORTPREPROCESS int Vship_CVVDPGetDetailedLastError(Vship_CVVDPHandler handler, char* out_message, int len);

#ifdef __cplusplus
}
#endif
#endif
