#ifndef VSHIP_VS_H
#define VSHIP_VS_H
/* VapourSynth plugin entry of libvship_amd.so.
 *
 * Replaces the reference's VapourSynthPluginInit2 (src/VshipLib.cpp:62-68):
 * a VapourSynth API v4 host that loads this library calls it with its plugin
 * handle and plugin API table. It registers, under id "com.lumen.vship" and
 * namespace "vship" (version 4.1):
 *   SSIMULACRA2(reference:vnode; distorted:vnode; numStream:int:opt;
 *               gpu_id:int:opt) -> clip:vnode
 *     (src/ssimu2/vapoursynth.hpp:8-195: inputs converted to RGBS by the
 *     host's resize.Bicubic when needed, frame property _SSIMULACRA2 on a
 *     copy of the distorted frame)
 *   GpuInfo(gpu_id:int:opt) -> gpu_human_data:data (src/VshipLib.cpp:9-60)
 * The pointer types are VapourSynth's VSPlugin and VSPLUGINAPI. */
#ifdef __cplusplus
extern "C" {
#endif

struct VSPlugin;
struct VSPLUGINAPI;

__attribute__((visibility("default"))) void VapourSynthPluginInit2(struct VSPlugin* plugin,
                                                                   const struct VSPLUGINAPI* vspapi);

#ifdef __cplusplus
}
#endif
#endif
