/* vszipcu_vs4.h — VapourSynth API v4 plugin entry of libvszipcu_amd.so.
 *
 * The library is itself a VapourSynth plugin: placed in a plugin directory (or loaded with
 * core.std.LoadPlugin) it registers namespace `vszipcu` with BM3D, BM3Dv2 and VAggregate,
 * their argument strings verbatim from src/plugin.zig:84-98, over the C ABI of
 * vszipcu_bm3d.h. VSPlugin and VSPLUGINAPI are VapourSynth's own (VapourSynth4.h).
 */
#ifndef VSZIPCU_VS4_H
#define VSZIPCU_VS4_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct VSPlugin VSPlugin;
typedef struct VSPLUGINAPI VSPLUGINAPI;

/* Replaces the reference's `export fn VapourSynthPluginInit2` (src/plugin.zig:15-23, 80-98):
 * configPlugin("com.julek.vszipcu", "vszipcu", "VapourSynth Zig Image Process CUDA", 0.1) and
 * registerFunction for BM3D, BM3Dv2 and VAggregate. */
void VapourSynthPluginInit2(VSPlugin *plugin, const VSPLUGINAPI *vspapi);

#ifdef __cplusplus
}
#endif

#endif /* VSZIPCU_VS4_H */
