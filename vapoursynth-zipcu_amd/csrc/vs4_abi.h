// The VapourSynth API v4 binary interface, as far as the BM3D plugin entry (vs_plugin.cpp) and
// the test host use it. VapourSynth's own header is not part of this tree, so the types are
// declared here: structs with their documented field order, enum values as published, and the
// VSAPI function table with every slot in its published position (API 4.0). Slots the plugin
// never calls are declared as untyped placeholders so the offsets of the used ones stay right.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSZ_VS_MAKE_VERSION(major, minor) (((major) << 16) | (minor))
#define VSZ_VS_API_VERSION VSZ_VS_MAKE_VERSION(4, 0)

typedef struct VSFrame VSFrame;
typedef struct VSNode VSNode;
typedef struct VSCore VSCore;
typedef struct VSPlugin VSPlugin;
typedef struct VSFrameContext VSFrameContext;
typedef struct VSMap VSMap;
typedef struct VSAPI VSAPI;
typedef struct VSPLUGINAPI VSPLUGINAPI;

enum { cfUndefined = 0, cfGray = 1, cfRGB = 2, cfYUV = 3 };          /* VSColorFamily */
enum { stInteger = 0, stFloat = 1 };                                  /* VSSampleType */
enum { fmParallel = 0, fmParallelRequests = 1, fmUnordered = 2, fmFrameState = 3 };
enum { mtVideo = 1, mtAudio = 2 };                                    /* VSMediaType */
enum { ptUnset = 0, ptInt = 1, ptFloat = 2, ptData = 3, ptFunction = 4, ptVideoNode = 5,
       ptAudioNode = 6, ptVideoFrame = 7, ptAudioFrame = 8 };         /* VSPropertyType */
enum { peSuccess = 0, peUnset = 1, peType = 2, peError = 3, peIndex = 4 };
enum { maReplace = 0, maAppend = 1 };
enum { arInitial = 0, arAllFramesReady = 1, arError = -1 };           /* VSActivationReason */
enum { mtDebug = 0, mtInformation = 1, mtWarning = 2, mtCritical = 3, mtFatal = 4 };
enum { dtUnknown = -1, dtBinary = 0, dtUtf8 = 1 };                    /* VSDataTypeHint */
enum { rpGeneral = 0, rpNoFrameReuse = 1, rpStrictSpatial = 2 };      /* VSRequestPattern */

typedef struct VSVideoFormat {
    int colorFamily;
    int sampleType;
    int bitsPerSample;
    int bytesPerSample;
    int subSamplingW;
    int subSamplingH;
    int numPlanes;
} VSVideoFormat;

typedef struct VSVideoInfo {
    VSVideoFormat format;
    int64_t fpsNum;
    int64_t fpsDen;
    int width;
    int height;
    int numFrames;
} VSVideoInfo;

typedef struct VSFilterDependency {
    VSNode *source;
    int requestPattern;
} VSFilterDependency;

typedef const VSFrame *(*VSFilterGetFrame)(int n, int activationReason, void *instanceData, void **frameData,
                                           VSFrameContext *frameCtx, VSCore *core, const VSAPI *vsapi);
typedef void (*VSFilterFree)(void *instanceData, VSCore *core, const VSAPI *vsapi);
typedef void (*VSPublicFunction)(const VSMap *in, VSMap *out, void *userData, VSCore *core, const VSAPI *vsapi);
typedef void (*VSApiSlot)(void); /* a table entry this code never calls */

struct VSPLUGINAPI {
    int (*getAPIVersion)(void);
    int (*configPlugin)(const char *identifier, const char *pluginNamespace, const char *name, int pluginVersion,
                        int apiVersion, int flags, VSPlugin *plugin);
    int (*registerFunction)(const char *name, const char *args, const char *returnType, VSPublicFunction argsFunc,
                            void *functionData, VSPlugin *plugin);
};

struct VSAPI {
    /* filters and nodes */
    void (*createVideoFilter)(VSMap *out, const char *name, const VSVideoInfo *vi, VSFilterGetFrame getFrame,
                              VSFilterFree free, int filterMode, const VSFilterDependency *dependencies, int numDeps,
                              void *instanceData, VSCore *core);
    VSApiSlot createVideoFilter2, createAudioFilter, createAudioFilter2, setLinearFilter, setCacheMode,
        setCacheOptions;
    void (*freeNode)(VSNode *node);
    VSNode *(*addNodeRef)(VSNode *node);
    int (*getNodeType)(VSNode *node);
    const VSVideoInfo *(*getVideoInfo)(VSNode *node);
    VSApiSlot getAudioInfo;
    /* frames */
    VSFrame *(*newVideoFrame)(const VSVideoFormat *format, int width, int height, const VSFrame *propSrc,
                              VSCore *core);
    VSApiSlot newVideoFrame2, newAudioFrame, newAudioFrame2;
    void (*freeFrame)(const VSFrame *f);
    const VSFrame *(*addFrameRef)(const VSFrame *f);
    VSApiSlot copyFrame;
    const VSMap *(*getFramePropertiesRO)(const VSFrame *f);
    VSMap *(*getFramePropertiesRW)(VSFrame *f);
    ptrdiff_t (*getStride)(const VSFrame *f, int plane);
    const uint8_t *(*getReadPtr)(const VSFrame *f, int plane);
    uint8_t *(*getWritePtr)(VSFrame *f, int plane);
    const VSVideoFormat *(*getVideoFrameFormat)(const VSFrame *f);
    VSApiSlot getAudioFrameFormat, getFrameType;
    int (*getFrameWidth)(const VSFrame *f, int plane);
    int (*getFrameHeight)(const VSFrame *f, int plane);
    VSApiSlot getFrameLength;
    /* formats */
    VSApiSlot getVideoFormatName, getAudioFormatName, queryVideoFormat, queryAudioFormat, queryVideoFormatID,
        getVideoFormatByID;
    /* frame requests inside getFrame */
    VSApiSlot getFrame, getFrameAsync;
    const VSFrame *(*getFrameFilter)(int n, VSNode *node, VSFrameContext *frameCtx);
    void (*requestFrameFilter)(int n, VSNode *node, VSFrameContext *frameCtx);
    VSApiSlot releaseFrameEarly, cacheFrame;
    void (*setFilterError)(const char *errorMessage, VSFrameContext *frameCtx);
    /* functions */
    VSApiSlot createFunction, freeFunction, addFunctionRef, callFunction;
    /* maps */
    VSMap *(*createMap)(void);
    void (*freeMap)(VSMap *map);
    VSApiSlot clearMap, copyMap;
    void (*mapSetError)(VSMap *map, const char *errorMessage);
    const char *(*mapGetError)(const VSMap *map);
    VSApiSlot mapNumKeys, mapGetKey, mapDeleteKey;
    int (*mapNumElements)(const VSMap *map, const char *key);
    VSApiSlot mapGetType, mapSetEmpty;
    int64_t (*mapGetInt)(const VSMap *map, const char *key, int index, int *error);
    int (*mapGetIntSaturated)(const VSMap *map, const char *key, int index, int *error);
    VSApiSlot mapGetIntArray;
    int (*mapSetInt)(VSMap *map, const char *key, int64_t i, int append);
    int (*mapSetIntArray)(VSMap *map, const char *key, const int64_t *i, int size);
    double (*mapGetFloat)(const VSMap *map, const char *key, int index, int *error);
    float (*mapGetFloatSaturated)(const VSMap *map, const char *key, int index, int *error);
    VSApiSlot mapGetFloatArray, mapSetFloat, mapSetFloatArray;
    const char *(*mapGetData)(const VSMap *map, const char *key, int index, int *error);
    int (*mapGetDataSize)(const VSMap *map, const char *key, int index, int *error);
    VSApiSlot mapGetDataTypeHint, mapSetData;
    VSNode *(*mapGetNode)(const VSMap *map, const char *key, int index, int *error);
    int (*mapSetNode)(VSMap *map, const char *key, VSNode *node, int append);
    int (*mapConsumeNode)(VSMap *map, const char *key, VSNode *node, int append);
    VSApiSlot mapGetFrame, mapSetFrame, mapConsumeFrame, mapGetFunction, mapSetFunction, mapConsumeFunction;
    /* plugins */
    VSApiSlot registerFunction;
    VSPlugin *(*getPluginByID)(const char *identifier, VSCore *core);
    VSApiSlot getPluginByNamespace, getNextPlugin, getPluginName, getPluginID, getPluginNamespace,
        getNextPluginFunction, getPluginFunctionByName, getPluginFunctionName, getPluginFunctionArguments,
        getPluginFunctionReturnType, getPluginPath, getPluginVersion;
    VSMap *(*invoke)(VSPlugin *plugin, const char *name, const VSMap *args);
    /* core */
    VSApiSlot createCore, freeCore, setMaxCacheSize, setThreadCount, getCoreInfo;
    int (*getAPIVersion)(void);
    void (*logMessage)(int msgType, const char *msg, VSCore *core);
};

#ifdef __cplusplus
}
#endif
