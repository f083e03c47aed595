/* The part of the VapourSynth plugin interface, API 3.6, that the bm3dcuda
 * plugin entry (bm3d_vsplugin.cpp) and its test host use. VapourSynth's own
 * header is not part of this image, so the declarations are written here
 * from the published API: the names, values, struct layouts and the order
 * of the VSAPI function table are the binary interface a VapourSynth R4x-R6x
 * core hands to a plugin, so they must stay exactly as the core has them.
 * Everything is x86-64 Linux: VS_CC is the default calling convention. */
#ifndef BM3D_VS_API3_H
#define BM3D_VS_API3_H

#include <stdint.h>

#define VAPOURSYNTH_API_MAJOR 3
#define VAPOURSYNTH_API_MINOR 6
#define VAPOURSYNTH_API_VERSION ((VAPOURSYNTH_API_MAJOR << 16) | (VAPOURSYNTH_API_MINOR))

#define VS_CC
#ifdef __cplusplus
#define VS_EXTERNAL_API(ret) extern "C" __attribute__((visibility("default"))) ret VS_CC
#else
#define VS_EXTERNAL_API(ret) __attribute__((visibility("default"))) ret VS_CC
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct VSFrameRef VSFrameRef;
typedef struct VSNodeRef VSNodeRef;
typedef struct VSCore VSCore;
typedef struct VSPlugin VSPlugin;
typedef struct VSNode VSNode;
typedef struct VSFuncRef VSFuncRef;
typedef struct VSMap VSMap;
typedef struct VSAPI VSAPI;
typedef struct VSFrameContext VSFrameContext;

typedef enum VSColorFamily {
    cmGray = 1000000,
    cmRGB = 2000000,
    cmYUV = 3000000,
    cmYCoCg = 4000000,
    cmCompat = 9000000
} VSColorFamily;

typedef enum VSSampleType { stInteger = 0, stFloat = 1 } VSSampleType;

typedef enum VSFilterMode {
    fmParallel = 100,
    fmParallelRequests = 200,
    fmUnordered = 300,
    fmSerial = 400
} VSFilterMode;

typedef struct VSFormat {
    char name[32];
    int id;
    int colorFamily;   /* VSColorFamily */
    int sampleType;    /* VSSampleType */
    int bitsPerSample;
    int bytesPerSample;
    int subSamplingW;  /* log2, planes 1 and 2 */
    int subSamplingH;
    int numPlanes;
} VSFormat;

typedef enum VSNodeFlags { nfNoCache = 1, nfIsCache = 2, nfMakeLinear = 4 } VSNodeFlags;

typedef enum VSGetPropErrors { peUnset = 1, peType = 2, peIndex = 4 } VSGetPropErrors;

typedef enum VSPropAppendMode { paReplace = 0, paAppend = 1, paTouch = 2 } VSPropAppendMode;

typedef struct VSCoreInfo {
    const char *versionString;
    int core;
    int api;
    int numThreads;
    int64_t maxFramebufferSize;
    int64_t usedFramebufferSize;
} VSCoreInfo;

typedef struct VSVideoInfo {
    const VSFormat *format;  /* NULL: variable format */
    int64_t fpsNum;
    int64_t fpsDen;
    int width;               /* 0: variable size */
    int height;
    int numFrames;
    int flags;
} VSVideoInfo;

typedef enum VSActivationReason {
    arInitial = 0,
    arFrameReady = 1,
    arAllFramesReady = 2,
    arError = -1
} VSActivationReason;

typedef enum VSMessageType { mtDebug = 0, mtWarning = 1, mtCritical = 2, mtFatal = 3 } VSMessageType;

typedef void (VS_CC *VSPublicFunction)(const VSMap *in, VSMap *out, void *userData, VSCore *core,
                                       const VSAPI *vsapi);
typedef void (VS_CC *VSRegisterFunction)(const char *name, const char *args, VSPublicFunction argsFunc,
                                         void *functionData, VSPlugin *plugin);
typedef void (VS_CC *VSConfigPlugin)(const char *identifier, const char *defaultNamespace, const char *name,
                                     int apiVersion, int readonly, VSPlugin *plugin);
typedef void (VS_CC *VSInitPlugin)(VSConfigPlugin configFunc, VSRegisterFunction registerFunc, VSPlugin *plugin);
typedef void (VS_CC *VSFreeFuncData)(void *userData);
typedef void (VS_CC *VSFilterInit)(VSMap *in, VSMap *out, void **instanceData, VSNode *node, VSCore *core,
                                   const VSAPI *vsapi);
typedef const VSFrameRef *(VS_CC *VSFilterGetFrame)(int n, int activationReason, void **instanceData,
                                                    void **frameData, VSFrameContext *frameCtx, VSCore *core,
                                                    const VSAPI *vsapi);
typedef void (VS_CC *VSFilterFree)(void *instanceData, VSCore *core, const VSAPI *vsapi);
typedef void (VS_CC *VSFrameDoneCallback)(void *userData, const VSFrameRef *f, int n, VSNodeRef *,
                                          const char *errorMsg);
typedef void (VS_CC *VSMessageHandler)(int msgType, const char *msg, void *userData);
typedef void (VS_CC *VSMessageHandlerFree)(void *userData);

/* The core's function table, in the core's order. */
struct VSAPI {
    VSCore *(VS_CC *createCore)(int threads);
    void (VS_CC *freeCore)(VSCore *core);
    const VSCoreInfo *(VS_CC *getCoreInfo)(VSCore *core);

    const VSFrameRef *(VS_CC *cloneFrameRef)(const VSFrameRef *f);
    VSNodeRef *(VS_CC *cloneNodeRef)(VSNodeRef *node);
    VSFuncRef *(VS_CC *cloneFuncRef)(VSFuncRef *f);

    void (VS_CC *freeFrame)(const VSFrameRef *f);
    void (VS_CC *freeNode)(VSNodeRef *node);
    void (VS_CC *freeFunc)(VSFuncRef *f);

    VSFrameRef *(VS_CC *newVideoFrame)(const VSFormat *format, int width, int height, const VSFrameRef *propSrc,
                                       VSCore *core);
    VSFrameRef *(VS_CC *copyFrame)(const VSFrameRef *f, VSCore *core);
    void (VS_CC *copyFrameProps)(const VSFrameRef *src, VSFrameRef *dst, VSCore *core);

    void (VS_CC *registerFunction)(const char *name, const char *args, VSPublicFunction argsFunc,
                                   void *functionData, VSPlugin *plugin);
    VSPlugin *(VS_CC *getPluginById)(const char *identifier, VSCore *core);
    VSPlugin *(VS_CC *getPluginByNs)(const char *ns, VSCore *core);
    VSMap *(VS_CC *getPlugins)(VSCore *core);
    VSMap *(VS_CC *getFunctions)(VSPlugin *plugin);
    void (VS_CC *createFilter)(const VSMap *in, VSMap *out, const char *name, VSFilterInit init,
                               VSFilterGetFrame getFrame, VSFilterFree free, int filterMode, int flags,
                               void *instanceData, VSCore *core);
    void (VS_CC *setError)(VSMap *map, const char *errorMessage);
    const char *(VS_CC *getError)(const VSMap *map);
    void (VS_CC *setFilterError)(const char *errorMessage, VSFrameContext *frameCtx);
    VSMap *(VS_CC *invoke)(VSPlugin *plugin, const char *name, const VSMap *args);

    const VSFormat *(VS_CC *getFormatPreset)(int id, VSCore *core);
    const VSFormat *(VS_CC *registerFormat)(int colorFamily, int sampleType, int bitsPerSample, int subSamplingW,
                                           int subSamplingH, VSCore *core);

    const VSFrameRef *(VS_CC *getFrame)(int n, VSNodeRef *node, char *errorMsg, int bufSize);
    void (VS_CC *getFrameAsync)(int n, VSNodeRef *node, VSFrameDoneCallback callback, void *userData);
    const VSFrameRef *(VS_CC *getFrameFilter)(int n, VSNodeRef *node, VSFrameContext *frameCtx);
    void (VS_CC *requestFrameFilter)(int n, VSNodeRef *node, VSFrameContext *frameCtx);
    void (VS_CC *queryCompletedFrame)(VSNodeRef **node, int *n, VSFrameContext *frameCtx);
    void (VS_CC *releaseFrameEarly)(VSNodeRef *node, int n, VSFrameContext *frameCtx);

    int (VS_CC *getStride)(const VSFrameRef *f, int plane);
    const uint8_t *(VS_CC *getReadPtr)(const VSFrameRef *f, int plane);
    uint8_t *(VS_CC *getWritePtr)(VSFrameRef *f, int plane);

    VSFuncRef *(VS_CC *createFunc)(VSPublicFunction func, void *userData, VSFreeFuncData free, VSCore *core,
                                   const VSAPI *vsapi);
    void (VS_CC *callFunc)(VSFuncRef *func, const VSMap *in, VSMap *out, VSCore *core, const VSAPI *vsapi);

    VSMap *(VS_CC *createMap)(void);
    void (VS_CC *freeMap)(VSMap *map);
    void (VS_CC *clearMap)(VSMap *map);

    const VSVideoInfo *(VS_CC *getVideoInfo)(VSNodeRef *node);
    void (VS_CC *setVideoInfo)(const VSVideoInfo *vi, int numOutputs, VSNode *node);
    const VSFormat *(VS_CC *getFrameFormat)(const VSFrameRef *f);
    int (VS_CC *getFrameWidth)(const VSFrameRef *f, int plane);
    int (VS_CC *getFrameHeight)(const VSFrameRef *f, int plane);
    const VSMap *(VS_CC *getFramePropsRO)(const VSFrameRef *f);
    VSMap *(VS_CC *getFramePropsRW)(VSFrameRef *f);

    int (VS_CC *propNumKeys)(const VSMap *map);
    const char *(VS_CC *propGetKey)(const VSMap *map, int index);
    int (VS_CC *propNumElements)(const VSMap *map, const char *key);
    char (VS_CC *propGetType)(const VSMap *map, const char *key);

    int64_t (VS_CC *propGetInt)(const VSMap *map, const char *key, int index, int *error);
    double (VS_CC *propGetFloat)(const VSMap *map, const char *key, int index, int *error);
    const char *(VS_CC *propGetData)(const VSMap *map, const char *key, int index, int *error);
    int (VS_CC *propGetDataSize)(const VSMap *map, const char *key, int index, int *error);
    VSNodeRef *(VS_CC *propGetNode)(const VSMap *map, const char *key, int index, int *error);
    const VSFrameRef *(VS_CC *propGetFrame)(const VSMap *map, const char *key, int index, int *error);
    VSFuncRef *(VS_CC *propGetFunc)(const VSMap *map, const char *key, int index, int *error);

    int (VS_CC *propDeleteKey)(VSMap *map, const char *key);
    int (VS_CC *propSetInt)(VSMap *map, const char *key, int64_t i, int append);
    int (VS_CC *propSetFloat)(VSMap *map, const char *key, double d, int append);
    int (VS_CC *propSetData)(VSMap *map, const char *key, const char *data, int size, int append);
    int (VS_CC *propSetNode)(VSMap *map, const char *key, VSNodeRef *node, int append);
    int (VS_CC *propSetFrame)(VSMap *map, const char *key, const VSFrameRef *f, int append);
    int (VS_CC *propSetFunc)(VSMap *map, const char *key, VSFuncRef *func, int append);

    int64_t (VS_CC *setMaxCacheSize)(int64_t bytes, VSCore *core);
    int (VS_CC *getOutputIndex)(VSFrameContext *frameCtx);
    VSFrameRef *(VS_CC *newVideoFrame2)(const VSFormat *format, int width, int height, const VSFrameRef **planeSrc,
                                        const int *planes, const VSFrameRef *propSrc, VSCore *core);
    void (VS_CC *setMessageHandler)(VSMessageHandler handler, void *userData);
    int (VS_CC *setThreadCount)(int threads, VSCore *core);

    const char *(VS_CC *getPluginPath)(const VSPlugin *plugin);

    /* API 3.1 */
    const int64_t *(VS_CC *propGetIntArray)(const VSMap *map, const char *key, int *error);
    const double *(VS_CC *propGetFloatArray)(const VSMap *map, const char *key, int *error);
    int (VS_CC *propSetIntArray)(VSMap *map, const char *key, const int64_t *i, int size);
    int (VS_CC *propSetFloatArray)(VSMap *map, const char *key, const double *d, int size);

    /* API 3.4 */
    void (VS_CC *logMessage)(int msgType, const char *msg);

    /* API 3.6 */
    int (VS_CC *addMessageHandler)(VSMessageHandler handler, VSMessageHandlerFree free, void *userData);
    int (VS_CC *removeMessageHandler)(int id);
    void (VS_CC *getCoreInfo2)(VSCore *core, VSCoreInfo *info);
};

#ifdef __cplusplus
}
#endif

#endif /* BM3D_VS_API3_H */
