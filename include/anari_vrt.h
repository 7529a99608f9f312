/*
 * anari_vrt.h -- the ANARI 1.0 C API subset implemented by libanari_vrt.so,
 * the ANARI front end of the gfx950 kernels (include/vrt.h).
 *
 * The reference is an ANARI device library (VisionarayLibrary.cpp:56-73,
 * VisionarayDevice.cpp:32-283) reached through the ANARI SDK's libanari.
 * The SDK is not available to this build, so this library exports the C API
 * functions themselves for its one device, with the SDK's names, argument
 * meanings and semantics, and the reference's direct device constructor
 * anariNewVisionarayDevice (VisionarayLibrary.cpp:69-73):
 *  - objects are reference counted (public references by the application,
 *    internal ones by parameters that hold them);
 *  - parameters are staged by anariSetParameter and take effect at the
 *    object's next commit; commits are deferred and flushed by
 *    anariRenderFrame (and by anariFrameReady / anariMapFrame);
 *  - anariRenderFrame is asynchronous; anariFrameReady(ANARI_WAIT) waits;
 *  - errors are reported through the device's status callback, never by
 *    return codes (functions that create objects return NULL on failure);
 *  - mapped frame channels are host copies (the reference maps device
 *    memory on HIP, Frame.h:83-95);
 *  - a frame's "frameCompletionCallback" (with
 *    "frameCompletionCallbackUserData") is called once the frame has
 *    finished on the GPU, from a device-owned thread (the reference calls it
 *    right after enqueueing the kernel on HIP, Frame.cpp:383-384).
 *
 * Supported subtypes: spatial fields "structuredRegular" (float32 data),
 * "nanovdb" (a float grid buffer), "unstructured" (tetrahedra); volume
 * "transferFunction1D"; geometry "triangle"; material "matte"; lights
 * "directional", "quad", "hdri"; camera "perspective"; renderers "raycast",
 * "pathtrace"; instance "transform".  Unknown subtypes yield an object that
 * is reported invalid when used, as in the reference.
 *
 * Enum values follow the ANARI 1.0 specification's tables; binary
 * compatibility with a particular libanari build cannot be verified here.
 */
#ifndef ANARI_VRT_H
#define ANARI_VRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANARI_VRT_API __attribute__((visibility("default")))

typedef struct anari_vrt_object *ANARIObject;
typedef ANARIObject ANARIDevice;
typedef ANARIObject ANARIArray;
typedef ANARIObject ANARIArray1D;
typedef ANARIObject ANARIArray2D;
typedef ANARIObject ANARIArray3D;
typedef ANARIObject ANARICamera;
typedef ANARIObject ANARIFrame;
typedef ANARIObject ANARIGeometry;
typedef ANARIObject ANARIGroup;
typedef ANARIObject ANARIInstance;
typedef ANARIObject ANARILight;
typedef ANARIObject ANARIMaterial;
typedef ANARIObject ANARIRenderer;
typedef ANARIObject ANARISurface;
typedef ANARIObject ANARISpatialField;
typedef ANARIObject ANARIVolume;
typedef ANARIObject ANARIWorld;

typedef int ANARIDataType;
#define ANARI_UNKNOWN 0
#define ANARI_DATA_TYPE 100
#define ANARI_STRING 101
#define ANARI_VOID_POINTER 102
#define ANARI_BOOL 103
#define ANARI_FRAME_COMPLETION_CALLBACK 203
#define ANARI_DEVICE 501
#define ANARI_OBJECT 502
#define ANARI_ARRAY 503
#define ANARI_ARRAY1D 504
#define ANARI_ARRAY2D 505
#define ANARI_ARRAY3D 506
#define ANARI_CAMERA 507
#define ANARI_FRAME 508
#define ANARI_GEOMETRY 509
#define ANARI_GROUP 510
#define ANARI_INSTANCE 511
#define ANARI_LIGHT 512
#define ANARI_MATERIAL 513
#define ANARI_RENDERER 514
#define ANARI_SURFACE 515
#define ANARI_SAMPLER 516
#define ANARI_SPATIAL_FIELD 517
#define ANARI_VOLUME 518
#define ANARI_WORLD 519
#define ANARI_UINT8 1004
#define ANARI_INT32 1016
#define ANARI_UINT32 1020
#define ANARI_UINT32_VEC2 1021
#define ANARI_UINT32_VEC3 1022
#define ANARI_UINT32_VEC4 1023
#define ANARI_UFIXED8_VEC4 1039
#define ANARI_FLOAT32 1068
#define ANARI_FLOAT32_VEC2 1069
#define ANARI_FLOAT32_VEC3 1070
#define ANARI_FLOAT32_VEC4 1071
#define ANARI_UFIXED8_RGBA_SRGB 2003
#define ANARI_FLOAT32_BOX1 2008
#define ANARI_FLOAT32_BOX2 2009
#define ANARI_FLOAT32_MAT4 2015

typedef int ANARIStatusSeverity;
#define ANARI_SEVERITY_FATAL_ERROR 0
#define ANARI_SEVERITY_ERROR 1
#define ANARI_SEVERITY_WARNING 2
#define ANARI_SEVERITY_PERFORMANCE_WARNING 3
#define ANARI_SEVERITY_INFO 4
#define ANARI_SEVERITY_DEBUG 5

typedef int ANARIStatusCode;
#define ANARI_STATUS_NO_ERROR 0
#define ANARI_STATUS_UNKNOWN_ERROR 1
#define ANARI_STATUS_INVALID_ARGUMENT 2
#define ANARI_STATUS_INVALID_OPERATION 3
#define ANARI_STATUS_OUT_OF_MEMORY 4
#define ANARI_STATUS_UNSUPPORTED_DEVICE 5

typedef int ANARIWaitMask;
#define ANARI_NO_WAIT 0
#define ANARI_WAIT 1

typedef void (*ANARIMemoryDeleter)(const void *userData, const void *appMemory);
typedef void (*ANARIFrameCompletionCallback)(const void *userData, ANARIDevice device,
                                             ANARIFrame frame);
typedef void (*ANARIStatusCallback)(const void *userData, ANARIDevice device,
                                    ANARIObject source, ANARIDataType sourceType,
                                    ANARIStatusSeverity severity, ANARIStatusCode code,
                                    const char *message);

/* the device (VisionarayLibrary.cpp:69-73); NULL when no GPU is usable */
ANARI_VRT_API ANARIDevice anariNewVisionarayDevice(ANARIStatusCallback defaultCallback,
                                                   const void *userPtr);

/* arrays: appMemory NULL -> device-managed memory (anariMapArray to fill) */
ANARI_VRT_API ANARIArray1D anariNewArray1D(ANARIDevice, const void *appMemory,
                                           ANARIMemoryDeleter deleter, const void *userData,
                                           ANARIDataType elementType, uint64_t numItems1);
ANARI_VRT_API ANARIArray2D anariNewArray2D(ANARIDevice, const void *appMemory,
                                           ANARIMemoryDeleter deleter, const void *userData,
                                           ANARIDataType elementType, uint64_t numItems1,
                                           uint64_t numItems2);
ANARI_VRT_API ANARIArray3D anariNewArray3D(ANARIDevice, const void *appMemory,
                                           ANARIMemoryDeleter deleter, const void *userData,
                                           ANARIDataType elementType, uint64_t numItems1,
                                           uint64_t numItems2, uint64_t numItems3);
ANARI_VRT_API void *anariMapArray(ANARIDevice, ANARIArray);
ANARI_VRT_API void anariUnmapArray(ANARIDevice, ANARIArray);

ANARI_VRT_API ANARIGeometry anariNewGeometry(ANARIDevice, const char *type);
ANARI_VRT_API ANARIMaterial anariNewMaterial(ANARIDevice, const char *type);
ANARI_VRT_API ANARISurface anariNewSurface(ANARIDevice);
ANARI_VRT_API ANARISpatialField anariNewSpatialField(ANARIDevice, const char *type);
ANARI_VRT_API ANARIVolume anariNewVolume(ANARIDevice, const char *type);
ANARI_VRT_API ANARILight anariNewLight(ANARIDevice, const char *type);
ANARI_VRT_API ANARIGroup anariNewGroup(ANARIDevice);
ANARI_VRT_API ANARIInstance anariNewInstance(ANARIDevice, const char *type);
ANARI_VRT_API ANARIWorld anariNewWorld(ANARIDevice);
ANARI_VRT_API ANARICamera anariNewCamera(ANARIDevice, const char *type);
ANARI_VRT_API ANARIRenderer anariNewRenderer(ANARIDevice, const char *type);
ANARI_VRT_API ANARIFrame anariNewFrame(ANARIDevice);

ANARI_VRT_API void anariSetParameter(ANARIDevice, ANARIObject, const char *name,
                                     ANARIDataType dataType, const void *mem);
ANARI_VRT_API void anariUnsetParameter(ANARIDevice, ANARIObject, const char *name);
ANARI_VRT_API void anariUnsetAllParameters(ANARIDevice, ANARIObject);
ANARI_VRT_API void anariCommitParameters(ANARIDevice, ANARIObject);
ANARI_VRT_API void anariRelease(ANARIDevice, ANARIObject);
ANARI_VRT_API void anariRetain(ANARIDevice, ANARIObject);

/* frame properties "duration" (FLOAT32, seconds of the last frame),
   "numSamples" (INT32 / UINT32, frames accumulated); returns 1 if written */
ANARI_VRT_API int anariGetProperty(ANARIDevice, ANARIObject, const char *name,
                                   ANARIDataType type, void *mem, uint64_t size,
                                   ANARIWaitMask mask);

ANARI_VRT_API void anariRenderFrame(ANARIDevice, ANARIFrame);
ANARI_VRT_API int anariFrameReady(ANARIDevice, ANARIFrame, ANARIWaitMask);
ANARI_VRT_API void anariDiscardFrame(ANARIDevice, ANARIFrame);
/* "channel.color", "channel.depth", "channel.normal" / "channel.albedo"
   (FLOAT32_VEC3), "channel.primitiveId" / "channel.objectId" /
   "channel.instanceId" (UINT32): a host copy valid until unmapped.  As in the
   reference (Frame::fillGBuffer) the normal, albedo and id channels are
   written only when the frame also has a depth channel. */
ANARI_VRT_API const void *anariMapFrame(ANARIDevice, ANARIFrame, const char *channel,
                                        uint32_t *width, uint32_t *height,
                                        ANARIDataType *pixelType);
ANARI_VRT_API void anariUnmapFrame(ANARIDevice, ANARIFrame, const char *channel);

#ifdef __cplusplus
}
#endif

#endif /* ANARI_VRT_H */
