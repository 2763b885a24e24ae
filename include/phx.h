/*
 * phx.h -- C ABI of the MI355X (gfx950) ray-query engine.
 *
 * This is the drop-in boundary for the CUDA_TRACER backend of PhantasyEngineTestbed. Every struct
 * below is byte-identical to the reference type it replaces, so the arrays the reference's host
 * code builds (buildStaticBVH + sanitizeBVH) can be handed over unchanged. Every entry point cites
 * the reference interface it replaces (paths are relative to the reference repository root).
 *
 * Conventions
 *  - All functions return phx_status; PHX_OK is 0. No exceptions cross the ABI and nothing calls
 *    exit/abort. The message of the last failure on the calling thread is phx_last_error().
 *  - Pointers flagged PHX_MEM_DEVICE are device pointers (hipMalloc'd or phx_device_alloc'd);
 *    otherwise they are host pointers and the call copies through a staging buffer.
 *  - A context owns one HIP stream. Calls on one context are issued in order on that stream.
 *    Calls without PHX_ASYNC return after the work has completed.
 *  - Threads: different contexts may be used from different threads at once. Calls on one
 *    context (and on the scenes, plans and renderers made from it) must not overlap in time: they
 *    share the context's staging buffers, ray counters and launch shape. Serialise them, or give
 *    each thread a context of its own. A server's dispatcher thread is internal to the server: it
 *    only enqueues on the server's slot streams (slot 0 is the context's stream) with the server's
 *    own counters, so calls on the context may run beside it.
 */
#ifndef PHX_H
#define PHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ */
/* Status codes                                                                               */
/* ------------------------------------------------------------------------------------------ */

typedef enum phx_status {
	PHX_OK = 0,
	PHX_ERR_INVALID_ARGUMENT = 1, /* null pointer, zero/oversized count, bad flag combination  */
	PHX_ERR_INVALID_HANDLE = 2,   /* released or foreign context/scene/bvh/server handle        */
	PHX_ERR_OUT_OF_MEMORY = 3,    /* host or device allocation failed                           */
	PHX_ERR_MALFORMED_BVH = 4,    /* child index out of range, missing end marker, too deep ... */
	PHX_ERR_HIP = 5,              /* a HIP runtime call failed (message holds hipGetErrorString)*/
	PHX_ERR_NO_DEVICE = 6,        /* no HIP device, or device index out of range                */
	PHX_ERR_BUDGET = 7,           /* served form: request can never fit the memory budget       */
	PHX_ERR_NOT_READY = 8,        /* served form: request still in flight (phx_poll)            */
	PHX_ERR_CANCELLED = 9,        /* served form: request was cancelled before it ran           */
	PHX_ERR_INVALID_SCENE = 10,   /* scene input unusable (no triangles, non-finite vertex ...) */
	PHX_ERR_DEVICE = 11           /* device arithmetic differs from what the kernels rely on     */
} phx_status;

/* Returns the message of the last failed call on this thread ("" if none). */
const char* phx_last_error(void);

/* Library version string, e.g. "phx 0.1.0 gfx950". */
const char* phx_version(void);

/* ------------------------------------------------------------------------------------------ */
/* Data layouts (byte-identical to the reference)                                             */
/* ------------------------------------------------------------------------------------------ */

/* BVHNode -- source/phantasy_engine/ray_tracer_common/BVHNode.hpp:23-149 (64 B).
 * fData = [lc.min.xyz, lc.max.x] [lc.max.yz, rc.min.xy] [rc.min.z, rc.max.xyz];
 * iData = [lcIndex, rcIndex, lcNumTriangles, rcNumTriangles]. A child with numTriangles != 0 is
 * a leaf whose index is the bitwise-negated (~) first triangle index; otherwise it is the index of
 * an inner node. Node 0 is the root. */
typedef struct PhxBVHNode {
	float fData[12];
	int32_t iData[4];
} PhxBVHNode;

/* TriangleVertices -- ray_tracer_common/Triangle.hpp:17-23 (48 B). v0..v2 are xyzw; w is 0 except
 * in the last triangle of a leaf, where all three w are -1 (end marker, StaticBVHBuilder.cpp:19-45). */
typedef struct PhxTriangleVertices {
	float v0[4];
	float v1[4];
	float v2[4];
} PhxTriangleVertices;

/* TriangleData -- ray_tracer_common/Triangle.hpp:25-37 (64 B). */
typedef struct PhxTriangleData {
	float n0[3];
	float n1[3];
	float n2[3];
	float uv0[2];
	float uv1[2];
	float uv2[2];
	uint32_t materialIndex;
} PhxTriangleData;

/* RayIn -- source/cuda_ray_tracer/kernels/RayCastKernel.hpp:22-37 (32 B):
 * (origin.xyz, minDist) (dir.xyz, maxDist). */
typedef struct PhxRayIn {
	float origin[3];
	float minDist;
	float dir[3];
	float maxDist;
} PhxRayIn;

/* RayHit -- kernels/RayCastKernel.hpp:39-45 (16 B), filled with the CPU traversal's
 * RayCastResult semantics (ray_tracer_common/BVHTraversal.hpp:18-27): on a miss
 * triangleIndex = 0xFFFFFFFF and t = u = v = FLT_MAX. */
typedef struct PhxRayHit {
	uint32_t triangleIndex;
	float t;
	float u;
	float v;
} PhxRayHit;

/* Material -- source/phantasy_engine/rendering/Material.hpp:17-55 (48 B):
 * iData = [albedoTex, roughnessTex, metallicTex, pad] (-1 = none), albedo rgba,
 * fData2 = [roughness, metallic, pad, pad]. */
typedef struct PhxMaterial {
	int32_t iData[4];
	float albedo[4];
	float fData2[4];
} PhxMaterial;

/* SphereLight -- source/phantasy_engine/level/SphereLight.hpp:13-26 (36 B with padding). */
typedef struct PhxSphereLight {
	float pos[3];
	float radius;
	float range;
	float strength[3];
	uint8_t staticShadows;
	uint8_t dynamicShadows;
	uint8_t pad[2];
} PhxSphereLight;

/* CameraDef -- ray_tracer_common/GenerateRays.hpp:16-25: ray = normalize(dir + cx*dX + cy*dY),
 * (cx, cy) = loc / res * 2 - 1 (pixel corner, no half-pixel offset). */
typedef struct PhxCameraDef {
	float origin[3];
	float dir[3];
	float dX[3];
	float dY[3];
} PhxCameraDef;

/* ------------------------------------------------------------------------------------------ */
/* Flags                                                                                      */
/* ------------------------------------------------------------------------------------------ */

#define PHX_MEM_DEVICE 0x1u /* ray/hit/image pointers are device pointers                    */
#define PHX_ASYNC 0x2u      /* return after enqueueing on the context stream (device ptrs only)*/

/* Traversal order of phx_cast_closest and phx_traversal_stats: always the CPU castRay visiting
 * order (left child first, BVHTraversal.hpp:37-128), so hits are bit-identical to it. The bits of
 * PHX_TRAVERSAL_MASK are reserved and must be zero. */
#define PHX_TRAVERSAL_REFERENCE_ORDER 0x00u
#define PHX_TRAVERSAL_MASK 0xF0u

/* Launch model of phx_cast_closest / phx_cast_anyhit / phx_traversal_stats. Default: persistent
 * wavefronts refilled from a shared ray counter (rayCastKernel's model, RayCastKernel.cu:179-411).
 * This flag selects one ray per lane (rayCastNoPersistenceKernel, RayCastKernel.cu:416-572).
 * Results are identical either way. */
#define PHX_KERNEL_ONE_RAY_PER_LANE 0x100u
/* Packet traversal for coherent rays (camera rays in 8x8-tile order, shadow rays towards one
 * light): a wave traces 64 consecutive rays over one shared stack, expanding the union of their
 * nodes; each ray's nodes, triangle tests and hit are still castRay's. No counterpart in the
 * reference. Not combinable with PHX_KERNEL_ONE_RAY_PER_LANE. */
#define PHX_KERNEL_PACKET 0x200u

/* Ray order of phx_gen_primary_rays. */
#define PHX_RAYS_ROW_MAJOR 0u  /* id = y * w + x (CPUTracerRenderer.cpp:84-97)                   */
#define PHX_RAYS_TILED_8X8 1u  /* 8x8-tile order of getRayIdx (CudaDeviceHelpers.cuh:49-62)       */

/* ------------------------------------------------------------------------------------------ */
/* Host-side SBVH build (replaces buildStaticFrom + sanitizeBVH,                              */
/* ray_tracer_common/StaticBVHBuilder.hpp:10-21, StaticBVHBuilder.cpp:213-260)                */
/* ------------------------------------------------------------------------------------------ */

typedef struct PhxBVH PhxBVH;

/* Builds the split BVH of `numTriangles` triangles exactly as buildStaticFrom + sanitizeBVH do
 * (Stich 2009 SBVH with the reference's parameters: splitAlpha 1e-5, unit SAH costs, min leaf 1,
 * 32 spatial bins, MaxDepth 64, MaxSpatialDepth 48). `datas` may be NULL. `numThreads` = 0 uses
 * all host cores; the output does not depend on it. Non-finite vertices -> PHX_ERR_INVALID_SCENE. */
phx_status phx_bvh_build(const PhxTriangleVertices* triangles, const PhxTriangleData* datas,
                         uint32_t numTriangles, uint32_t numThreads, PhxBVH** outBvh);

/* Read-only views of the built arrays (valid until phx_bvh_release). */
phx_status phx_bvh_arrays(const PhxBVH* bvh, const PhxBVHNode** nodes, uint32_t* numNodes,
                          const PhxTriangleVertices** triangles, const PhxTriangleData** datas,
                          uint32_t* numTriangles);

void phx_bvh_release(PhxBVH* bvh);

/* Checks a sanitized BVH: child indices in range, leaf ranges in range and terminated by exactly
 * one end marker, every inner node reachable once, depth <= 64. Writes the depth (root = 1). */
phx_status phx_bvh_validate(const PhxBVHNode* nodes, uint32_t numNodes,
                            const PhxTriangleVertices* triangles, uint32_t numTriangles,
                            uint32_t* outDepth);

/* ------------------------------------------------------------------------------------------ */
/* Context, device memory                                                                     */
/* ------------------------------------------------------------------------------------------ */

typedef struct PhxContext PhxContext;
typedef struct PhxScene PhxScene;

/* Replaces the CudaTracerRenderer constructor's device setup (CudaTracerRenderer.cpp:123-167):
 * binds `device`, reads its properties, creates the context stream. No device reset on destroy. */
phx_status phx_create(int device, PhxContext** outCtx);
/* The context handle is invalid once this returns. Scenes, plans, renderers, instance sets and
 * servers made from it stay valid until they are released: calls that take the context itself
 * (casts, plan launches, copies) fail with PHX_ERR_INVALID_HANDLE, a server keeps serving. The
 * context's stream and buffers are freed with the last of them, so any release order is safe.
 * Device buffers from phx_device_alloc are plain hipMalloc allocations: free them first, or with
 * hipFree afterwards. */
void phx_destroy(PhxContext* ctx);

/* The context's hipStream_t, as an opaque pointer (for events / interop). */
void* phx_context_stream(PhxContext* ctx);

/* Number of compute units of the context's device. */
uint32_t phx_context_num_cus(PhxContext* ctx);

/* Launch shape of the persistent traversal kernels on this context: grid size in blocks for
 * closest-hit and any-hit, block size (lane-mode threads, a multiple of 64 up to 1024), the number
 * of top tree nodes each block keeps in LDS (0..2048; the upload lays out the top of the tree
 * breadth first), the refill threshold (a wave refills its idle lanes once at least `refillMin` of
 * its 64 lanes are idle; 1..64), and helper waves: extra waves per block that trace no ray of
 * their own and take over, one wave per ray, the last rays the block's lane-mode waves hand out
 * (blockThreads + 64 * helperWaves <= 1024), rays per fetch (the rays a wave claims from the
 * launch's shared counter with one atomic; 8..4096) and the stack split (the traversal stack
 * entries per lane kept in LDS, 4..32; deeper entries live in HBM). phx_create's defaults: one
 * block of 512 threads per CU, no treelet, refill 8, no helper waves, 64 rays per fetch, 16 LDS
 * stack entries; phx_autotune picks them per scene and ray mix. In phx_set_launch_params a zero
 * field (0xFFFFFFFF for topNodes and helperWaves) keeps the current value; combinations that do
 * not fit a CU are rejected. In phx_plan_set_launch_params 0xFFFFFFFF helper waves means none and
 * a zero raysPerFetch or stackLdsEntries the default. Results do not depend on these parameters. */
typedef struct PhxLaunchParams {
	uint32_t gridBlocksClosest;
	uint32_t gridBlocksAnyhit;
	uint32_t refillMin;
	uint32_t blockThreads;
	uint32_t topNodes;
	uint32_t residentGridBlocks; /* read only: blocks of this shape the device holds at once */
	uint32_t helperWaves;
	uint32_t raysPerFetch;
	uint32_t stackLdsEntries;
} PhxLaunchParams;

phx_status phx_get_launch_params(PhxContext* ctx, PhxLaunchParams* out);
phx_status phx_set_launch_params(PhxContext* ctx, const PhxLaunchParams* params);

/* Autotuner for single-batch casts: times phx_cast_closest (or, with PHX_AUTOTUNE_ANYHIT,
 * phx_cast_anyhit) on the given device rays under candidate launch shapes -- block 256/512,
 * treelet 0/512/1024 nodes, 1/4, 1/2 or all of the resident grid, refill 8/32, 0 or 4 helper
 * waves, the one-ray-per-lane kernel and the packet kernel -- with HIP events; if a persistent
 * shape wins it then refines rays per fetch (16..256) and the stack split (8..32 LDS entries) one
 * at a time around it, and applies the fastest persistent shape to the context. `chosen` (may be NULL) receives it; residentGridBlocks
 * = 0xFFFFFFFF there means the one-ray-per-lane kernel won (pass PHX_KERNEL_ONE_RAY_PER_LANE to
 * the casts) and 0xFFFFFFFE the packet kernel (pass PHX_KERNEL_PACKET). `rays` must be device
 * memory (PHX_MEM_DEVICE). */
#define PHX_AUTOTUNE_ANYHIT 0x1000u
phx_status phx_autotune(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                        uint32_t numRays, uint32_t flags, PhxLaunchParams* chosen);

/* Device self-check; no counterpart in the reference. The traversal kernels compute 1/a as
 * v_rcp_f32 refined by one Newton-Raphson step, which on gfx950 equals IEEE division for every a
 * with 2^-126 <= |a| < 2^126; outside that range they divide. This compares the two on the device:
 * every significand at the exponents of 2^-126, 2^-1, 2^0 and 2^125, both signs (flags 0; phx_create
 * runs this and fails with PHX_ERR_DEVICE on a mismatch), or every float of the range
 * (PHX_CHECK_EXHAUSTIVE, about 4.2e9 inputs). counts[0] = inputs compared, counts[1] = mismatches. */
#define PHX_CHECK_EXHAUSTIVE 0x1u
phx_status phx_check_reciprocal(PhxContext* ctx, uint32_t flags, uint64_t counts[2]);

/* Device buffers (CudaBuffer.hpp:93-108 equivalents). Allocations are 256-B aligned. */
phx_status phx_device_alloc(PhxContext* ctx, size_t bytes, void** outPtr);
phx_status phx_device_free(PhxContext* ctx, void* ptr);
phx_status phx_copy_to_device(PhxContext* ctx, void* dst, const void* src, size_t bytes);
phx_status phx_copy_to_host(PhxContext* ctx, void* dst, const void* src, size_t bytes);
phx_status phx_copy_on_device(PhxContext* ctx, void* dst, const void* src, size_t bytes);
phx_status phx_synchronize(PhxContext* ctx);

/* ------------------------------------------------------------------------------------------ */
/* Static scene upload (replaces CudaTracerRenderer::setStaticScene, CudaTracerRenderer.cpp:251-   */
/* 270, and setMaterialsAndTextures' material upload, :192-239)                                */
/* ------------------------------------------------------------------------------------------ */

/* Copies the three SBVH arrays (+ optional materials and lights) into device memory owned by the
 * returned scene. The BVH is validated first (PHX_ERR_MALFORMED_BVH). `datas`, `materials` and
 * `lights` may be NULL with a zero count. The scene stays resident until phx_scene_release. */
phx_status phx_scene_upload(PhxContext* ctx, const PhxBVHNode* nodes, uint32_t numNodes,
                            const PhxTriangleVertices* triangles, const PhxTriangleData* datas,
                            uint32_t numTriangles, const PhxMaterial* materials,
                            uint32_t numMaterials, const PhxSphereLight* lights,
                            uint32_t numLights, PhxScene** outScene);
/* The handle is invalid after this call. A plan (phx_plan_create) or server (phx_server_add_scene)
 * that uses the scene keeps its device arrays alive until the plan is released or the server
 * destroyed, so their launches stay valid. */
void phx_scene_release(PhxScene* scene);

/* Device bytes held by a scene. */
size_t phx_scene_device_bytes(const PhxScene* scene);

/* RawImage (rendering/RawImage.hpp:21-29): `height` rows of `pitch` bytes, row 0 first, each
 * texel `bytesPerPixel` consecutive bytes. */
typedef struct PhxTexture {
	const uint8_t* data;
	uint32_t width;
	uint32_t height;
	uint32_t bytesPerPixel; /* 1, 2 or 4 (CudaBindlessTexture.cpp:42-53 rejects 3) */
	uint32_t pitch;         /* bytes per row, >= width * bytesPerPixel */
} PhxTexture;

/* Material textures. Replaces the texture half of setMaterialsAndTextures
 * (CudaTracerRenderer.cpp:192-239) and CudaBindlessTexture::load (CudaBindlessTexture.cpp:33-67).
 * Copies `count` images (rows packed to width * bytesPerPixel) into one device arena owned by the
 * scene, replacing the textures set before; count 0 removes them. A material's iData[0..2]
 * index this list; an index < 0 or >= count means no texture for that channel. Sampling (used
 * by phx_interpret_hits and phx_render_frame): nearest texel, wrapping, normalised coordinates,
 * texel x = floor(fmod(fmod(u * width, width) + width, width)) and y likewise from v with
 * height (CPURayTracerRenderer::sampleImage, CPUTracerRenderer.cpp:160-180; non-finite
 * coordinates read texel 0). Albedo reads rgba from 4-byte textures, rgb linearised as
 * pow(c / 255, 2.2) and rounded once to float, alpha c / 255 (InterpretRayHitKernel.cu:86-93);
 * textures of 1 or 2 bytes per texel leave the constant albedo (CPUTracerRenderer.cpp:200-206).
 * Roughness and metallic read byte 0 of a texel as c / 255 (:96-111). */
phx_status phx_scene_set_textures(PhxContext* ctx, PhxScene* scene, const PhxTexture* textures,
                                  uint32_t count);

/* ------------------------------------------------------------------------------------------ */
/* Scene files: the reference's .bvhcache (three raw array dumps, testbed_application/         */
/* Main.cpp:105-154) as one versioned file with a checksum, plus an importer of the raw dumps  */
/* ------------------------------------------------------------------------------------------ */

typedef struct PhxSceneArrays {
	const PhxBVHNode* nodes;
	uint32_t numNodes;
	uint32_t numTriangles;
	const PhxTriangleVertices* triangles;
	const PhxTriangleData* datas; /* may be NULL */
	const PhxMaterial* materials;
	uint32_t numMaterials;
	uint32_t numLights;
	const PhxSphereLight* lights;
} PhxSceneArrays;

typedef struct PhxSceneFile PhxSceneFile;

/* Writes the arrays (the BVH is validated first) as a version-1 scene file: a 64-byte header
 * (magic "PHXSCENE", version, counts, flags, FNV-1a 64 checksum of the payload) followed by the
 * nodes, triangles, datas, materials and lights, each section 64-byte aligned. */
phx_status phx_scene_file_write(const char* path, const PhxSceneArrays* arrays);
/* Reads a scene file; rejects a wrong magic or version, a size that differs from what the
 * header's counts give (truncation, trailing bytes; checked before any buffer is sized), nonzero
 * section padding, a checksum mismatch and a structurally invalid BVH with PHX_ERR_INVALID_SCENE /
 * PHX_ERR_MALFORMED_BVH. */
phx_status phx_scene_file_read(const char* path, PhxSceneFile** outFile);
/* Imports the reference's raw cache files (BVHNode, TriangleVertices and, if datasPath is not
 * NULL, TriangleData records) with the same validation (Main.cpp:120-134 checks sizes only). */
phx_status phx_scene_file_read_bvhcache(const char* nodesPath, const char* vertsPath,
                                        const char* datasPath, PhxSceneFile** outFile);
/* Views of a read file's arrays (valid until phx_scene_file_release), ready for phx_scene_upload. */
phx_status phx_scene_file_arrays(const PhxSceneFile* file, PhxSceneArrays* arrays);
void phx_scene_file_release(PhxSceneFile* file);

/* ------------------------------------------------------------------------------------------ */
/* Cast rays                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* Closest hit. Replaces launchRayCastKernel / launchRayCastNoPersistenceKernel
 * (kernels/RayCastKernel.hpp:57-61, RayCastKernel.cu:806-851). Hit/t/u/v follow the CPU castRay
 * (BVHTraversal.hpp:37-128) with tMin = ray.minDist, tMax = ray.maxDist: a triangle is accepted
 * iff t < closest.t && tMin <= t <= tMax; a child box is entered iff it is hit, tOut > tMin and
 * tIn < min(tMax, closest.t). `flags` = PHX_MEM_DEVICE | PHX_ASYNC | PHX_TRAVERSAL_*. */
phx_status phx_cast_closest(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                            uint32_t numRays, PhxRayHit* hits, uint32_t flags);

/* Diagnostics, no counterpart in the reference: phx_cast_closest on device buffers (`flags` =
 * PHX_MEM_DEVICE) with the persistent kernel at the context's launch shape, which also reports
 * when each ray finished: msSinceStart[i] (host array of numRays floats) = milliseconds from the
 * launch's first block start to the write of ray i's result, on the device's constant-rate clock.
 * Hits equal phx_cast_closest's. For tail analysis (when the last rays finish); the clock writes
 * make this launch slightly slower than the plain one. At most 2^31 - 1 rays. */
phx_status phx_cast_closest_timeline(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                                     uint32_t numRays, PhxRayHit* hits, float* msSinceStart,
                                     uint32_t flags);

/* phx_cast_closest on device buffers without allocating anything: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Results equal phx_cast_closest's. `flags` must contain
 * PHX_MEM_DEVICE and may add PHX_ASYNC; `rays` and `hits` are device pointers. The persistent
 * launch's ray counters and stack spill live in a static device array of the library (256 B +
 * 20 MiB per GPU, held until the process ends): the launch uses the context's shape with a
 * 32-entry LDS stack split where it fits and at most 5,242,880 / ((72 - split) * blockThreads)
 * blocks. At most 2^31 - 1 rays (PHX_ERR_INVALID_ARGUMENT beyond). Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other contexts or threads, or PHX_ASYNC calls on
 * different streams) race; serialise them. */
phx_status phx_cast_closest_noalloc(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                                    uint32_t numRays, PhxRayHit* hits, uint32_t flags);

/* Any hit (shadow rays). Replaces launchShadowRayCastKernel (RayCastKernel.hpp:63-64,
 * RayCastKernel.cu:577-801, 853-886): inLight[i] = 1 iff no triangle has tMin <= t < tMax, where
 * boxes are culled with the castRay box predicate (hit, tOut > tMin, tIn < tMax). One byte/ray. */
phx_status phx_cast_anyhit(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                           uint32_t numRays, uint8_t* inLight, uint32_t flags);

/* castDebugRay's counters (BVHTraversal.hpp:29-35, 130-235) of phx_cast_closest on the same rays,
 * summed over the batch, computed on the GPU by a counting variant of the traversal kernel:
 * counters = {nodesVisited, trianglesIntersected, triangleIntersectionTests,
 * aabbIntersectionTests, aabbIntersections}. Synchronous; hits are discarded. With
 * PHX_STATS_ANYHIT the counters are those of phx_cast_anyhit's traversal instead (castRay's
 * order with tCurrMax = tMax, ending at the first triangle with tMin <= t < tMax;
 * trianglesIntersected = rays occluded); no counterpart in the reference. */
#define PHX_STATS_ANYHIT 0x4000u
phx_status phx_traversal_stats(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                               uint32_t numRays, uint32_t flags, uint64_t counters[5]);

/* ------------------------------------------------------------------------------------------ */
/* Dynamic meshes: instances of uploaded meshes under a top tree (the reference's DynamicBVH /   */
/* OuterBVH, ray_tracer_common/DynamicBVH.hpp:13-26, DynamicBVH.cpp:16-329, for DynObject       */
/* {meshIndex, transform}, level/DynObject.hpp; its CUDA backend leaves them untraced,          */
/* CudaTracerRenderer.cpp:317-327, :412-415)                                                   */
/* ------------------------------------------------------------------------------------------ */

/* One placed mesh: `mesh` is an uploaded scene holding the mesh's object-space SBVH (and its
 * TriangleData for hit records); objectToWorld is the affine transform, rows (R | t):
 * world = R * p + t. */
typedef struct PhxInstance {
	const PhxScene* mesh;
	float objectToWorld[12];
} PhxInstance;

typedef struct PhxInstanceSet PhxInstanceSet;

/* Builds the two-level structure of `count` instances (0 allowed): world -> object transforms
 * (the inverse in double, rounded once to float), world boxes (the mesh root box's corners mapped
 * in double, widened by 1e-5 of the extent + 1e-6 of the magnitude, rounded outwards) and a top
 * tree with one instance per leaf. Rebuild it when instances move (the meshes' SBVHs stay). The
 * set holds its meshes until phx_instances_release. PHX_ERR_INVALID_ARGUMENT for a singular or
 * non-finite transform. */
phx_status phx_instances_build(PhxContext* ctx, const PhxInstance* instances, uint32_t count,
                               PhxInstanceSet** outSet);
void phx_instances_release(PhxInstanceSet* set);

/* Closest hit against the instances (and, if `scene` is not NULL, a static scene). An instance is
 * traced iff the ray enters its world box (castRay's box test with tIn <= min(tMax, best t)); it
 * runs castRay on its mesh with the ray in object space (o' = A o + b, d' = A d, the same tMin and
 * tMax, so t is the same parameter), from a miss. The closest result wins; on equal t the static
 * scene, then the lowest instance index. hits[i].triangleIndex is the triangle in the winning
 * mesh (or the static scene); instanceIds (may be NULL) receives the instance index, 0xFFFFFFFF
 * for the static scene or a miss. `flags` = PHX_MEM_DEVICE | PHX_ASYNC. */
phx_status phx_cast_closest_instances(PhxContext* ctx, const PhxScene* scene,
                                      const PhxInstanceSet* set, const PhxRayIn* rays,
                                      uint32_t numRays, PhxRayHit* hits, uint32_t* instanceIds,
                                      uint32_t flags);

/* Any hit against the instances (and `scene` if not NULL): inLight[i] = 0 iff some triangle of the
 * static scene or of an instance whose world box the ray enters has tMin <= t < tMax. */
phx_status phx_cast_anyhit_instances(PhxContext* ctx, const PhxScene* scene,
                                     const PhxInstanceSet* set, const PhxRayIn* rays,
                                     uint32_t numRays, uint8_t* inLight, uint32_t flags);

/* ------------------------------------------------------------------------------------------ */
/* Batched form: many ray batches in one persistent fused launch                               */
/* ------------------------------------------------------------------------------------------ */

#define PHX_BATCH_CLOSEST 0u /* out is PhxRayHit[numRays] (phx_cast_closest semantics)          */
#define PHX_BATCH_ANYHIT 1u  /* out is uint8_t[numRays] inLight (phx_cast_anyhit semantics)     */

/* One batch of a plan; rays and out are device pointers. */
typedef struct PhxBatch {
	const PhxScene* scene;
	const PhxRayIn* rays;
	void* out;
	uint32_t numRays;
	uint32_t kind;
} PhxBatch;

typedef struct PhxPlan PhxPlan;

#define PHX_PLAN_TRACK_COMPLETION 0x10u /* phx_plan_create: record each batch's completion time  */
#define PHX_PLAN_USE_GRAPH 0x20u        /* phx_plan_launch: replay the launch as a hipGraph      */

/* Replaces the per-batch launches of launchRayCastKernel / launchShadowRayCastKernel
 * (RayCastKernel.cu:806-886, one launch + device sync each, CudaTracerRenderer.cpp:610-721) with
 * one persistent launch over all batches (up to 1024 per plan, any mix of closest-hit and any-hit,
 * on one or more scenes of the context): the batches' rays are pulled from one shared counter, so
 * the tail of one batch overlaps the work of the others. Per ray the results equal
 * phx_cast_closest / phx_cast_anyhit. The plan keeps device copies of the batch table; the
 * scenes and buffers must outlive it. */
phx_status phx_plan_create(PhxContext* ctx, const PhxBatch* batches, uint32_t numBatches,
                           uint32_t flags, PhxPlan** outPlan);

/* Launches the plan on the context stream. PHX_PLAN_USE_GRAPH captures the launch into a hipGraph
 * the first time and replays it afterwards (the launch leaves its counters zeroed, so replays need
 * no reset). PHX_ASYNC returns after enqueueing. */
phx_status phx_plan_launch(PhxContext* ctx, PhxPlan* plan, uint32_t flags);

/* After a tracked launch completed: milliseconds from the launch's first block start to each
 * batch's completion (numBatches floats). */
phx_status phx_plan_batch_times(PhxContext* ctx, PhxPlan* plan, float* msSinceStart);

/* Launch shape of one plan (a fused launch is throughput bound, so plans default to the resident
 * grid without a treelet). gridBlocksClosest is the grid; gridBlocksAnyhit is ignored. */
phx_status phx_plan_set_launch_params(PhxContext* ctx, PhxPlan* plan,
                                      const PhxLaunchParams* params);

/* Times the plan under a set of candidate shapes (block 256/512, treelet 0/512/1024 nodes, 1/4,
 * 1/2 or all of the resident grid, refill 8/32) and keeps the fastest. Runs the plan's batches
 * several times (outputs are overwritten with the same values). */
phx_status phx_plan_autotune(PhxContext* ctx, PhxPlan* plan, PhxLaunchParams* chosen);

void phx_plan_release(PhxPlan* plan);

/* ------------------------------------------------------------------------------------------ */
/* Ray generation                                                                             */
/* ------------------------------------------------------------------------------------------ */

/* generateCameraDef (ray_tracer_common/GenerateRays.cpp:12-35) on the host. */
phx_status phx_camera_def(const float pos[3], const float dir[3], const float up[3],
                          float vertFovRad, float resX, float resY, PhxCameraDef* outCam);

/* ViewFrustum::set's orthogonalisation (rendering/ViewFrustum.cpp:126-131: dir = normalize(target -
 * pos), up = normalize(up - dot(up, dir) * dir)) followed by generateCameraDef. */
phx_status phx_camera_look_at(const float pos[3], const float target[3], const float up[3],
                              float vertFovRad, float resX, float resY, PhxCameraDef* outCam);

/* Pinhole primary rays (genPrimaryRaysKernel, RayCastKernel.cu:891-919; CPUTracerRenderer.cpp:
 * 93-97) with minDist = tMin, maxDist = tMax, in PHX_RAYS_ROW_MAJOR or PHX_RAYS_TILED_8X8 order
 * (the tiled order needs w and h divisible by 8). `rays` holds w*h records. */
phx_status phx_gen_primary_rays(PhxContext* ctx, const PhxCameraDef* cam, uint32_t width,
                                uint32_t height, float tMin, float tMax, uint32_t order,
                                PhxRayIn* rays, uint32_t flags);

/* One cosine-weighted diffuse bounce per ray (createSecondaryRaysKernel's Lambertian branch,
 * kernels/MaterialKernel.cu:239-312): origin = o + t*d, direction sampleLambertian around the
 * interpolated TriangleData normal turned to face the incoming ray, minDist 0.001, maxDist
 * FLT_MAX. Rays that missed (triangleIndex 0xFFFFFFFF), and hits whose triangle index is not a
 * triangle of `scene`, get the reference's dummy ray (MaterialKernel.cu:41-47). Random numbers: PCG32 stream `i` seeded from (seed, frame). The
 * scene must have TriangleData. */
phx_status phx_gen_bounce_rays(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                               const PhxRayHit* hits, uint32_t numRays, uint64_t seed,
                               uint32_t frame, PhxRayIn* outRays, uint32_t flags);

/* RayHitInfo -- kernels/InterpretRayHitKernel.hpp:21-58 (48 B): [pos, roughness],
 * [normal, metallic], [albedo rgb, alpha]; alpha < 0 marks a miss. */
typedef struct PhxRayHitInfo {
	float data1[4];
	float data2[4];
	float data3[4];
} PhxRayHitInfo;

/* Replaces launchInterpretRayHitKernel (InterpretRayHitKernel.hpp:76-78, .cu:58-136). A hit
 * gets position o + d*t, normal normalize((n0 + (n1-n0)u) + (n2-n0)v), uv interpolated the same
 * way, and its material: constant values, or texture samples as phx_scene_set_textures describes;
 * a material index outside the table reads albedo 1, alpha 1, roughness 1, metallic 0. A miss,
 * or a hit whose triangle index is not a triangle of `scene`, gets zeros and alpha -1. The scene must have TriangleData. `flags` = PHX_MEM_DEVICE |
 * PHX_ASYNC. */
phx_status phx_interpret_hits(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                              const PhxRayHit* hits, uint32_t numRays, PhxRayHitInfo* out,
                              uint32_t flags);

/* ------------------------------------------------------------------------------------------ */
/* Served form: a request loop over resident scenes with continuous batching. No counterpart in  */
/* the reference, whose renderer issues one synchronous launch per ray batch                  */
/* (RayCastKernel.cu:806-886); these calls replace that pattern for many concurrent clients.   */
/* ------------------------------------------------------------------------------------------ */

typedef struct PhxServer PhxServer;

/* One ray-query request. With PHX_MEM_DEVICE in flags, rays/out are device pointers used in place;
 * otherwise they are host pointers (pinned memory from phx_host_alloc copies fastest) and the
 * server copies them through its request arena. out: PhxRayHit per ray (closest) or one byte per
 * ray, 1 = unoccluded (any hit). */
typedef struct PhxRequest {
	const PhxScene* scene;
	const PhxRayIn* rays;
	void* out;
	uint32_t numRays;
	uint32_t kind; /* PHX_BATCH_CLOSEST or PHX_BATCH_ANYHIT */
	uint32_t flags;
	uint32_t pad;
} PhxRequest;

typedef struct PhxServerStats {
	uint64_t requestsSubmitted;
	uint64_t requestsDone;
	uint64_t raysDone;
	uint64_t launches;
	uint64_t rejected;      /* requests larger than the whole request budget */
	uint64_t bytesInUse;    /* request arena bytes held by admitted requests */
	uint64_t peakBytesInUse;
	uint64_t requestBudget; /* budget minus resident scene bytes */
	uint64_t sceneBytes;
	uint32_t queued;        /* submitted requests not yet launched */
	uint32_t cancelled;     /* requests cancelled before their launch */
} PhxServerStats;

#define PHX_SERVE_PLAIN 0x1u /* one request per launch, one launch in flight (per-batch baseline) */

/* A server on ctx's device with an HBM budget (resident scenes + request buffers) and `slots`
 * (1..4) launch streams. Requests are admitted in arrival order while their buffers fit, and
 * admitted requests are coalesced, in order, into the next fused persistent launch on an idle
 * slot, up to 4M rays per launch (a larger request gets a launch of its own), so a backlog spreads
 * over the slots and one launch's copies overlap another's traversal; a launch formed while
 * another runs fills that launch's tail. Slot 0 launches on the context's stream (other work on
 * that stream runs in order with it); each further slot has a stream of its own. HIP spreads a
 * process's streams over a few hardware queues (4 by default) and slots that share a queue
 * serialise, so more than 3 slots only pay where more queues exist. A dispatcher thread owns the
 * device work. The server
 * launches with the default stack split and rays per fetch and no treelet or helper waves; it
 * takes the block size and refill threshold from the context. */
phx_status phx_server_create(PhxContext* ctx, uint64_t budgetBytes, uint32_t slots,
                             uint32_t flags, PhxServer** outServer);
/* Waits for every submitted request, then frees the server. */
void phx_server_destroy(PhxServer* server);
/* Makes a scene of the server's context resident: its bytes count against the budget
 * (PHX_ERR_BUDGET if they do not fit); call before submitting requests. */
phx_status phx_server_add_scene(PhxServer* server, const PhxScene* scene);
/* Queues a request; never blocks. Fails with PHX_ERR_BUDGET if the request alone exceeds the
 * request budget, PHX_ERR_INVALID_HANDLE if its scene is not resident or was released, and
 * PHX_ERR_INVALID_ARGUMENT for more than 2^31 - 1 rays. */
phx_status phx_server_submit(PhxServer* server, const PhxRequest* request, uint64_t* ticket);
/* done = 1 once the request's outputs are in place; latency = submit to completion (ms). A
 * cancelled request reports done = 1 and returns PHX_ERR_CANCELLED. */
phx_status phx_server_poll(PhxServer* server, uint64_t ticket, int* done, double* latencyMs);
/* Blocks until the request completes, then forgets the ticket. Returns PHX_ERR_CANCELLED (and
 * forgets the ticket) for a cancelled request. */
phx_status phx_server_wait(PhxServer* server, uint64_t ticket, double* latencyMs);
/* Cancels a request that has not been launched yet: it leaves the queue, its arena space is
 * returned, its outputs are never written, and poll / wait report PHX_ERR_CANCELLED.
 * *cancelled (may be NULL) = 1 if the request is cancelled, 0 if it was already on the device or
 * complete (it then completes normally). No counterpart in the reference. */
phx_status phx_server_cancel(PhxServer* server, uint64_t ticket, int* cancelled);
/* pause != 0: the dispatcher launches nothing new until called again with 0 (requests are still
 * admitted while their buffers fit); launches already on the device complete. For draining, and
 * for holding requests before their launch. */
phx_status phx_server_pause(PhxServer* server, int pause);
phx_status phx_server_stats(PhxServer* server, PhxServerStats* stats);
/* Load driver: submits reqs[i] at arrivalMs[i] after the call starts (open loop), waits for all,
 * and returns each request's latency and the time from the call's start to the last completion. */
phx_status phx_server_replay(PhxServer* server, const PhxRequest* reqs, const double* arrivalMs,
                             uint32_t n, double* latencyMs, double* makespanMs);

/* Pinned host memory (hipHostMalloc), for request buffers that copy at full PCIe rate. */
phx_status phx_host_alloc(size_t bytes, void** out);
phx_status phx_host_free(void* p);

/* ------------------------------------------------------------------------------------------ */
/* Render frame: the wavefront path of CudaTracerRenderer::render (CudaTracerRenderer.cpp:637-   */
/* 720) and the TAA pass of GameScreen::render (GameScreen.cpp:187-251)                        */
/* ------------------------------------------------------------------------------------------ */

typedef struct PhxRenderer PhxRenderer;

#define PHX_FRAME_TAA 0x10000u       /* jitter by Halton(2,3)[frame % 16], resolve into history  */
#define PHX_FRAME_TAA_RESET 0x20000u /* start the TAA history at this frame                      */
#define PHX_FRAME_NO_CLAMP 0x40000u  /* TAA without the 3x3 neighbourhood clamp of the history   */
#define PHX_FRAME_GRAPH 0x80000u     /* replay the frame's launches as a captured hipGraph         */

/* Frame buffers of one resolution (targetResolutionUpdated, CudaTracerRenderer.cpp:735-793).
 * Per-light buffers grow on the first frame that needs them. */
phx_status phx_renderer_create(PhxContext* ctx, uint32_t width, uint32_t height,
                               PhxRenderer** outRenderer);
void phx_renderer_release(PhxRenderer* renderer);

/* Renders one frame of `scene` (its materials and lights) seen through `cam`, camera position
 * cam->origin: primary rays (Halton jitter with PHX_FRAME_TAA), closest hit, hit records
 * (interpretRayHitKernel), shading of every light with a shadow ray each (gBufferMaterialKernel),
 * any hit, shadow logic; one bounce ray per 2x2 pixel block (createSecondaryRaysKernel), closest
 * hit, shading x path throughput with shadow rays (materialKernel), any hit, shadow logic added
 * at half resolution; then, with PHX_FRAME_TAA, the TAA resolve (taa.frag:25-71) with camera
 * motion (velocity_gen.frag:49-71): each pixel's velocity is the screen position of its primary
 * hit (for a miss, of its ray direction) under `cam` minus that under the camera of this
 * renderer's previous frame; the FP32 history is read bilinearly at the pixel centre minus the
 * velocity (a position off the image starts the pixel afresh), trusted by velocity and roughness
 * as taa.frag does, and clamped to the 3x3 neighbourhood. The primary hits' shadow rays
 * (coherent: a pixel tile's rays towards one light) and, when width and height are multiples of 8
 * (camera rays in 8x8-tile order), camera rays are traced as packets (PHX_KERNEL_PACKET); bounce
 * rays, and camera rays in row-major order, by the persistent kernel; the bounce hits' shadow
 * rays (scattered origins) one ray per lane (PHX_KERNEL_ONE_RAY_PER_LANE). `rgba`
 * (width*height*4 floats, row-major; host or device with PHX_MEM_DEVICE; may be NULL) receives
 * the output. Random numbers: PCG32 streams keyed by (seed, frameIndex, pass, record). The
 * frame's kernels read the per-frame values (camera, previous camera, jitter, frame index, seed,
 * TAA reset and history buffer) from a small device block written before the frame, so with
 * PHX_FRAME_GRAPH the launches are captured once as a hipGraph -- again whenever the scene,
 * PHX_FRAME_TAA / PHX_FRAME_NO_CLAMP or the context's launch shape change -- and replayed for the
 * following frames, a moving camera included; the output is the same as without it. */
phx_status phx_render_frame(PhxContext* ctx, PhxRenderer* renderer, const PhxScene* scene,
                            const PhxCameraDef* cam, uint32_t frameIndex, uint64_t seed,
                            uint32_t flags, float* rgba);

/* Instances drawn by this renderer's frames besides the frame's scene (NULL: none): camera and
 * bounce rays take the closest of the scene's and the instances' hits (phx_cast_closest_instances
 * with the scene), shadow rays are blocked by either, and a hit of an instance reads its mesh's
 * TriangleData (material index into the frame scene's materials) with the interpolated normal
 * mapped to world space by the inverse transpose of the instance transform. The set must stay
 * valid while it is set (a released set makes phx_render_frame fail with
 * PHX_ERR_INVALID_HANDLE); rebuild and set it again when instances move. */
phx_status phx_renderer_set_instances(PhxRenderer* renderer, const PhxInstanceSet* set);

#define PHX_BUFFER_COLOR 0u       /* float4 per pixel, row-major: the frame before TAA        */
#define PHX_BUFFER_OUTPUT 1u      /* float4 per pixel: the last output (TAA or frame colour)  */
#define PHX_BUFFER_GBUFFER 2u     /* 3 float4 per primary record (pixel_id order)             */
#define PHX_BUFFER_BOUNCE_RAYS 3u /* PhxRayIn per half-resolution record                      */
#define PHX_BUFFER_HISTORY 4u     /* float4 per pixel: TAA history (rgb, sample count)        */
#define PHX_BUFFER_VELOCITY 5u    /* float2 per pixel, row-major: the last TAA frame's velocity */

/* Device address and size of one of the renderer's buffers (valid until the next frame). */
phx_status phx_renderer_buffer(PhxRenderer* renderer, uint32_t which, const void** devPtr,
                               size_t* bytes);

/* Pixel jitter of frame f in [0, 1)^2: the radical inverses of (f % 16) + 1 in bases 2 and 3
 * (HALTON_SEQ, GameScreen.cpp:21-38). The sample of pixel (x, y) lies at (x + jx, y + jy). */
void phx_halton23(uint32_t frameIndex, float* jx, float* jy);

/* ------------------------------------------------------------------------------------------ */
/* Synthetic scenes of the benchmark workloads (host)                                         */
/* ------------------------------------------------------------------------------------------ */

/* `n` random triangles: centroid ~ U[0,1]^3, edge lengths ~ U[eMin, eMax] (resampled until they
 * form a triangle), planar triangle by the law of cosines centred on its centroid, rotated by a
 * uniform random quaternion. Writes n triangles (w = 0) and, if datas != NULL, their data
 * (geometric normal in n0..n2, zero uvs, material 0). */
phx_status phx_scene_random_triangles(uint32_t n, uint64_t seed, float eMin, float eMax,
                                      PhxTriangleVertices* triangles, PhxTriangleData* datas);

/* `count` constant materials: albedo ~ U[0,1]^3 (alpha 1), roughness ~ U[0.05, 1], metallic 0,
 * no textures (iData = -1). */
phx_status phx_scene_materials_random(uint32_t count, uint64_t seed, PhxMaterial* out);

/* `count` point lights (radius 0) at U[0,1] x U[0.3,1.5] x U[0,1], range 3, strength
 * ~ U[0.5, 2]^3, static and dynamic shadows on. */
phx_status phx_scene_lights_random(uint32_t count, uint64_t seed, PhxSphereLight* out);

/* Triangle count of phx_scene_heightfield_boxes. */
uint32_t phx_scene_heightfield_boxes_count(uint32_t grid, uint32_t numBoxes);

/* grid x grid heightfield on [0,1]^2 (x, z), 2 triangles per cell, height = amp * fBm(value
 * noise, 4 octaves, lacunarity 2, gain 0.5), followed by `numBoxes` axis-aligned boxes (12
 * triangles each) resting on the surface with lognormal(mu, sigma) edge lengths per axis.
 * If numMaterials > 0 each triangle gets materialIndex ~ U{0..numMaterials-1}. */
phx_status phx_scene_heightfield_boxes(uint32_t grid, uint32_t numBoxes, uint64_t seed, float amp,
                                       float boxMu, float boxSigma, uint32_t numMaterials,
                                       PhxTriangleVertices* triangles, PhxTriangleData* datas);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PHX_H */
