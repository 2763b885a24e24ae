// Host-callable launchers of the HIP kernels (defined in the .hip translation units).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "../../include/phx.h"

namespace phx {

// Device view of an uploaded scene. trisPre holds each triangle as (p0, e1 = p1 - p0,
// e2 = p2 - p0) with the end marker in p0.w.
struct SceneView {
	const PhxBVHNode* nodes;
	const PhxTriangleVertices* trisPre;
	uint32_t numNodes;
	uint32_t numTris;
	uint32_t finiteBounds;  // every node bound within +-2^126 (enables the NaN-free slab path)
	uint32_t topCount;      // nodes [0, topCount) are the top levels in breadth-first order
};

// Material textures of a scene: one device arena holding the linearisation table, the
// descriptors and the texels (rows packed to width * bpp bytes, each image 16-byte aligned).
struct TexDesc {
	uint64_t offset;  // byte offset of texel (0, 0) from `texels`
	uint32_t width, height, pitch, bpp;
	uint32_t pad[2];
};
static_assert(sizeof(TexDesc) == 32, "TexDesc is 32 bytes");

struct TexSet {
	const float* linear = nullptr;  // linear[c] = pow(c / 255, 2.2), rounded once to float
	const TexDesc* descs = nullptr;
	const uint8_t* texels = nullptr;
	uint32_t count = 0;
};

enum class TraceKernel { Persistent, Simple, Packet };

// Launch parameters of the traversal kernels. `work` points at two zeroed uint32 words owned by
// the launching stream (the persistent kernels leave them zeroed); `spill` holds
// persistentSpillWords(gridBlocks) words.
struct TraceLaunch {
	TraceKernel kernel = TraceKernel::Persistent;
	uint32_t gridBlocks = 0;     // persistent grid
	uint32_t blockThreads = 256; // persistent block size (multiple of 64, <= 1024)
	uint32_t topNodes = 0;       // treelet nodes held in LDS (capped by the scene's topCount)
	uint32_t refillMin = 8;      // refill a wave once this many of its lanes are idle
	uint32_t spillLanes = 0;     // lanes the spill area is sized for
	uint32_t helperWaves = 0;    // extra waves per block that only trace offered rays solo
	uint32_t ldsDepth = 16;      // stack entries per lane held in LDS (the rest spill to HBM)
	uint32_t window = 64;        // rays a wave claims per fetch from the shared counter
	uint32_t* work = nullptr;
	uint32_t* spill = nullptr;
};

// Words of the stack spill area for `lanes` resident lanes.
uint32_t persistentSpillWords(uint32_t lanes);
// Stack entries per lane of the persistent kernels (LDS part + spill part).
constexpr uint32_t kPersistentStackEntries = 72;

// Scratch of the allocation-free closest-hit cast: a static device array per GPU holding the
// persistent launch's ray counters (first kNoAllocCounterBytes) and its stack spill area
// (kNoAllocSpillWords words: 40 spilled entries for 131,072 lanes, 20 MiB).
constexpr size_t kNoAllocCounterBytes = 256;
constexpr uint32_t kNoAllocSpillWords = 40u * 131072u;
constexpr int kNoAllocMaxDevices = 64;
// Looks up the scratch array's address on `device` (which must be current) once; phx_create
// calls it. noAllocScratch returns the address found (nullptr before).
hipError_t noAllocScratchInit(int device);
void* noAllocScratch(int device);
// Zeroes n <= 64 words with a one-block kernel (graph-capturable, allocates nothing).
hipError_t launchZeroWords(uint32_t* p, uint32_t n, hipStream_t stream);
// Dynamic LDS bytes of a persistent block.
uint32_t persistentLdsBytes(uint32_t blockThreads, uint32_t topNodes, uint32_t ldsDepth);
// True if `depth` LDS stack entries per lane is a supported split (4..32).
bool validLdsDepth(uint32_t depth);
// Resident persistent blocks per CU for a block size, treelet size, helper waves and LDS stack
// depth (0 if it does not fit).
hipError_t persistentBlocksPerCU(uint32_t blockThreads, uint32_t topNodes, uint32_t helperWaves,
                                 uint32_t ldsDepth, uint32_t* blocks);

// One instance of a two-level scene on the device: its SBVH (the mesh scene's node and
// pre-transformed triangle arrays) and the world-to-object affine transform, rows (A | b).
struct DevInstance {
	const float4* nodes;
	const float4* tris;
	const PhxTriangleData* datas;  // for hit records (may be null)
	float worldToObject[12];
	uint32_t index;    // the instance's position in the caller's list (ties: lower wins)
	uint32_t numTris;  // triangles of the mesh (hit records check indices against it)
};

// Closest hit (inLight == nullptr; hits and optionally instOut written) or any hit (inLight) of
// every ray against the instances under the top tree `top` (one instance per leaf; `inst` in
// top-tree leaf order). `seeded`: hits / inLight hold the static scene's results on entry, which
// win ties (closest) or already occlude (any hit).
hipError_t launchInstances(const PhxBVHNode* top, const DevInstance* inst, uint32_t numInst,
                           const PhxRayIn* rays, uint32_t n, PhxRayHit* hits, uint32_t* instOut,
                           uint8_t* inLight, bool seeded, hipStream_t stream);

// Any hit of every ray (one ray per lane) adding castDebugRay-style counters of the any-hit
// traversal (trianglesIntersected = rays occluded) to counters[0..4].
hipError_t launchAnyHitCounted(const SceneView& s, const PhxRayIn* rays, uint32_t n,
                               uint8_t* inLight, unsigned long long* counters, hipStream_t stream);

// Closest hit in the CPU castRay visiting order. If `counters` is not null the castDebugRay
// counters of the batch are added to counters[0..4] (device memory).
// Counts the floats a (exhaustive: all of 2^-126 <= |a| < 2^126; else every significand at four
// exponents, both signs) whose fast reciprocal differs from IEEE 1/a; adds them to *mismatches.
// Returns the number of inputs compared in *tested.
hipError_t launchCheckReciprocal(bool exhaustive, unsigned long long* mismatches, uint64_t* tested,
                                 hipStream_t stream);

hipError_t launchClosest(const SceneView& s, const PhxRayIn* rays, uint32_t n, PhxRayHit* hits,
                         const TraceLaunch& launch, unsigned long long* counters,
                         hipStream_t stream);

// The persistent closest-hit launch of `launch`'s shape, also writing each ray's device clock
// (wall_clock64) when its result is written to clock[i] and the earliest block start to clock[n]
// (which the caller sets to ~0 first).
hipError_t launchClosestClock(const SceneView& s, const PhxRayIn* rays, uint32_t n, PhxRayHit* hits,
                              const TraceLaunch& launch, unsigned long long* clock,
                              hipStream_t stream);

// Any hit (shadow rays).
hipError_t launchAnyHit(const SceneView& s, const PhxRayIn* rays, uint32_t n, uint8_t* inLight,
                        const TraceLaunch& launch, hipStream_t stream);

// Device descriptor of one batch of a fused launch (48 B).
struct BatchDesc {
	const float4* nodes;
	const float4* tris;       // pre-transformed triangles
	const float4* rays;
	void* out;                // PhxRayHit* (closest) or uint8_t* (any hit)
	uint32_t count;
	uint32_t kind;            // 0 closest hit, 1 any hit
	uint32_t finiteBounds;
	uint32_t pad;
};
static_assert(sizeof(BatchDesc) == 48, "BatchDesc layout");

// Per-batch completion tracking of a fused launch (device memory).
struct BatchClock {
	unsigned long long start;      // earliest block start of the running launch (~0 between launches)
	unsigned long long lastStart;  // start of the last completed launch
	uint32_t done[1024];           // rays completed per batch (reset by the launch's last wave)
	unsigned long long finish[1024];  // device time of each batch's last completed ray (max)
};

constexpr uint32_t kMaxFusedBatches = 1024;

uint32_t batchLdsBytes(uint32_t blockThreads, uint32_t topNodes, uint32_t numBatches,
                       bool tracked, uint32_t ldsDepth);
// Resident blocks per CU of the fused kernel for a shape and batch count, with or without
// completion tracking (0 if it does not fit).
hipError_t batchBlocksPerCU(uint32_t blockThreads, uint32_t topNodes, uint32_t helperWaves,
                            uint32_t ldsDepth, uint32_t numBatches, bool tracked,
                            uint32_t* blocks);

// One persistent launch over `numBatches` batches; offsets[b] = first ray of batch b in the fused
// index space, offsets[numBatches] = total. `clock` may be null (no completion tracking).
hipError_t launchBatches(const BatchDesc* descs, const uint32_t* offsets, uint32_t numBatches,
                         uint32_t total, const TraceLaunch& launch, const float4* treeletNodes,
                         uint32_t treeletCount, BatchClock* clock, hipStream_t stream);

// Per-frame values of the wavefront frame. Its kernels read them from device memory, so the
// frame's launches are the same for every frame and can be captured once and replayed as a
// hipGraph with only this block rewritten (by launchSetFrameParams, before the frame).
struct FrameParams {
	float jx, jy;       // sub-pixel jitter of the primary rays
	uint32_t frame;     // frame index (random-number streams)
	uint32_t taaReset;  // TAA: start the history afresh
	uint64_t seed;
	uint32_t histIn;    // TAA: history (and velocity) buffer read (0 or 1); the other is written
	uint32_t pad;
	PhxCameraDef cam;      // this frame's camera
	PhxCameraDef prevCam;  // the previous frame's camera (TAA velocity)
};
static_assert(sizeof(FrameParams) == 128, "FrameParams layout");

// Writes `value` to *dst on the stream (a one-thread kernel: its argument is copied at launch).
hipError_t launchSetFrameParams(FrameParams* dst, const FrameParams& value, hipStream_t stream);

// Pinhole primary rays. With `fp` the camera and jitter are read from fp->cam, fp->jx, fp->jy
// instead of cam, jx, jy.
hipError_t launchGenPrimaryRays(const PhxCameraDef& cam, uint32_t w, uint32_t h, float jx,
                                float jy, float tMin, float tMax, uint32_t order, PhxRayIn* rays,
                                hipStream_t stream, const FrameParams* fp = nullptr);

// Wavefront frame kernels (frame_kernels.hip). Hit records are 3 float4 each:
// (position, roughness) (normal, metallic) (albedo, alpha; alpha < 0 = no hit).
// A hit whose triangle index is not below numTris gets the miss record. With instIds, a hit of
// instance k < numInst (inst in instance-index order) reads that instance's mesh data.
hipError_t launchInterpret(const PhxRayIn* rays, const PhxRayHit* hits, uint32_t n,
                           const PhxTriangleData* datas, uint32_t numTris, const PhxMaterial* mats,
                           uint32_t numMats, const TexSet& tex, float4* info, hipStream_t stream,
                           const uint32_t* instIds = nullptr, const DevInstance* inst = nullptr,
                           uint32_t numInst = 0);
// The seed and frame index of the random-number streams, and the camera position, come from fp.
hipError_t launchShade(const float4* info, const PhxRayIn* rays, const float4* throughput,
                       const PhxSphereLight* lights, uint32_t numLights, uint32_t count,
                       const FrameParams* fp, uint32_t purpose, float4* contrib, PhxRayIn* shadow,
                       hipStream_t stream);
hipError_t launchGenSecondary(const float4* gbuf, uint32_t w, uint32_t h, const FrameParams* fp,
                              PhxRayIn* rays, float4* throughput, hipStream_t stream);
hipError_t launchShadowLogic(float4* img, uint32_t w, uint32_t h, uint32_t scale, bool addTo,
                             const uint8_t* inLight, const float4* contrib, uint32_t numLights,
                             hipStream_t stream);
// TAA resolve with camera motion: writes every pixel's velocity (fp->cam against fp->prevCam)
// to vel[1 - fp->histIn], reprojects the history hist[fp->histIn] and the previous velocity
// vel[fp->histIn] (both ignored if fp->taaReset), writes the other history buffer.
hipError_t launchTaa(const float4* cur, float4* hist0, float4* hist1, float2* vel0, float2* vel1,
                     const float4* gbuf, const PhxRayIn* prim, uint32_t w, uint32_t h,
                     bool clampHistory, const FrameParams* fp, float4* out, hipStream_t stream);

}  // namespace phx

namespace phx {

// One cosine-weighted diffuse bounce per primary hit (createSecondaryRaysKernel's Lambertian
// branch, MaterialKernel.cu:239-312); misses, and hits whose triangle index is not below
// numTris, get the reference's dummy ray (:41-47).
hipError_t launchGenBounceRays(const PhxRayIn* rays, const PhxRayHit* hits,
                               const PhxTriangleData* datas, uint32_t numTris, uint32_t n,
                               uint64_t seed, uint32_t frame, PhxRayIn* out, hipStream_t stream);

}  // namespace phx
