// Host-callable launchers of the HIP kernels (defined in the .hip translation units).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "../../include/phx.h"

namespace phx {

// Device view of an uploaded scene.
struct SceneView {
	const PhxBVHNode* nodes;
	const PhxTriangleVertices* tris;
	uint32_t numNodes;
	uint32_t numTris;
};

// Closest hit in the CPU castRay visiting order; one ray per lane. If `counters` is not null the
// castDebugRay counters of the batch are added to counters[0..4] (device memory).
hipError_t launchClosestReferenceOrder(const SceneView& s, const PhxRayIn* rays, uint32_t n,
                                       PhxRayHit* hits, unsigned long long* counters,
                                       hipStream_t stream);

// Any hit (shadow); one ray per lane.
hipError_t launchAnyHit(const SceneView& s, const PhxRayIn* rays, uint32_t n, uint8_t* inLight,
                        hipStream_t stream);

// Pinhole primary rays.
hipError_t launchGenPrimaryRays(const PhxCameraDef& cam, uint32_t w, uint32_t h, float tMin,
                                float tMax, uint32_t order, PhxRayIn* rays, hipStream_t stream);

}  // namespace phx

namespace phx {

// One cosine-weighted diffuse bounce per primary hit (createSecondaryRaysKernel's Lambertian
// branch, MaterialKernel.cu:239-312); misses get the reference's dummy ray (:41-47).
hipError_t launchGenBounceRays(const PhxRayIn* rays, const PhxRayHit* hits,
                               const PhxTriangleData* datas, uint32_t n, uint64_t seed,
                               uint32_t frame, PhxRayIn* out, hipStream_t stream);

}  // namespace phx
