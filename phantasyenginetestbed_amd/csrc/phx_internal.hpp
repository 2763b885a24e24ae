// Internal declarations shared by the host-side translation units of the phx library.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/phx.h"

static_assert(sizeof(PhxBVHNode) == 64, "PhxBVHNode must match BVHNode (64 B)");
static_assert(sizeof(PhxTriangleVertices) == 48, "PhxTriangleVertices must match (48 B)");
static_assert(sizeof(PhxTriangleData) == 64, "PhxTriangleData must match (64 B)");
static_assert(sizeof(PhxRayIn) == 32, "PhxRayIn must match RayIn (32 B)");
static_assert(sizeof(PhxRayHit) == 16, "PhxRayHit must match RayHit (16 B)");
static_assert(sizeof(PhxMaterial) == 48, "PhxMaterial must match Material (48 B)");
static_assert(sizeof(PhxSphereLight) == 36, "PhxSphereLight must match SphereLight (36 B)");
static_assert(sizeof(PhxCameraDef) == 48, "PhxCameraDef must match CameraDef (48 B)");

namespace phx {

// Records `status` and a printf-style message as the calling thread's last error; returns status.
phx_status fail(phx_status status, const char* fmt, ...);

// Clears the calling thread's last error and returns PHX_OK.
phx_status ok();

// Runs f (which returns a phx_status) and turns a C++ exception -- a failed host allocation, or a
// thread that could not be started -- into PHX_ERR_OUT_OF_MEMORY, so that none crosses the C ABI.
template <class F>
phx_status guarded(const char* what, F&& f)
{
	try {
		return f();
	} catch (const std::bad_alloc&) {
		return fail(PHX_ERR_OUT_OF_MEMORY, "%s: out of host memory", what);
	} catch (const std::exception& e) {
		return fail(PHX_ERR_OUT_OF_MEMORY, "%s: %s", what, e.what());
	}
}

// Host-side result of the SBVH build: the three arrays of the reference's BVH container
// (ray_tracer_common/BVH.hpp:19-26) after sanitizeBVH.
struct BuiltBVH {
	std::vector<PhxBVHNode> nodes;
	std::vector<PhxTriangleVertices> triangles;
	std::vector<PhxTriangleData> datas;
};

// Builds the sanitized SBVH; see phx_bvh_build.
phx_status buildSBVH(const PhxTriangleVertices* tris, const PhxTriangleData* datas, uint32_t n,
                     uint32_t numThreads, BuiltBVH& out);

// Structural check of a sanitized BVH; see phx_bvh_validate.
phx_status validateBVH(const PhxBVHNode* nodes, uint32_t numNodes,
                       const PhxTriangleVertices* tris, uint32_t numTris, uint32_t* depthOut);

}  // namespace phx
