// ref_driver.cpp -- C entry points around the reference's own CPU ray-query code, compiled from
// the reference sources where they lie (see oracle/Makefile, target `ref`). TEST INFRASTRUCTURE
// ONLY: it writes the golden fixtures under tests/golden and serves as the "reference" CPU
// baseline in bench.py. It includes the reference headers; it contains no reference source.
//
// Calls: buildStaticFrom + sanitizeBVH (ray_tracer_common/StaticBVHBuilder.cpp:213-260),
// castRay<144> and castDebugRay<144> (ray_tracer_common/BVHTraversal.hpp:37-235), and the CPU
// renderer's primary-ray formula on generateCameraDef (GenerateRays.cpp:12-35,
// CPUTracerRenderer.cpp:93-97).
#include <atomic>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include <unistd.h>

#include "phantasy_engine/ray_tracer_common/BVHTraversal.hpp"
#include "phantasy_engine/ray_tracer_common/GenerateRays.hpp"
#include "phantasy_engine/ray_tracer_common/StaticBVHBuilder.hpp"

using namespace phe;

struct RefRay {
	float o[3];
	float tMin;
	float d[3];
	float tMax;
};

struct RefHit {
	uint32_t tri;
	float t, u, v;
};

extern "C" {

// Builds the sanitized SBVH of n triangles (48-B TriangleVertices, 64-B TriangleData). The
// reference prints build statistics to stdout; they are redirected to stderr here.
void* phxref_build(const void* tris, const void* datas, uint32_t n, uint32_t* numNodes,
                   uint32_t* numTris)
{
	DynArray<TriangleVertices> inTris;
	DynArray<TriangleData> inDatas;
	inTris.setCapacity(n);
	inDatas.setCapacity(n);
	const TriangleVertices* tv = static_cast<const TriangleVertices*>(tris);
	const TriangleData* td = static_cast<const TriangleData*>(datas);
	for (uint32_t i = 0; i < n; i++) {
		inTris.add(tv[i]);
		TriangleData d;
		std::memset(&d, 0, sizeof(d));
		if (td) d = td[i];
		inDatas.add(d);
	}
	fflush(stdout);
	int saved = dup(1);
	dup2(2, 1);
	BVH* bvh = new BVH(buildStaticFrom(inTris, inDatas));
	sanitizeBVH(*bvh);
	fflush(stdout);
	dup2(saved, 1);
	close(saved);
	*numNodes = bvh->nodes.size();
	*numTris = bvh->triangleVerts.size();
	return bvh;
}

void phxref_get(void* handle, void* nodes, void* tris, void* datas)
{
	BVH* bvh = static_cast<BVH*>(handle);
	std::memcpy(nodes, bvh->nodes.data(), bvh->nodes.size() * sizeof(BVHNode));
	std::memcpy(tris, bvh->triangleVerts.data(), bvh->triangleVerts.size() * sizeof(TriangleVertices));
	if (datas)
		std::memcpy(datas, bvh->triangleDatas.data(), bvh->triangleDatas.size() * sizeof(TriangleData));
}

void phxref_free(void* handle) { delete static_cast<BVH*>(handle); }

// castRay<144> per ray with the ray's tMin/tMax; `numThreads` workers pull chunks of 1024 rays
// from an atomic counter (0 = all cores).
void phxref_cast(const void* nodes, const void* tris, const RefRay* rays, uint32_t n,
                 RefHit* hits, uint32_t numThreads)
{
	const BVHNode* nd = static_cast<const BVHNode*>(nodes);
	const TriangleVertices* tv = static_cast<const TriangleVertices*>(tris);
	if (numThreads == 0) numThreads = std::max(1u, std::thread::hardware_concurrency());
	std::atomic<uint32_t> next{0};
	auto work = [&]() {
		for (;;) {
			uint32_t b = next.fetch_add(1024);
			if (b >= n) break;
			uint32_t e = std::min(n, b + 1024);
			for (uint32_t i = b; i < e; i++) {
				const RefRay& r = rays[i];
				Ray ray(vec3(r.o[0], r.o[1], r.o[2]), vec3(r.d[0], r.d[1], r.d[2]));
				RayCastResult res = castRay(nd, tv, ray, r.tMin, r.tMax);
				hits[i] = {res.triangleIndex, res.t, res.u, res.v};
			}
		}
	};
	std::vector<std::thread> th;
	for (uint32_t t = 1; t < numThreads; t++) th.emplace_back(work);
	work();
	for (auto& t : th) t.join();
}

// castRay<144> per ray split the way the reference's CPU renderer splits a frame
// (cpu_tracer_renderer/CPUTracerRenderer.cpp:77-87): `numThreads` threads, thread i takes the
// contiguous band [i * (n / numThreads), (i + 1) * (n / numThreads)), the last one also the
// remainder; no work stealing.
void phxref_cast_bands(const void* nodes, const void* tris, const RefRay* rays, uint32_t n,
                       RefHit* hits, uint32_t numThreads)
{
	const BVHNode* nd = static_cast<const BVHNode*>(nodes);
	const TriangleVertices* tv = static_cast<const TriangleVertices*>(tris);
	if (numThreads == 0) numThreads = 1;
	const uint32_t per = n / numThreads;
	auto band = [&](uint32_t t) {
		const uint32_t b = t * per, e = t + 1 == numThreads ? n : b + per;
		for (uint32_t i = b; i < e; i++) {
			const RefRay& r = rays[i];
			Ray ray(vec3(r.o[0], r.o[1], r.o[2]), vec3(r.d[0], r.d[1], r.d[2]));
			RayCastResult res = castRay(nd, tv, ray, r.tMin, r.tMax);
			hits[i] = {res.triangleIndex, res.t, res.u, res.v};
		}
	};
	std::vector<std::thread> th;
	for (uint32_t t = 0; t < numThreads; t++) th.emplace_back(band, t);
	for (auto& t : th) t.join();
}

// castDebugRay<144>: hits plus per-ray counters (nodesVisited, trianglesIntersected,
// triangleIntersectionTests, aabbIntersectionTests, aabbIntersections).
void phxref_cast_debug(const void* nodes, const void* tris, const RefRay* rays, uint32_t n,
                       RefHit* hits, uint32_t* counters)
{
	const BVHNode* nd = static_cast<const BVHNode*>(nodes);
	const TriangleVertices* tv = static_cast<const TriangleVertices*>(tris);
	for (uint32_t i = 0; i < n; i++) {
		const RefRay& r = rays[i];
		Ray ray(vec3(r.o[0], r.o[1], r.o[2]), vec3(r.d[0], r.d[1], r.d[2]));
		DebugRayCastData dbg;
		RayCastResult res = castDebugRay(nd, tv, ray, &dbg, r.tMin, r.tMax);
		hits[i] = {res.triangleIndex, res.t, res.u, res.v};
		uint32_t* c = counters + 5 * size_t(i);
		c[0] = dbg.nodesVisited;
		c[1] = dbg.trianglesIntersected;
		c[2] = dbg.triangleIntersectionTests;
		c[3] = dbg.aabbIntersectionTests;
		c[4] = dbg.aabbIntersections;
	}
}

// generateCameraDef + the CPU renderer's per-pixel ray direction, row-major.
void phxref_primary_rays(const float pos[3], const float dir[3], const float up[3], float fovRad,
                         uint32_t w, uint32_t h, float tMin, float tMax, RefRay* rays,
                         float camOut[12])
{
	vec2 res = vec2(float(w), float(h));
	CameraDef cam = generateCameraDef(vec3(pos[0], pos[1], pos[2]), vec3(dir[0], dir[1], dir[2]),
	                                  vec3(up[0], up[1], up[2]), fovRad, res);
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			vec2 loc = vec2(float(x), float(y));
			vec2 locNormalized = loc / res;
			vec2 centerOffsCoord = locNormalized * 2.0f - vec2(1.0f);
			vec3 rayDir = normalize(cam.dir + centerOffsCoord.x * cam.dX + centerOffsCoord.y * cam.dY);
			RefRay& r = rays[size_t(y) * w + x];
			r = {{cam.origin.x, cam.origin.y, cam.origin.z}, tMin,
			     {rayDir.x, rayDir.y, rayDir.z}, tMax};
		}
	const float c[12] = {cam.origin.x, cam.origin.y, cam.origin.z, cam.dir.x, cam.dir.y, cam.dir.z,
	                     cam.dX.x,     cam.dX.y,     cam.dX.z,     cam.dY.x,  cam.dY.y,  cam.dY.z};
	std::memcpy(camOut, c, sizeof(c));
}

}  // extern "C"
