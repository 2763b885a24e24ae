// SBVH traversal kernels for gfx950 (wave64).
//
// Node layout (64 B, BVHNode.hpp:23-149): four 16-B words
//   w0 = lc.min.xyz, lc.max.x   w1 = lc.max.yz, rc.min.xy   w2 = rc.min.z, rc.max.xyz
//   w3 = (lcIndex, rcIndex, lcNumTriangles, rcNumTriangles)
// read as four 128-bit loads; a triangle (48 B) is three 128-bit loads.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace phx {
namespace {

// Deepest stack the reference-order traversal can need: one pending entry per level of a tree of
// depth <= 64 (enforced at upload) plus the two children of the node being expanded.
constexpr int kStackSize = 72;

struct NodeWords {
	float4 w0, w1, w2;
	int4 w3;
};

__device__ __forceinline__ NodeWords load_node(const float4* __restrict__ nodes, uint32_t i)
{
	NodeWords n;
	const float4* p = nodes + 4u * i;
	n.w0 = p[0];
	n.w1 = p[1];
	n.w2 = p[2];
	n.w3 = reinterpret_cast<const int4*>(p)[3];
	return n;
}

struct Closest {
	uint32_t tri;
	float t, u, v;
};

// Tests the `count` triangles of a leaf starting at `first` (BVHTraversal.hpp:78-93).
__device__ __forceinline__ void test_leaf(const dev::Ray& r, const float4* __restrict__ tris,
                                          uint32_t first, int32_t count, Closest& c,
                                          uint32_t* accepted)
{
	for (int32_t k = 0; k < count; k++) {
		const float4* p = tris + 3u * (first + uint32_t(k));
		float4 p0 = p[0], p1 = p[1], p2 = p[2];
		float t, u, v;
		if (dev::triangle(r, p0, p1, p2, t, u, v)) {
			if (t < c.t && r.tMin <= t && t <= r.tMax) {
				c.tri = first + uint32_t(k);
				if (accepted) (*accepted)++;
				c.t = t;
				c.u = u;
				c.v = v;
			}
		}
	}
}

// Closest hit with castRay's visiting order (BVHTraversal.hpp:37-128): pop a node, decide both
// children against min(tMax, closest.t) taken before either child is processed, then handle the
// left child (push it if inner, test its triangles if a leaf) and then the right child. The
// stack is LIFO, so the right subtree is explored before the left one.
// kCount adds castDebugRay's counters (BVHTraversal.hpp:130-235), summed over the batch into
// counters[0..4] = nodesVisited, trianglesIntersected, triangleIntersectionTests,
// aabbIntersectionTests, aabbIntersections.
template <bool kCount>
__global__ __launch_bounds__(256) void k_closest_reference_order(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t n,
	unsigned long long* __restrict__ counters)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);

	uint32_t stack[kStackSize];
	stack[0] = 0u;
	int sp = 1;
	Closest c = {~0u, FLT_MAX, FLT_MAX, FLT_MAX};
	uint32_t nNodes = 0, nAccepted = 0, nTriTests = 0, nBoxHits = 0;

	while (sp > 0) {
		sp--;
		const NodeWords nd = load_node(nodes, stack[sp]);
		if (kCount) nNodes++;
		float inL, outL, inR, outR;
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
		const float tCurrMax = dev::min_first(r.tMax, c.t);
		const bool visitL = (outL >= inL) && outL > r.tMin && inL < tCurrMax;
		const bool visitR = (outR >= inR) && outR > r.tMin && inR < tCurrMax;

		if (visitL) {
			if (kCount) nBoxHits++;
			const int32_t raw = nd.w3.x;
			const uint32_t idx = uint32_t(raw < 0 ? ~raw : raw);
			if (nd.w3.z == 0) {
				if (sp < kStackSize) stack[sp++] = idx;
			} else {
				test_leaf(r, tris, idx, nd.w3.z, c, kCount ? &nAccepted : nullptr);
				if (kCount) nTriTests += uint32_t(nd.w3.z);
			}
		}
		if (visitR) {
			if (kCount) nBoxHits++;
			const int32_t raw = nd.w3.y;
			const uint32_t idx = uint32_t(raw < 0 ? ~raw : raw);
			if (nd.w3.w == 0) {
				if (sp < kStackSize) stack[sp++] = idx;
			} else {
				test_leaf(r, tris, idx, nd.w3.w, c, kCount ? &nAccepted : nullptr);
				if (kCount) nTriTests += uint32_t(nd.w3.w);
			}
		}
	}
	uint4 out;
	out.x = c.tri;
	out.y = __float_as_uint(c.t);
	out.z = __float_as_uint(c.u);
	out.w = __float_as_uint(c.v);
	hits[i] = out;
	if (kCount) {
		atomicAdd(&counters[0], (unsigned long long)nNodes);
		atomicAdd(&counters[1], (unsigned long long)nAccepted);
		atomicAdd(&counters[2], (unsigned long long)nTriTests);
		atomicAdd(&counters[3], 2ull * nNodes);
		atomicAdd(&counters[4], (unsigned long long)nBoxHits);
	}
}

// Any hit: occluded iff some triangle has tMin <= t < tMax (the reference's shadow predicate,
// RayCastKernel.cu:759-765), with boxes culled by castRay's predicate (hit, tOut > tMin,
// tIn < tMax). The answer does not depend on the visiting order.
__global__ __launch_bounds__(256) void k_anyhit(const float4* __restrict__ nodes,
                                                const float4* __restrict__ tris,
                                                const float4* __restrict__ rays,
                                                uint8_t* __restrict__ inLight, uint32_t n)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);

	uint32_t stack[kStackSize];
	stack[0] = 0u;
	int sp = 1;
	bool occluded = false;

	while (sp > 0 && !occluded) {
		sp--;
		const NodeWords nd = load_node(nodes, stack[sp]);
		float inL, outL, inR, outR;
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
		const bool visit[2] = {(outL >= inL) && outL > r.tMin && inL < r.tMax,
		                       (outR >= inR) && outR > r.tMin && inR < r.tMax};
		for (int side = 0; side < 2 && !occluded; side++) {
			if (!visit[side]) continue;
			const int32_t raw = side == 0 ? nd.w3.x : nd.w3.y;
			const int32_t count = side == 0 ? nd.w3.z : nd.w3.w;
			const uint32_t idx = uint32_t(raw < 0 ? ~raw : raw);
			if (count == 0) {
				if (sp < kStackSize) stack[sp++] = idx;
				continue;
			}
			for (int32_t k = 0; k < count; k++) {
				const float4* p = tris + 3u * (idx + uint32_t(k));
				float t, u, v;
				if (dev::triangle(r, p[0], p[1], p[2], t, u, v) && r.tMin <= t && t < r.tMax) {
					occluded = true;
					break;
				}
			}
		}
	}
	inLight[i] = occluded ? 0 : 1;
}

}  // namespace

hipError_t launchClosestReferenceOrder(const SceneView& s, const PhxRayIn* rays, uint32_t n,
                                       PhxRayHit* hits, unsigned long long* counters,
                                       hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const uint32_t block = 256;
	const uint32_t grid = n / block + (n % block != 0u);
	if (counters)
		hipLaunchKernelGGL(k_closest_reference_order<true>, dim3(grid), dim3(block), 0, stream,
		                   reinterpret_cast<const float4*>(s.nodes),
		                   reinterpret_cast<const float4*>(s.tris),
		                   reinterpret_cast<const float4*>(rays), reinterpret_cast<uint4*>(hits),
		                   n, counters);
	else
		hipLaunchKernelGGL(k_closest_reference_order<false>, dim3(grid), dim3(block), 0, stream,
		                   reinterpret_cast<const float4*>(s.nodes),
		                   reinterpret_cast<const float4*>(s.tris),
		                   reinterpret_cast<const float4*>(rays), reinterpret_cast<uint4*>(hits),
		                   n, counters);
	return hipGetLastError();
}

hipError_t launchAnyHit(const SceneView& s, const PhxRayIn* rays, uint32_t n, uint8_t* inLight,
                        hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const uint32_t block = 256;
	const uint32_t grid = n / block + (n % block != 0u);
	hipLaunchKernelGGL(k_anyhit, dim3(grid), dim3(block), 0, stream,
	                   reinterpret_cast<const float4*>(s.nodes),
	                   reinterpret_cast<const float4*>(s.tris),
	                   reinterpret_cast<const float4*>(rays), inLight, n);
	return hipGetLastError();
}

}  // namespace phx
