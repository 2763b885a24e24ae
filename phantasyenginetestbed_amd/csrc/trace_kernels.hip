// SBVH traversal kernels for gfx950 (wave64).
//
// Node layout (64 B, BVHNode.hpp:23-149): four 16-B words read as four 128-bit loads
//   w0 = lc.min.xyz, lc.max.x   w1 = lc.max.yz, rc.min.xy   w2 = rc.min.z, rc.max.xyz
//   w3 = (lcIndex, rcIndex, lcNumTriangles, rcNumTriangles)
// Triangles are the pre-transformed copy made at upload: (p0, e1 = p1 - p0, e2 = p2 - p0), three
// 128-bit loads per triangle.
//
// Visiting order of every closest-hit kernel here is castRay's (BVHTraversal.hpp:37-128): pop a
// node, decide both children against min(tMax, closest.t) taken before either is processed,
// handle the left child (push if inner, test its triangles if a leaf), then the right child; the
// stack is LIFO. The persistent kernels keep the node castRay would pop next in a register
// ("next") and the rest of castRay's stack in LDS, so the order -- and every hit -- is unchanged.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace phx {
namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kBlock = 256;

// Address-space views: records are loaded through pointers the compiler knows to be global (or
// LDS), so the loads are global_load / ds_read rather than flat loads, whose latency is higher
// and which wait on both the vector-memory and LDS counters.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const float4 __attribute__((address_space(1))) GlobalF4;
typedef const float __attribute__((address_space(1))) GlobalF1;
typedef const float4 __attribute__((address_space(3))) LdsF4;
#else  // the host pass only needs the declarations to parse
typedef const float4 GlobalF4;
typedef const float GlobalF1;
typedef const float4 LdsF4;
#endif
__device__ __forceinline__ GlobalF4* as_global(const float4* p) { return (GlobalF4*)(p); }
__device__ __forceinline__ GlobalF1* as_global(const float* p) { return (GlobalF1*)(p); }
__device__ __forceinline__ LdsF4* as_lds(const float4* p) { return (LdsF4*)(p); }

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

struct Counts {
	uint32_t nodes = 0, accepted = 0, triTests = 0, boxHits = 0;
};

// Tests the `count` triangles of a leaf starting at `first` (BVHTraversal.hpp:78-93).
template <bool kCount>
__device__ __forceinline__ void closest_leaf(const dev::Ray& r, const float4* __restrict__ tris,
                                             uint32_t first, int32_t count, Closest& c, Counts& k)
{
	for (int32_t j = 0; j < count; j++) {
		const float4* p = tris + 3u * (first + uint32_t(j));
		const float4 p0 = p[0], e1 = p[1], e2 = p[2];
		float t, u, v;
		if (dev::triangle_pre(r, p0, e1, e2, t, u, v) && t < c.t && r.tMin <= t && t <= r.tMax) {
			c.tri = first + uint32_t(j);
			c.t = t;
			c.u = u;
			c.v = v;
			if (kCount) k.accepted++;
		}
	}
	if (kCount) k.triTests += uint32_t(count);
}

// Any-hit leaf: true iff a triangle has tMin <= t < tMax (RayCastKernel.cu:759-765).
__device__ __forceinline__ bool anyhit_leaf(const dev::Ray& r, const float4* __restrict__ tris,
                                            uint32_t first, int32_t count)
{
	for (int32_t j = 0; j < count; j++) {
		const float4* p = tris + 3u * (first + uint32_t(j));
		float t, u, v;
		if (dev::triangle_pre(r, p[0], p[1], p[2], t, u, v) && r.tMin <= t && t < r.tMax)
			return true;
	}
	return false;
}

__device__ __forceinline__ uint32_t child_index(int32_t raw)
{
	return uint32_t(raw < 0 ? ~raw : raw);
}

__device__ __forceinline__ void add_counts(unsigned long long* counters, const Counts& k)
{
	atomicAdd(&counters[0], (unsigned long long)k.nodes);
	atomicAdd(&counters[1], (unsigned long long)k.accepted);
	atomicAdd(&counters[2], (unsigned long long)k.triTests);
	atomicAdd(&counters[3], 2ull * k.nodes);
	atomicAdd(&counters[4], (unsigned long long)k.boxHits);
}

// ---- one ray per lane -------------------------------------------------------------------------

// Deepest castRay stack for a tree of depth <= 64 (enforced at upload): one pending entry per
// level plus the two children of the node being expanded.
constexpr int kSimpleStack = 72;

// castRay<144> of one ray (BVHTraversal.hpp:37-128) into `c` (which holds the miss result on
// entry): pop, decide both children, left first; a visited leaf is tested at once, a visited
// inner node pushed.
// castRay's pending stack for one lane: its first kLds entries in the lane's LDS column (entry e
// at col[e * kBlock], one 4-B word per lane per row, so a wave's accesses are conflict-free), the
// rest in private memory. A pop is on the path to the next node's load, and an LDS pop returns
// sooner than a private (scratch) one.
template <int kLds>
struct ColumnStack {
	uint32_t* col;
	uint32_t deep[kSimpleStack - kLds];
	__device__ __forceinline__ void put(int e, uint32_t v)
	{
		if (e < kLds)
			col[e * kBlock] = v;
		else
			deep[e - kLds] = v;
	}
	__device__ __forceinline__ uint32_t get(int e) const
	{
		return e < kLds ? col[e * kBlock] : deep[e - kLds];
	}
};

// LDS stack entries per lane: 24 for one-ray-per-lane casts (24 KB per 256-thread block); 12 each
// for an instance cast's top-tree stack and its entered instance's castRay stack (24 KB, so the
// five blocks per CU of its launch fit). Larger splits lower the blocks a CU holds and measured
// slower.
constexpr int kSimpleLds = 24;
constexpr int kInstanceLds = 12;

template <bool kCount, class Stack>
__device__ __forceinline__ void closest_one(const float4* __restrict__ nodes,
                                            const float4* __restrict__ tris, const dev::Ray& r,
                                            Closest& c, Counts& k, Stack& stack)
{
	stack.put(0, 0u);
	int sp = 1;
	while (sp > 0) {
		sp--;
		const NodeWords nd = load_node(nodes, stack.get(sp));
		if (kCount) k.nodes++;
		float inL, outL, inR, outR;
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
		const float tCurrMax = dev::min_first(r.tMax, c.t);
		const bool visitL = (outL >= inL) && outL > r.tMin && inL < tCurrMax;
		const bool visitR = (outR >= inR) && outR > r.tMin && inR < tCurrMax;
		if (visitL) {
			if (kCount) k.boxHits++;
			if (nd.w3.z == 0) {
				if (sp < kSimpleStack) stack.put(sp++, child_index(nd.w3.x));
			} else {
				closest_leaf<kCount>(r, tris, child_index(nd.w3.x), nd.w3.z, c, k);
			}
		}
		if (visitR) {
			if (kCount) k.boxHits++;
			if (nd.w3.w == 0) {
				if (sp < kSimpleStack) stack.put(sp++, child_index(nd.w3.y));
			} else {
				closest_leaf<kCount>(r, tris, child_index(nd.w3.y), nd.w3.w, c, k);
			}
		}
	}
}

// Any-hit leaf that counts its triangle tests (up to and including the first hit).
__device__ __forceinline__ bool anyhit_leaf_counted(const dev::Ray& r,
                                                    const float4* __restrict__ tris,
                                                    uint32_t first, int32_t count, Counts& k)
{
	for (int32_t j = 0; j < count; j++) {
		const float4* p = tris + 3u * (first + uint32_t(j));
		float t, u, v;
		k.triTests++;
		if (dev::triangle_pre(r, p[0], p[1], p[2], t, u, v) && r.tMin <= t && t < r.tMax) {
			k.accepted++;
			return true;
		}
	}
	return false;
}

// Any hit of one ray: castRay's traversal with tCurrMax = tMax, stopping at the first triangle
// with tMin <= t < tMax. kCount adds castDebugRay-style counters of that traversal to k.
template <bool kCount, class Stack>
__device__ __forceinline__ bool anyhit_one(const float4* __restrict__ nodes,
                                           const float4* __restrict__ tris, const dev::Ray& r,
                                           Counts& k, Stack& stack)
{
	stack.put(0, 0u);
	int sp = 1;
	bool occluded = false;
	while (sp > 0 && !occluded) {
		sp--;
		const NodeWords nd = load_node(nodes, stack.get(sp));
		if (kCount) k.nodes++;
		float inL, outL, inR, outR;
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
		const bool visitL = (outL >= inL) && outL > r.tMin && inL < r.tMax;
		const bool visitR = (outR >= inR) && outR > r.tMin && inR < r.tMax;
		if (visitL) {
			if (kCount) k.boxHits++;
			if (nd.w3.z == 0) {
				if (sp < kSimpleStack) stack.put(sp++, child_index(nd.w3.x));
			} else {
				occluded = kCount ? anyhit_leaf_counted(r, tris, child_index(nd.w3.x), nd.w3.z, k)
				                  : anyhit_leaf(r, tris, child_index(nd.w3.x), nd.w3.z);
			}
		}
		if (visitR && !occluded) {
			if (kCount) k.boxHits++;
			if (nd.w3.w == 0) {
				if (sp < kSimpleStack) stack.put(sp++, child_index(nd.w3.y));
			} else {
				occluded = kCount ? anyhit_leaf_counted(r, tris, child_index(nd.w3.y), nd.w3.w, k)
				                  : anyhit_leaf(r, tris, child_index(nd.w3.y), nd.w3.w);
			}
		}
	}
	return occluded;
}

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_closest_simple(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t n,
	unsigned long long* __restrict__ counters)
{
	__shared__ uint32_t stackLds[kSimpleLds * kBlock];
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);
	Closest c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
	Counts k;
	ColumnStack<kSimpleLds> st;
	st.col = stackLds + threadIdx.x;
	closest_one<kCount>(nodes, tris, r, c, k, st);
	hits[i] = make_uint4(c.tri, __float_as_uint(c.t), __float_as_uint(c.u), __float_as_uint(c.v));
	if (kCount) add_counts(counters, k);
}

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_anyhit_simple(const float4* __restrict__ nodes,
                                                          const float4* __restrict__ tris,
                                                          const float4* __restrict__ rays,
                                                          uint8_t* __restrict__ inLight, uint32_t n,
                                                          unsigned long long* __restrict__ counters)
{
	__shared__ uint32_t stackLds[kSimpleLds * kBlock];
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);
	Counts k;
	ColumnStack<kSimpleLds> st;
	st.col = stackLds + threadIdx.x;
	inLight[i] = anyhit_one<kCount>(nodes, tris, r, k, st) ? 0 : 1;
	if (kCount) add_counts(counters, k);
}

// ---- instances (two-level traversal) ------------------------------------------------------------
//
// A top tree over the instances' world boxes (one instance per leaf, PhxBVHNode layout) is walked
// in castRay's order. An instance is entered iff the ray's slab test on its world box hits with
// tOut > tMin and tIn <= min(tMax, best t) -- `<=`, so an instance that could tie the best hit is
// still traced. Each entered instance runs castRay on its own SBVH with the ray moved into object
// space (o' = A o + b, d' = A d, the same t range; t is the same parameter in both spaces), from a
// miss, so its result is the one castRay gives for that instance alone. The closest of those
// results wins, the lower instance index on equal t. With `hits` holding a static-scene result on
// entry (seeded), that result counts as instance -1: an instance must be strictly closer.

__device__ __forceinline__ dev::Ray object_ray(const DevInstance& in, const dev::Ray& w)
{
	const float* A = in.worldToObject;
	float o[3], d[3];
	for (int k = 0; k < 3; k++) {
		float a = A[4 * k] * w.ox + A[4 * k + 1] * w.oy;
		a = a + A[4 * k + 2] * w.oz;
		o[k] = a + A[4 * k + 3];
		float b = A[4 * k] * w.dx + A[4 * k + 1] * w.dy;
		d[k] = b + A[4 * k + 2] * w.dz;
	}
	return dev::make_ray(make_float4(o[0], o[1], o[2], w.tMin), make_float4(d[0], d[1], d[2], w.tMax));
}

// The kernel is a per-lane state machine: every iteration a lane does one unit of its ray's work
// -- a top-tree node, entering an instance, or one node of the entered instance's castRay (with
// that node's leaf triangles) -- and a lane whose ray is done takes the next ray of its wave's
// range (`raysPerWave` consecutive rays per wave), so the lanes of a wave do not wait for the
// wave's costliest ray. (Written as nested loops, one ray per lane, the same casts took 3.8x as
// long for closest hit and 2.7x for any hit on 1M rays into 300 instances.) Per ray, the order is
// the one above: both children of a top node are decided with the best t on reaching it, inner
// children are pushed left then right, the visited leaves' instances are entered left leaf
// first, and each entry's castRay result is merged when its traversal ends.
template <bool kAnyHit>
__global__ __launch_bounds__(kBlock) void k_trace_instances(
	const float4* __restrict__ top, const DevInstance* __restrict__ inst, uint32_t numInst,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t* __restrict__ instOut,
	uint8_t* __restrict__ inLight, uint32_t n, uint32_t seeded, uint32_t raysPerWave)
{
	__shared__ uint32_t innerLds[kInstanceLds * kBlock];
	__shared__ uint32_t topLds[kInstanceLds * kBlock];
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	uint32_t next = wave * raysPerWave;  // wave-uniform
	const uint32_t end = next < n ? (n - next < raysPerWave ? n : next + raysPerWave) : next;
	ColumnStack<kInstanceLds> inner;
	inner.col = innerLds + threadIdx.x;
	ColumnStack<kInstanceLds> topStk;
	topStk.col = topLds + threadIdx.x;
	uint32_t ray = kInvalid;
	dev::Ray r = dev::make_ray(make_float4(0.f, 0.f, 0.f, 0.f), make_float4(1.f, 1.f, 1.f, 0.f));
	dev::Ray o = r;
	Closest best = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX}, c = best;
	uint32_t bestInst = kInvalid, curId = 0u;
	bool seededHit = false, occluded = false, inInner = false;
	int tsp = 0, isp = 0;
	uint32_t pend0 = 0u, pend0End = 0u, pend1 = 0u, pend1End = 0u;  // leaf instance ranges
	const float4* iNodes = nullptr;
	const float4* iTris = nullptr;
	Counts none;
	while (true) {
		// Lanes without a ray take the next rays of the wave's range, in lane order.
		const uint64_t idle = __ballot(ray == kInvalid);
		bool fresh = false;
		if (idle != 0 && next < end) {
			const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
			                                                __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
			if (ray == kInvalid && next + rank < end) {
				ray = next + rank;
				fresh = true;
			}
			const uint32_t taken = uint32_t(__popcll(idle));
			next = end - next < taken ? end : next + taken;
		}
		if (__ballot(ray != kInvalid) == 0) break;
		if (ray == kInvalid) continue;
		if (fresh) {
			r = dev::make_ray(rays[2 * size_t(ray)], rays[2 * size_t(ray) + 1]);
			best = Closest{kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			bestInst = kInvalid;
			seededHit = false;
			occluded = false;
			if (!kAnyHit && seeded) {
				const uint4 h = hits[ray];
				best = Closest{h.x, __uint_as_float(h.y), __uint_as_float(h.z), __uint_as_float(h.w)};
				seededHit = h.x != kInvalid;
			}
			if (kAnyHit && seeded) occluded = inLight[ray] == 0;
			tsp = numInst ? 1 : 0;
			topStk.put(0, 0u);
			pend0 = pend0End = pend1 = pend1End = 0u;
			inInner = false;
		}
		// A lane may take a top-tree node, enter the instance it leaves pending and take that
		// instance's first node in one iteration; the order of each ray's work is unchanged.
		bool done = false;
		if (!inInner && !occluded && pend0 >= pend0End && pend1 >= pend1End && tsp > 0) {
			// One top-tree node: decide both children with the best t so far.
			tsp--;
			const NodeWords nd = load_node(top, topStk.get(tsp));
			float inL, outL, inR, outR;
			dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
			dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
			const float lim = kAnyHit ? r.tMax : dev::min_first(r.tMax, best.t);
			const bool visitL = (outL >= inL) && outL > r.tMin && (kAnyHit ? inL < lim : inL <= lim);
			const bool visitR = (outR >= inR) && outR > r.tMin && (kAnyHit ? inR < lim : inR <= lim);
			if (visitL && nd.w3.z == 0 && tsp < kSimpleStack) topStk.put(tsp++, child_index(nd.w3.x));
			if (visitR && nd.w3.w == 0 && tsp < kSimpleStack) topStk.put(tsp++, child_index(nd.w3.y));
			if (visitL && nd.w3.z != 0) {
				pend0 = child_index(nd.w3.x);
				pend0End = pend0 + uint32_t(nd.w3.z);
			}
			if (visitR && nd.w3.w != 0) {
				pend1 = child_index(nd.w3.y);
				pend1End = pend1 + uint32_t(nd.w3.w);
			}
		}
		if (!inInner && !occluded && (pend0 < pend0End || pend1 < pend1End)) {
			// Enter the next instance of the visited leaves (left leaf first).
			const uint32_t j = pend0 < pend0End ? pend0++ : pend1++;
			const DevInstance& in = inst[j];
			curId = in.index;
			iNodes = in.nodes;
			iTris = in.tris;
			o = object_ray(in, r);
			c = Closest{kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			inner.put(0, 0u);
			isp = 1;
			inInner = true;
		}
		if (inInner) {
			// One node of the entered instance's castRay (closest_one / anyhit_one's loop body).
			isp--;
			const NodeWords nd = load_node(iNodes, inner.get(isp));
			float inL, outL, inR, outR;
			dev::slab(o, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
			dev::slab(o, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
			const float lim = kAnyHit ? o.tMax : dev::min_first(o.tMax, c.t);
			const bool visitL = (outL >= inL) && outL > o.tMin && inL < lim;
			const bool visitR = (outR >= inR) && outR > o.tMin && inR < lim;
			bool hitAny = false;
			if (visitL) {
				if (nd.w3.z == 0) {
					if (isp < kSimpleStack) inner.put(isp++, child_index(nd.w3.x));
				} else if (kAnyHit) {
					hitAny = anyhit_leaf(o, iTris, child_index(nd.w3.x), nd.w3.z);
				} else {
					closest_leaf<false>(o, iTris, child_index(nd.w3.x), nd.w3.z, c, none);
				}
			}
			if (visitR && !hitAny) {
				if (nd.w3.w == 0) {
					if (isp < kSimpleStack) inner.put(isp++, child_index(nd.w3.y));
				} else if (kAnyHit) {
					hitAny = anyhit_leaf(o, iTris, child_index(nd.w3.y), nd.w3.w);
				} else {
					closest_leaf<false>(o, iTris, child_index(nd.w3.y), nd.w3.w, c, none);
				}
			}
			if (kAnyHit && hitAny) {
				occluded = true;
				done = true;
			} else if (isp == 0) {
				inInner = false;
				if (!kAnyHit) {
					const bool tieWins = c.t == best.t && !seededHit &&
					                     (bestInst == kInvalid || curId < bestInst);
					if (c.tri != kInvalid && (c.t < best.t || tieWins)) {
						best = c;
						bestInst = curId;
						seededHit = false;
					}
				}
			}
		}
		if (!done && !inInner)
			done = occluded || (pend0 >= pend0End && pend1 >= pend1End && tsp == 0);
		if (done) {
			if (kAnyHit) {
				inLight[ray] = occluded ? 0 : 1;
			} else {
				if (bestInst != kInvalid || !seeded)
					hits[ray] = make_uint4(best.tri, __float_as_uint(best.t), __float_as_uint(best.u),
					                       __float_as_uint(best.v));
				if (instOut) instOut[ray] = bestInst;
			}
			ray = kInvalid;
		}
	}
}

// ---- persistent wavefronts ----------------------------------------------------------------------
//
// A fixed grid of waves pulls rays from a shared counter, one atomic per wavefront for a window
// of `window` rays (rays per fetch, a launch parameter). Every iteration, lanes without a ray are
// refilled from the window when at least `refillMin` of the wave's 64 lanes are idle (always when
// all are). The last wave to finish resets the counter, so every launch starts from zero without a
// memset.
//
// LDS (dynamic) holds, per block:
//   - castRay's pending stack, [wave][entry][lane] (one 4-B column per lane, so each push/pop is a
//     conflict-free 256-B row), `depth` entries deep (the stack split, a launch parameter) plus
//     one row that lanes without a push write to; deeper entries spill to a global array laid out
//     [entry][lane] (coalesced). Trees are at most 64 levels deep (checked at upload), so the
//     pending stack never exceeds kPersistentStack entries;
//   - the top treelet: nodes [0, topCount) of the device node array, which the upload lays out in
//     breadth-first order so they are the top levels of the tree -- where most node visits of
//     incoherent rays happen. They are copied once per block and read from LDS afterwards.

constexpr int kMinLdsDepth = 4;    // smallest stack split; the spill area is sized for it
constexpr int kMaxLdsDepth = 32;
constexpr int kPersistentStack = 72;
constexpr int kMaxBlock = 1024;

struct WaveQueue {
	uint32_t base = 0;   // next unclaimed ray of this wave's window
	uint32_t left = 0;   // rays left in the window
	uint32_t start = 0;  // first ray of the window
	bool exhausted = false;
};

// Assigns ray indices to the idle lanes of the wave (`idle` = ballot of idle lanes). Returns the
// lane's new ray index, or kInvalid. Wave-uniform state lives in `q`.
__device__ __forceinline__ uint32_t refill(WaveQueue& q, uint64_t idle, uint32_t* work,
                                           uint32_t n, bool isIdle, uint32_t window)
{
	const uint32_t need = uint32_t(__popcll(idle));
	const uint32_t take1 = need < q.left ? need : q.left;
	const uint32_t base1 = q.base;
	q.base += take1;
	q.left -= take1;
	uint32_t take2 = 0, base2 = 0;
	if (need > take1) {
		const int leader = __ffsll((unsigned long long)idle) - 1;
		uint32_t b = 0;
		if (int(__lane_id()) == leader) b = atomicAdd(work, window);
		b = __shfl(b, leader);
		if (b >= n) {
			q.exhausted = true;
			q.left = 0;
		} else {
			const uint32_t avail = (n - b) < window ? (n - b) : window;
			take2 = (need - take1) < avail ? (need - take1) : avail;
			base2 = b;
			q.start = b;
			q.base = b + take2;
			q.left = avail - take2;
		}
	}
	if (!isIdle) return kInvalid;
	const uint32_t rank = uint32_t(__popcll(idle & __lanemask_lt()));
	if (rank < take1) return base1 + rank;
	if (rank < take1 + take2) return base2 + (rank - take1);
	return kInvalid;
}

// Called by every wave as it exits: the last one resets the shared counters to zero.
__device__ __forceinline__ void retire_wave(uint32_t* work)
{
	if (__lane_id() == 0) {
		__threadfence();
		const uint32_t totalWaves = gridDim.x * (blockDim.x / 64u);
		if (atomicAdd(&work[1], 1u) == totalWaves - 1u) {
			atomicExch(&work[0], 0u);
			atomicExch(&work[1], 0u);
		}
	}
}

struct LaneStack {
	uint32_t* lds;    // this wave's [entry][lane] block: depth + 1 rows
	uint32_t* spill;  // this lane's spill column
	uint32_t stride;  // spill row stride (lanes in the grid)
	uint32_t lane;
	int depth;        // entries kept in LDS
	int sp;

	// Pushes v if doPush. Lanes that do not push (or whose entry spills) store to the extra row, so
	// the LDS store needs no branch; only a spilling entry takes one.
	__device__ __forceinline__ void push_if(bool doPush, uint32_t v)
	{
		const bool toLds = doPush && sp < depth;
		lds[(toLds ? sp : depth) * 64 + int(lane)] = v;
		if (doPush && sp >= depth && sp < kPersistentStack)
			spill[size_t(sp - depth) * stride] = v;
		sp += doPush ? 1 : 0;
	}
	// Pops if doPop (the value is meaningless otherwise): every lane reads one LDS word -- lanes
	// that do not pop, or whose entry spilled, read the extra row -- and only a spilled entry
	// takes a branch.
	__device__ __forceinline__ uint32_t pop_if(bool doPop)
	{
		sp -= doPop ? 1 : 0;
		uint32_t v = lds[(doPop && sp < depth ? sp : depth) * 64 + int(lane)];
		if (doPop && sp >= depth)
			v = sp < kPersistentStack ? spill[size_t(sp - depth) * stride] : 0u;
		return v;
	}
};

// Visit decision of both children of a node (BVHTraversal.hpp:59-64; any-hit: tCurrMax = tMax).
__device__ __forceinline__ void decide(const dev::Ray& r, bool fastSlab, const NodeWords& nd,
                                       float tCurrMax, bool& visitL, bool& visitR)
{
	float inL, outL, inR, outR;
	if (fastSlab) {
		dev::slab_fast(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab_fast(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
	} else {
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
	}
	visitL = (outL >= inL) && outL > r.tMin && inL < tCurrMax;
	visitR = (outR >= inR) && outR > r.tMin && inR < tCurrMax;
}

// A triangle step tests up to kTriBatch consecutive triangles of the current leaf range, all
// loaded by the same step, in castRay's order (BVHTraversal.hpp:78-93). Leaves on the paths of
// costly incoherent rays hold about five triangles, so this shortens their chains of dependent
// steps. The words of a second triangle past a leaf's last one lie in the next triangle or in the
// 64 B of padding after the array.
constexpr uint32_t kTriBatch = 2;

// Per-lane traversal state of the persistent kernels. Every iteration a lane does one unit of
// work in castRay's per-ray order: expand a node (decide both children, queue the triangles of
// visited leaf children -- left, then right -- and pick the next node), or test up to kTriBatch
// queued triangles. Nodes are expanded only once the queued triangles are done, so each box
// decision sees the closest.t that castRay would see.
struct LaneTrace {
	uint32_t triCur, triEnd;    // triangles being tested
	uint32_t pendCur, pendEnd;  // right-leaf triangles queued behind them
	int32_t next;               // node castRay pops next, or -1
};

__host__ __device__ __forceinline__ uint32_t handoff_bytes(uint32_t blockThreads);

// LDS bytes of a persistent launch: the stacks of every wave, the treelet, the handoff area.
__host__ __device__ __forceinline__ uint32_t persistent_lds_bytes(uint32_t blockThreads,
                                                                  uint32_t topCount,
                                                                  uint32_t depth)
{
	return (blockThreads / 64u) * (depth + 1u) * 64u * 4u + topCount * 64u +
	       handoff_bytes(blockThreads);
}

// ---- packet traversal (coherent rays) ------------------------------------------------------------
//
// For coherent rays -- camera rays in 8x8-tile order, shadow rays towards one light -- a wave traces
// 64 consecutive rays together over one shared stack of (node, lane mask) entries. Each node is
// loaded once for the wave (every lane reads the same address); each lane decides both child boxes
// with its own ray and closest hit, and tests a visited leaf's triangles only where it decided to.
// Per lane, the nodes expanded and the triangles tested, in their order, are exactly castRay's
// (BVHTraversal.hpp:37-128): both child boxes are decided before either child is used, a visited
// leaf's triangles are tested at once, left child first, and inner children are pushed left then
// right, so the right subtree is expanded first; a lane takes part only in the entries it pushed.
// The wave expands the union of its lanes' nodes, which for coherent rays is close to one ray's,
// and every step does the same kind of work. Stack entry e lives in lane e % 64 of the register
// triple e / 64 (node, mask low, mask high); trees are at most 64 levels deep (checked at upload),
// and the stack holds at most one pending left child per level.

constexpr uint32_t kPacketBlock = 256;

template <bool kAnyHit>
__global__ __launch_bounds__(kPacketBlock) void k_packet(const float4* __restrict__ nodes,
                                                         const float4* __restrict__ tris,
                                                         const float4* __restrict__ rays,
                                                         void* __restrict__ out, uint32_t n,
                                                         uint32_t finiteBounds)
{
	const uint32_t lane = __lane_id();
	const uint32_t first = blockIdx.x * kPacketBlock + (threadIdx.x & ~63u);
	if (first >= n) return;  // whole wave
	const uint32_t i = first + lane;
	const bool active = i < n;
	const uint32_t ri = active ? i : first;  // lanes past the end read a valid ray, never used
	const dev::Ray r = dev::make_ray(rays[2 * size_t(ri)], rays[2 * size_t(ri) + 1]);
	// The NaN-free slab path only when every active lane may take it (the decision is uniform).
	const bool fast = __ballot(active && !(finiteBounds != 0u && r.nanFree)) == 0ull;
	Closest c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
	Counts none;
	bool occluded = false;
	uint32_t sN0 = 0u, sL0 = 0u, sH0 = 0u, sN1 = 0u, sL1 = 0u, sH1 = 0u;
	int sp = 0;
	auto push = [&](uint32_t node, uint64_t m) {
		const uint32_t lo = uint32_t(m), hi = uint32_t(m >> 32);
		const bool low = sp < 64;
		const bool mine = int(lane) == (low ? sp : sp - 64);
		sN0 = low && mine ? node : sN0;
		sL0 = low && mine ? lo : sL0;
		sH0 = low && mine ? hi : sH0;
		sN1 = !low && mine ? node : sN1;
		sL1 = !low && mine ? lo : sL1;
		sH1 = !low && mine ? hi : sH1;
		sp++;
	};
	uint32_t cur = 0u;
	uint64_t mask = __ballot(active);
	while (true) {
		const bool in = ((mask >> lane) & 1ull) != 0;
		NodeWords nd = load_node(nodes, cur);
		const int32_t lRaw = __builtin_amdgcn_readfirstlane(nd.w3.x);
		const int32_t rRaw = __builtin_amdgcn_readfirstlane(nd.w3.y);
		const int32_t lCnt = __builtin_amdgcn_readfirstlane(nd.w3.z);
		const int32_t rCnt = __builtin_amdgcn_readfirstlane(nd.w3.w);
		const float tCurrMax = kAnyHit ? r.tMax : dev::min_first(r.tMax, c.t);
		bool visitL, visitR;
		decide(r, fast, nd, tCurrMax, visitL, visitR);
		visitL = visitL && in;
		visitR = visitR && in;
		uint64_t pushL = 0ull, pushR = 0ull;
		if (lCnt == 0)
			pushL = __ballot(visitL);
		else if (visitL)
			kAnyHit ? void(occluded = anyhit_leaf(r, tris, child_index(lRaw), lCnt))
			        : closest_leaf<false>(r, tris, child_index(lRaw), lCnt, c, none);
		if (rCnt == 0)
			pushR = __ballot(visitR && !(kAnyHit && occluded));
		else if (visitR && !(kAnyHit && occluded))
			kAnyHit ? void(occluded = anyhit_leaf(r, tris, child_index(rRaw), rCnt))
			        : closest_leaf<false>(r, tris, child_index(rRaw), rCnt, c, none);
		const uint64_t done = kAnyHit ? __ballot(occluded) : 0ull;
		pushL &= ~done;
		if (pushL != 0ull && sp < 128) push(child_index(lRaw), pushL);
		if (pushR != 0ull && sp < 128) push(child_index(rRaw), pushR);
		// Next: the top entry that still has lanes (any hit: lanes without an occluder).
		mask = 0ull;
		while (mask == 0ull && sp > 0) {
			sp--;
			const bool low = sp < 64;
			const int l = low ? sp : sp - 64;
			cur = uint32_t(__builtin_amdgcn_readlane(int(low ? sN0 : sN1), l));
			const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(low ? sL0 : sL1), l));
			const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(low ? sH0 : sH1), l));
			mask = ((uint64_t(hi) << 32) | lo) & ~done;
		}
		if (mask == 0ull) break;
	}
	if (active) {
		if (kAnyHit)
			static_cast<uint8_t*>(out)[i] = occluded ? 0 : 1;
		else
			static_cast<uint4*>(out)[i] = make_uint4(c.tri, __float_as_uint(c.t),
			                                         __float_as_uint(c.u), __float_as_uint(c.v));
	}
}

// ---- software-pipelined persistent loop -------------------------------------------------------
//
// Each iteration of a lane consumes the record it loaded in the previous iteration and issues the
// load of the next one as soon as its address is known: right after a node's expansion, and for
// a triangle step before the triangles are tested, so the test arithmetic runs while the next
// record is in flight. The loop body is instantiated twice with the two record buffers swapped
// (A -> B, then B -> A), so no record is ever copied between registers -- a copy would wait for
// the load it copies. A new ray is loaded together with the root node and its first iteration
// expands the root, so a ray costs one iteration less than a node of its own would.

constexpr uint32_t kRecIdle = 0, kRecRay = 1, kRecTri = 2, kRecNode = 3;

// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt = 0 only: expcnt 7 and lgkmcnt 15 are
// the "no wait" values of those fields.
constexpr int kWaitVmcntZero = (0x7 << 4) | (0xF << 8);

// One record's six 16-B words: a node (words 0-3), a triangle pair (p0, e1, e2 of the first
// triangle, then of the second), or the root node (words 0-3) with a new ray (words 4-5).
struct Rec {
	float4 w0, w1, w2, w3, w4, w5;
};

struct PipeArgs {
	const float4* nodes;
	const float4* tris;
	const float4* rays;
	uint4* hits;
	uint8_t* inLight;
	uint32_t n;
	uint32_t* work;
	uint32_t refillMin;
	uint32_t topCount;
	bool finiteBounds;
	const float4* treelet;  // LDS copy of nodes [0, topCount)
	const uint32_t* handoff;  // the block's handoff area (LDS)
	uint32_t window;          // rays per fetch
	unsigned long long* clock;  // completion clock per ray (completion-clock launches only)
};

// Loop-carried state of one lane.
struct Tracer {
	dev::Ray r;
	Closest c;
	LaneTrace s;
	uint32_t rayIdx;   // ray being traced
	uint32_t doneIdx;  // ray that ended in the previous step, its result still to be written
	uint32_t kind;    // record in flight (kRec*)
	bool fastSlab;
};

// Writes a ray's result: its closest hit, or (any hit) 0 if occluded -- c.tri == 0 marks a hit
// there -- and 1 if not.
template <bool kAnyHit>
__device__ __forceinline__ void store_result(void* out, uint32_t idx, const Closest& c)
{
	if (kAnyHit)
		static_cast<uint8_t*>(out)[idx] = c.tri == 0u ? 0 : 1;
	else
		static_cast<uint4*>(out)[idx] = make_uint4(c.tri, __float_as_uint(c.t),
		                                           __float_as_uint(c.u), __float_as_uint(c.v));
}
template <bool kAnyHit>
__device__ __forceinline__ void store_result(const PipeArgs& P, uint32_t idx, const Closest& c)
{
	store_result<kAnyHit>(kAnyHit ? static_cast<void*>(P.inLight) : static_cast<void*>(P.hits),
	                      idx, c);
}

// Both triangles of a pair, branch-free, accepted in castRay's order (BVHTraversal.hpp:78-93).
// The second is skipped when the pair holds one triangle (its words are the next triangle's or
// the array's padding).
template <bool kCount>
__device__ __forceinline__ void test_pair_closest(const dev::Ray& r, const Rec& A, uint32_t cnt,
                                                  uint32_t first, Closest& c, Counts& k)
{
	const dev::TriDet d0 = dev::tri_det(r, A.w1, A.w2);
	const dev::TriDet d1 = dev::tri_det(r, A.w4, A.w5);
	float f0 = dev::recip_newton(d0.a), f1 = dev::recip_newton(d1.a);
	if (!(dev::recip_newton_ok(d0.a) && dev::recip_newton_ok(d1.a))) {
		f0 = 1.0f / d0.a;
		f1 = 1.0f / d1.a;
	}
	float t0, u0, v0, t1, u1, v1;
	const bool h0 = dev::tri_finish(r, d0, f0, A.w0, A.w1, A.w2, t0, u0, v0);
	const bool h1 = dev::tri_finish(r, d1, f1, A.w3, A.w4, A.w5, t1, u1, v1) && cnt > 1u;
	const bool acc0 = h0 && t0 < c.t && r.tMin <= t0 && t0 <= r.tMax;
	c.tri = acc0 ? first : c.tri;
	c.t = acc0 ? t0 : c.t;
	c.u = acc0 ? u0 : c.u;
	c.v = acc0 ? v0 : c.v;
	const bool acc1 = h1 && t1 < c.t && r.tMin <= t1 && t1 <= r.tMax;
	c.tri = acc1 ? first + 1u : c.tri;
	c.t = acc1 ? t1 : c.t;
	c.u = acc1 ? u1 : c.u;
	c.v = acc1 ? v1 : c.v;
	if (kCount) {
		k.accepted += uint32_t(acc0) + uint32_t(acc1);
		k.triTests += cnt;
	}
}

// Any-hit pair: true iff one of the triangles has tMin <= t < tMax (RayCastKernel.cu:759-765).
__device__ __forceinline__ bool test_pair_any(const dev::Ray& r, const Rec& A, uint32_t cnt)
{
	const dev::TriDet d0 = dev::tri_det(r, A.w1, A.w2);
	const dev::TriDet d1 = dev::tri_det(r, A.w4, A.w5);
	float f0 = dev::recip_newton(d0.a), f1 = dev::recip_newton(d1.a);
	if (!(dev::recip_newton_ok(d0.a) && dev::recip_newton_ok(d1.a))) {
		f0 = 1.0f / d0.a;
		f1 = 1.0f / d1.a;
	}
	float t0, u0, v0, t1, u1, v1;
	const bool h0 = dev::tri_finish(r, d0, f0, A.w0, A.w1, A.w2, t0, u0, v0);
	const bool h1 = dev::tri_finish(r, d1, f1, A.w3, A.w4, A.w5, t1, u1, v1) && cnt > 1u;
	return (h0 && r.tMin <= t0 && t0 < r.tMax) || (h1 && r.tMin <= t1 && t1 < r.tMax);
}

// Node expansion (selects, not branches): visited leaf children become the triangle range (left
// first, the right one queued behind it), the right inner child is castRay's next pop, a left
// inner child with it is pushed, and with no inner child the stack is popped -- one LDS read for
// every lane (lanes that do not pop read the spare row); only a spilled entry takes a branch.
__device__ __forceinline__ void expand_pipe(const NodeWords& nd, bool visitL, bool visitR,
                                            LaneTrace& s, LaneStack& st)
{
	const bool innerL = visitL && nd.w3.z == 0;
	const bool innerR = visitR && nd.w3.w == 0;
	const bool leafL = visitL && !innerL;
	const bool leafR = visitR && !innerR;
	const uint32_t lFirst = child_index(nd.w3.x), rFirst = child_index(nd.w3.y);
	const uint32_t lEnd = lFirst + uint32_t(nd.w3.z), rEnd = rFirst + uint32_t(nd.w3.w);
	s.triCur = leafL ? lFirst : leafR ? rFirst : s.triCur;
	s.triEnd = leafL ? lEnd : leafR ? rEnd : s.triEnd;
	const bool both = leafL && leafR;
	s.pendCur = both ? rFirst : s.pendCur;
	s.pendEnd = both ? rEnd : s.pendEnd;
	st.push_if(innerL && innerR, lFirst);
	const bool doPop = !innerL && !innerR && st.sp > 0;
	const uint32_t popped = st.pop_if(doPop);
	s.next = innerR ? int32_t(rFirst) : innerL ? int32_t(lFirst) : doPop ? int32_t(popped) : -1;
}

// What the wave does after a step: go on, stop (no ray left), or trace its last few rays solo.
constexpr int kStepGo = 0, kStepDone = 1, kStepSolo = 2;

// Rays left in a wave, once the ray queue is empty, at or below which the wave goes solo: it traces
// them one at a time with all 64 lanes instead of one lane each, keeping one and offering the
// others to the block's other waves (below). While waves of the block are idle, the threshold is
// one more than their number, up to kMaxSoloRays, so each idle wave can take a ray at once.
constexpr uint32_t kSoloRays = 6;
constexpr uint32_t kMaxSoloRays = 16;
__device__ __forceinline__ uint32_t solo_threshold(const uint32_t* handoff);

// One iteration: consume record A, issue the load of the lane's next record into B, test A's
// triangles, retire finished rays and refill lanes. Every path a lone lane can take is kept short
// -- once few rays are left, a wave's time per step is the number of instructions it issues --
// so work that only some lanes need sits in the branch of those lanes. kExitCheck: whether this
// step tests if the wave is done or should go solo (kStepGo otherwise); the loop runs two steps
// per turn and checks after the second.
template <bool kAnyHit, bool kCount, bool kTreelet, bool kExitCheck, bool kClock = false>
__device__ __forceinline__ int pipe_step(const PipeArgs& P, Tracer& L, LaneStack& st,
                                          WaveQueue& q, const Rec& A, Rec& B, Counts& k)
{
	// Wait for A -- and for the stores of the previous step -- here, before anything else. The
	// compiler's own waits treat the vector-memory counter as unordered while a store is pending
	// among loads and then wait for everything, B's loads included; with nothing outstanding at
	// the top of the step, the only loads in flight during the triangle tests are B's.
	__builtin_amdgcn_s_waitcnt(kWaitVmcntZero);
	const uint32_t kind = L.kind;
	// The result of a ray that ended in the previous step is still in L.c until a new ray starts.
	const Closest ended = L.c;
	uint32_t triFirst = 0u, triCnt = 0u;
	bool anyHit = false;
	if (kind == kRecTri) {
		// Up to kTriBatch triangles of the current range; when the range is used up, the right
		// leaf's range queued behind it (if any) becomes current.
		triFirst = L.s.triCur;
		const uint32_t left = L.s.triEnd - triFirst;
		triCnt = left < kTriBatch ? left : kTriBatch;
		const bool rangeDone = left <= kTriBatch;
		L.s.triCur = rangeDone ? L.s.pendCur : triFirst + kTriBatch;
		L.s.triEnd = rangeDone ? L.s.pendEnd : L.s.triEnd;
		L.s.pendCur = rangeDone ? 0u : L.s.pendCur;
		L.s.pendEnd = rangeDone ? 0u : L.s.pendEnd;
		if (kAnyHit) anyHit = test_pair_any(L.r, A, triCnt);
	} else if (kind != kRecIdle) {
		if (kind == kRecRay) {
			L.r = dev::make_ray(A.w4, A.w5);
			L.fastSlab = P.finiteBounds && L.r.nanFree;
			L.c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			st.sp = 0;
			L.s = {0u, 0u, 0u, 0u, -1};
		}
		NodeWords nd;
		nd.w0 = A.w0;
		nd.w1 = A.w1;
		nd.w2 = A.w2;
		nd.w3 = make_int4(__float_as_int(A.w3.x), __float_as_int(A.w3.y), __float_as_int(A.w3.z),
		                  __float_as_int(A.w3.w));
		const float tCurrMax = kAnyHit ? L.r.tMax : dev::min_first(L.r.tMax, L.c.t);
		bool visitL, visitR;
		decide(L.r, L.fastSlab, nd, tCurrMax, visitL, visitR);
		if (kCount) {
			k.nodes++;
			k.boxHits += uint32_t(visitL) + uint32_t(visitR);
		}
		expand_pipe(nd, visitL, visitR, L.s, st);
	}

	// Next record: more of this ray, or (for lanes whose ray ends here, and idle lanes) a new ray.
	const bool more = kind != kRecIdle && !anyHit && (L.s.triCur < L.s.triEnd || L.s.next >= 0);
	uint32_t fresh = kInvalid;
	if (!q.exhausted) {
		const uint64_t freeLanes = __ballot(!more);
		if (freeLanes != 0 && (uint32_t(__popcll(freeLanes)) >= P.refillMin || freeLanes == ~0ull))
			fresh = refill(q, freeLanes, P.work, P.n, !more, P.window);
	}
	const bool tri = more && L.s.triCur < L.s.triEnd;
	const bool node = more && !tri;
	const bool ray = fresh != kInvalid;
	const uint32_t nextKind = tri ? kRecTri : node ? kRecNode : ray ? kRecRay : kRecIdle;
	// The ended ray's result is written just before the loads, after the waits of this step (the
	// stack pop's, the refill's), so its latency overlaps the triangle tests, not a wait.
	if (L.doneIdx != kInvalid) {
		store_result<kAnyHit>(P, L.doneIdx, ended);
		if (kClock) P.clock[L.doneIdx] = wall_clock64();
	}
	// Every lane issues the same six loads: a node (words 0-3), a triangle pair (words 0-5), or
	// the root node and a new ray's 32 B; idle lanes and words 4-5 of a node read the root node
	// (always cached). Offsets are in 16-B words: 4 per node, 3 per triangle.
	const uint32_t idx = tri ? L.s.triCur : node ? uint32_t(L.s.next) : 0u;
	const uint32_t word = tri ? idx + (idx << 1) : idx << 2;
	const float4* rec = (tri ? P.tris : P.nodes) + word;
	const float4* src45 = tri ? rec + 4 : ray ? P.rays + 2u * size_t(ray ? fresh : 0u) : P.nodes;
	if (kTreelet && !tri && idx < P.topCount) {
		const LdsF4* t = as_lds(P.treelet) + 4u * idx;
		B.w0 = t[0];
		B.w1 = t[1];
		B.w2 = t[2];
		B.w3 = t[3];
	} else {
		B.w0 = as_global(rec)[0];
		B.w1 = as_global(rec)[1];
		B.w2 = as_global(rec)[2];
		B.w3 = as_global(rec)[3];
	}
	B.w4 = as_global(src45)[0];
	B.w5 = as_global(src45)[1];

	if (!kAnyHit && kind == kRecTri) test_pair_closest<kCount>(L.r, A, triCnt, triFirst, L.c, k);

	if (kAnyHit && anyHit) L.c.tri = 0u;
	L.doneIdx = kind != kRecIdle && !more ? L.rayIdx : kInvalid;
	if (!more) L.rayIdx = fresh;
	L.kind = nextKind;
	if (!kExitCheck || !q.exhausted) return kStepGo;
	const uint64_t busy = __ballot(nextKind != kRecIdle);
	if (busy == 0) return kStepDone;
	return !kCount && uint32_t(__popcll(busy)) <= solo_threshold(P.handoff) ? kStepSolo : kStepGo;
}

// ---- solo mode: one ray per wave ---------------------------------------------------------------
//
// Once the ray queue is empty, a wave's last rays are the costliest ones, and a wave with one busy
// lane still issues every instruction of a step. A wave with at most kSoloRays rays left therefore
// traces them one after the other with all 64 lanes: a node's twelve slab planes are computed by
// lanes 0-11 at once and combined across lanes (DPP), the triangles of visited leaves are tested
// one per lane in a single pass, and the traversal state -- the ray, the closest hit, the stack
// (entry e in lane e of two registers) -- is uniform, so every decision is a scalar branch.
// Visiting order and arithmetic are castRay's: the same slab operations with std::min/max
// semantics, the same triangle test, and a leaf pass accepts candidates in castRay's order
// (ascending lane = ascending position in the left-then-right leaf sequence), each against the
// closest hit so far.

__device__ __forceinline__ uint32_t lane_u(uint32_t v, uint32_t l)
{
	return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}
__device__ __forceinline__ float lane_f(float v, uint32_t l)
{
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(l)));
}
// The value of lane (this lane + k) of the same 16-lane row (DPP row_shl:k).
template <int k>
__device__ __forceinline__ float from_lane_plus(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 | k, 0xF, 0xF, false));
}

struct SoloHit {
	uint32_t tri;
	float t, u, v;
	bool any;  // any hit found (any-hit traversal)
};

// Tests the n1 triangles from a1 and then the n2 from a2 -- castRay's order -- one per lane, 64 per
// pass. Closest hit: each candidate (hit, t < closest t on entry, tMin <= t <= tMax) is accepted
// in order if t is still below the closest t, exactly as the sequential loop does. Any hit:
// records a hit with tMin <= t < tMax.
template <bool kAnyHit>
__device__ __forceinline__ void solo_tris(const float4* tris, const dev::Ray& r, uint32_t a1,
                                          uint32_t n1, uint32_t a2, uint32_t n2, SoloHit& c)
{
	const uint32_t lane = __lane_id();
	for (uint32_t base = 0; base < n1 + n2; base += 64u) {
		const uint32_t j = base + lane;
		const bool have = j < n1 + n2;
		const uint32_t tri = j < n1 ? a1 + j : a2 + (j - n1);
		float t = 0.0f, u = 0.0f, v = 0.0f;
		bool hit = false;
		if (have) {
			const float4* p = tris + 3u * tri;
			const float4 p0 = as_global(p)[0], e1 = as_global(p)[1], e2 = as_global(p)[2];
			const dev::TriDet d = dev::tri_det(r, e1, e2);
			float f = dev::recip_newton(d.a);
			if (!dev::recip_newton_ok(d.a)) f = 1.0f / d.a;
			hit = dev::tri_finish(r, d, f, p0, e1, e2, t, u, v);
		}
		if (kAnyHit) {
			if (__ballot(have && hit && r.tMin <= t && t < r.tMax) != 0) {
				c.any = true;
				return;
			}
		} else {
			uint64_t cand = __ballot(have && hit && t < c.t && r.tMin <= t && t <= r.tMax);
			while (cand != 0) {
				const uint32_t w = uint32_t(__builtin_ctzll(cand));
				cand &= cand - 1u;
				const float tw = lane_f(t, w);
				if (tw < c.t) {
					c.t = tw;
					c.u = lane_f(u, w);
					c.v = lane_f(v, w);
					c.tri = lane_u(tri, w);
				}
			}
		}
	}
}

// The scene and output of a solo ray (uniform).
struct SoloScene {
	const float4* nodes;
	const float4* tris;
	const float4* rays;
	void* out;  // uint4 hits (closest hit) or uint8_t (any hit), indexed by the ray's index
	unsigned long long* clock = nullptr;  // completion clock per ray (completion-clock launches)
};

// Where a solo ray starts: the state the pipelined loop left for it (the record in flight is
// loaded again). Uniform values.
struct SoloStart {
	float4 oMin, dMax;            // origin and tMin, direction and tMax (1/dir is recomputed)
	Closest c;                    // closest hit so far
	uint32_t triCur, triEnd, pendCur, pendEnd;  // queued triangle ranges
	int32_t next;                 // node to expand next, or -1
	int sp;                       // stack entries
	uint32_t rayIdx;
	uint32_t kind;                // kRecRay: the ray is not set up yet (load it from rays[rayIdx])
	const uint32_t* ldsCol;       // the owner lane's LDS stack column (entry e at ldsCol[64 e])
	const uint32_t* spillCol;     // its spill column (entry e >= depth at [(e - depth) stride])
	uint32_t stride;
	int depth;
};

// The start state of lane lsel of this wave.
__device__ __forceinline__ SoloStart solo_start_lane(const Tracer& L, const LaneStack& st,
                                                     uint32_t lsel)
{
	SoloStart S;
	S.oMin = make_float4(lane_f(L.r.ox, lsel), lane_f(L.r.oy, lsel), lane_f(L.r.oz, lsel),
	                     lane_f(L.r.tMin, lsel));
	S.dMax = make_float4(lane_f(L.r.dx, lsel), lane_f(L.r.dy, lsel), lane_f(L.r.dz, lsel),
	                     lane_f(L.r.tMax, lsel));
	S.c = {lane_u(L.c.tri, lsel), lane_f(L.c.t, lsel), lane_f(L.c.u, lsel), lane_f(L.c.v, lsel)};
	S.triCur = lane_u(L.s.triCur, lsel);
	S.triEnd = lane_u(L.s.triEnd, lsel);
	S.pendCur = lane_u(L.s.pendCur, lsel);
	S.pendEnd = lane_u(L.s.pendEnd, lsel);
	S.next = int32_t(lane_u(uint32_t(L.s.next), lsel));
	S.sp = int(lane_u(uint32_t(st.sp), lsel));
	S.rayIdx = lane_u(L.rayIdx, lsel);
	S.kind = lane_u(L.kind, lsel);
	S.ldsCol = st.lds + lsel;
	S.spillCol = st.spill - __lane_id() + lsel;
	S.stride = st.stride;
	S.depth = st.depth;
	return S;
}

// Traces a ray to its end from S and writes its result.
template <bool kAnyHit, bool kClock = false>
__device__ __forceinline__ void solo_run(const SoloScene& P, const SoloStart& S)
{
	const uint32_t lane = __lane_id();
	const uint32_t rayIdx = S.rayIdx;
	dev::Ray r;
	SoloHit c;
	uint32_t triCur, triEnd, pendCur, pendEnd;
	int32_t cur;
	int sp;
	if (S.kind == kRecRay) {
		const float4* src = P.rays + 2u * size_t(rayIdx);
		const float4 a = as_global(src)[0], b = as_global(src)[1];
		r = dev::make_ray(make_float4(lane_f(a.x, 0), lane_f(a.y, 0), lane_f(a.z, 0), lane_f(a.w, 0)),
		                  make_float4(lane_f(b.x, 0), lane_f(b.y, 0), lane_f(b.z, 0), lane_f(b.w, 0)));
		c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX, false};
		triCur = triEnd = pendCur = pendEnd = 0u;
		cur = 0;
		sp = 0;
	} else {
		r = dev::make_ray(S.oMin, S.dMax);
		c = {S.c.tri, S.c.t, S.c.u, S.c.v, false};
		triCur = S.triCur, triEnd = S.triEnd, pendCur = S.pendCur, pendEnd = S.pendEnd;
		cur = S.next;
		sp = S.sp;
	}
	// The owner's stack (LDS column, then spill column) into two registers: entry e in lane e of
	// stk0 (e < 64) or of stk1.
	uint32_t stk0 = 0u, stk1 = 0u;
	for (int e = int(lane), h = 0; h < 2; h++, e += 64) {
		uint32_t v = 0u;
		if (e < sp)
			v = e < S.depth ? S.ldsCol[e * 64] : S.spillCol[size_t(e - S.depth) * S.stride];
		if (h == 0)
			stk0 = v;
		else
			stk1 = v;
	}
	// Slab lanes: lane i < 12 holds plane i of a node -- lc.min.xyz, lc.max.xyz, rc.min.xyz,
	// rc.max.xyz -- whose axis is i mod 3.
	const uint32_t axis = lane % 3u;
	const float oA = axis == 0u ? r.ox : axis == 1u ? r.oy : r.oz;
	const float iA = axis == 0u ? r.ix : axis == 1u ? r.iy : r.iz;
	const float* nodeWords = reinterpret_cast<const float*>(P.nodes);

	// Triangles queued when the loop handed over (a leaf pair's ranges), then the nodes.
	solo_tris<kAnyHit>(P.tris, r, triCur, triEnd - triCur, pendCur, pendEnd - pendCur, c);
	while (!(kAnyHit && c.any)) {
		if (cur < 0) {
			if (sp == 0) break;
			sp--;
			cur = int32_t(sp < 64 ? lane_u(stk0, uint32_t(sp)) : lane_u(stk1, uint32_t(sp - 64)));
		}
		const float w = as_global(nodeWords)[16u * uint32_t(cur) + (lane & 15u)];
		// t of plane i (lane i), and of plane i + 3 (the max plane of the same axis and box).
		const float t1 = (w - oA) * iA;
		const float t2 = from_lane_plus<3>(t1);
		// min/max with std::min/std::max semantics (BVHTraversal via Intersection.hpp:77-90).
		const float lo = dev::min_first(t1, t2), hi = dev::max_first(t1, t2);
		// lanes 0 (left box) and 6 (right box): combine the x, y, z slabs.
		const float tIn = dev::max_first(dev::max_first(lo, from_lane_plus<1>(lo)), from_lane_plus<2>(lo));
		const float tOut = dev::min_first(dev::min_first(hi, from_lane_plus<1>(hi)), from_lane_plus<2>(hi));
		const float tCurrMax = kAnyHit ? r.tMax : dev::min_first(r.tMax, c.t);
		const uint64_t vis = __ballot((tOut >= tIn) && tOut > r.tMin && tIn < tCurrMax);
		const bool visitL = (vis & 1ull) != 0, visitR = (vis & (1ull << 6)) != 0;
		const int32_t lRaw = int32_t(lane_u(__float_as_uint(w), 12u));
		const int32_t rRaw = int32_t(lane_u(__float_as_uint(w), 13u));
		const uint32_t lCnt = lane_u(__float_as_uint(w), 14u);
		const uint32_t rCnt = lane_u(__float_as_uint(w), 15u);
		const uint32_t lFirst = child_index(lRaw), rFirst = child_index(rRaw);
		const bool innerL = visitL && lCnt == 0u, innerR = visitR && rCnt == 0u;
		if (innerL && innerR) {
			stk0 = int(lane) == sp ? lFirst : stk0;
			stk1 = int(lane) + 64 == sp ? lFirst : stk1;
			sp++;
		}
		cur = innerR ? int32_t(rFirst) : innerL ? int32_t(lFirst) : -1;
		const bool leafL = visitL && !innerL, leafR = visitR && !innerR;
		if (leafL || leafR)
			solo_tris<kAnyHit>(P.tris, r, lFirst, leafL ? lCnt : 0u, rFirst, leafR ? rCnt : 0u, c);
	}
	if (lane == 0u) {
		const Closest res = {kAnyHit ? (c.any ? 0u : kInvalid) : c.tri, c.t, c.u, c.v};
		store_result<kAnyHit>(P.out, rayIdx, res);
		if (kClock) P.clock[rayIdx] = wall_clock64();
	}
}

// ---- handing solo rays to the block's other waves ------------------------------------------------
//
// A wave that goes solo keeps one ray and offers the others in LDS slots; waves of the block that
// have nothing left -- and the block's helper waves, which trace nothing else -- claim them and
// trace them solo, so a wave's last rays run side by side instead of one after another. The stack of an offered ray stays where it is (its owner's LDS
// column and spill column, which the owner no longer writes). Area layout (32-bit words):
// [0] slots allocated, [1] slots claimed, [2] waves that may still offer, [3] waves waiting to
// claim, then
// kSlotWords per slot. A claimed slot index is below the allocated count, and its owner marks it
// ready right after writing it, so a claimant's wait ends; the last claim attempts end once no
// wave may offer and every slot is claimed.

constexpr uint32_t kHandoffPerWave = kMaxSoloRays - 1u;  // a wave offers at most this many rays
constexpr uint32_t kSlotWords = 24;

// `laneThreads`: the block's lane-mode threads (helper waves offer nothing).
__host__ __device__ __forceinline__ uint32_t handoff_bytes(uint32_t laneThreads)
{
	return 16u + (laneThreads / 64u) * kHandoffPerWave * kSlotWords * 4u;
}

__device__ __forceinline__ uint32_t lds_load(const uint32_t* p)
{
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The solo threshold of a wave of the block (see kSoloRays).
__device__ __forceinline__ uint32_t solo_threshold(const uint32_t* H)
{
	const uint32_t idle = __builtin_amdgcn_readfirstlane(lds_load(&H[3]));
	return idle + 1u > kSoloRays ? (idle + 1u < kMaxSoloRays ? idle + 1u : kMaxSoloRays) : kSoloRays;
}

// Every thread of the block calls this before the block's first barrier: resets the counters and
// every slot's ready word (LDS holds whatever an earlier block left there, and a stale ready word
// would let a claimant read a slot before its owner writes it). `laneWaves`: the waves that run
// the lane-mode loop and may offer rays.
__device__ __forceinline__ void handoff_init(uint32_t* H, uint32_t laneWaves)
{
	if (threadIdx.x == 0) {
		H[0] = 0u;
		H[1] = 0u;
		H[2] = laneWaves;
		H[3] = 0u;
	}
	const uint32_t slots = laneWaves * kHandoffPerWave;
	for (uint32_t i = threadIdx.x; i < slots; i += blockDim.x) H[4u + i * kSlotWords + 22u] = 0u;
}

// Offers the rays of the lanes in `offer` (each lane writes its own slot), then retires this wave
// as a producer.
__device__ __forceinline__ void handoff_offer(uint32_t* H, uint64_t offer, const Tracer& L,
                                              const LaneStack& st, uint32_t batch)
{
	const uint32_t lane = __lane_id();
	const uint32_t n = uint32_t(__popcll(offer));
	if (n != 0u) {
		uint32_t base = 0u;
		if (lane == 0u) base = atomicAdd(&H[0], n);
		base = lane_u(base, 0u);
		const bool mine = ((offer >> lane) & 1ull) != 0;
		uint32_t* w = H + 4u + (base + uint32_t(__popcll(offer & __lanemask_lt()))) * kSlotWords;
		if (mine) {
			w[0] = __float_as_uint(L.r.ox);
			w[1] = __float_as_uint(L.r.oy);
			w[2] = __float_as_uint(L.r.oz);
			w[3] = __float_as_uint(L.r.tMin);
			w[4] = __float_as_uint(L.r.dx);
			w[5] = __float_as_uint(L.r.dy);
			w[6] = __float_as_uint(L.r.dz);
			w[7] = __float_as_uint(L.r.tMax);
			w[8] = L.c.tri;
			w[9] = __float_as_uint(L.c.t);
			w[10] = __float_as_uint(L.c.u);
			w[11] = __float_as_uint(L.c.v);
			w[12] = L.s.triCur;
			w[13] = L.s.triEnd;
			w[14] = L.s.pendCur;
			w[15] = L.s.pendEnd;
			w[16] = uint32_t(L.s.next);
			w[17] = uint32_t(st.sp);
			w[18] = L.rayIdx;
			w[19] = L.kind;
			w[20] = ((threadIdx.x >> 6) << 6) | lane;  // owner: wave in block, lane
			w[21] = batch;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
		if (mine) __hip_atomic_store(&w[22], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
	}
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
	if (lane == 0u) atomicSub(&H[2], 1u);
}

// Claims an offered ray: its slot, or kInvalid once none is left and no wave can offer more.
__device__ __forceinline__ uint32_t handoff_claim(uint32_t* H)
{
	uint32_t got = kInvalid;
	if (__lane_id() == 0u) {
		atomicAdd(&H[3], 1u);
		while (true) {
			const uint32_t c = lds_load(&H[1]);
			if (c < lds_load(&H[0])) {
				if (atomicCAS(&H[1], c, c + 1u) == c) {
					got = c;
					break;
				}
				continue;
			}
			if (lds_load(&H[2]) == 0u && lds_load(&H[1]) >= lds_load(&H[0])) break;
			__builtin_amdgcn_s_sleep(2);
		}
		atomicSub(&H[3], 1u);
	}
	return lane_u(got, 0u);
}

// The start state in a claimed slot (waits for its owner to mark it ready).
__device__ __forceinline__ SoloStart solo_start_slot(const uint32_t* H, uint32_t slot,
                                                     const uint32_t* ldsStacks,
                                                     const uint32_t* spill, uint32_t stride,
                                                     int depth, uint32_t& batch)
{
	const uint32_t* w = H + 4u + slot * kSlotWords;
	while (lane_u(lds_load(&w[22]), 0u) == 0u) __builtin_amdgcn_s_sleep(1);
	uint32_t v[22];
#pragma unroll
	for (int i = 0; i < 22; i++) v[i] = __builtin_amdgcn_readfirstlane(w[i]);
	SoloStart S;
	S.oMin = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
	                     __uint_as_float(v[3]));
	S.dMax = make_float4(__uint_as_float(v[4]), __uint_as_float(v[5]), __uint_as_float(v[6]),
	                     __uint_as_float(v[7]));
	S.c = {v[8], __uint_as_float(v[9]), __uint_as_float(v[10]), __uint_as_float(v[11])};
	S.triCur = v[12];
	S.triEnd = v[13];
	S.pendCur = v[14];
	S.pendEnd = v[15];
	S.next = int32_t(v[16]);
	S.sp = int(v[17]);
	S.rayIdx = v[18];
	S.kind = v[19];
	const uint32_t wave = v[20] >> 6, lane = v[20] & 63u;
	S.ldsCol = ldsStacks + wave * (uint32_t(depth) + 1u) * 64u + lane;
	// Spill columns are numbered over the grid's lane-mode threads (`stride` of them).
	S.spillCol = spill + (blockIdx.x * (stride / gridDim.x) + wave * 64u + lane);
	S.stride = stride;
	S.depth = depth;
	batch = v[21];
	return S;
}

// The block's last `helpers` waves (helper waves) run no lane-mode loop: they wait for offered rays
// and trace them solo, at raised priority so that the waves still in lane mode on their SIMD delay
// them little. A block with helpers has more waves for a solo tail than waves in lane mode.
__device__ __forceinline__ bool is_helper(uint32_t helpers)
{
	return (threadIdx.x >> 6) >= blockDim.x / 64u - helpers;
}

// Shared skeleton of the persistent closest-hit (kAnyHit = false) and any-hit kernels. Without a
// treelet (kTreelet false) every record comes from global memory and no LDS path is compiled.
// LDS stacks and spill columns belong to the lane-mode threads only.
// kClock: `counters` is instead a completion clock of n + 1 words: each ray's device clock
// (wall_clock64) when its result is written, and in word n the earliest block start.
template <bool kAnyHit, bool kCount, bool kTreelet, bool kClock = false>
__device__ __forceinline__ void persistent_trace(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint8_t* __restrict__ inLight,
	uint32_t n, uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	bool finiteBounds, uint32_t topCount, uint32_t helpers, int depth, uint32_t window,
	unsigned long long* __restrict__ counters)
{
	extern __shared__ float4 ldsRaw[];
	const uint32_t laneWaves = blockDim.x / 64u - helpers;
	const uint32_t laneThreads = laneWaves * 64u;
	uint32_t* ldsStacks = reinterpret_cast<uint32_t*>(ldsRaw);
	float4* treelet = ldsRaw + laneWaves * (uint32_t(depth) + 1u) * 64u / 4u;
	uint32_t* handoff = reinterpret_cast<uint32_t*>(treelet + 4u * topCount);
	if (kClock && threadIdx.x == 0) atomicMin(&counters[n], wall_clock64());
	if (kTreelet)
		for (uint32_t i = threadIdx.x; i < 4u * topCount; i += blockDim.x) treelet[i] = nodes[i];
	handoff_init(handoff, laneWaves);
	__syncthreads();

	const SoloScene S = {nodes, tris, rays,
	                     kAnyHit ? static_cast<void*>(inLight) : static_cast<void*>(hits),
	                     kClock ? counters : nullptr};
	const uint32_t stride = gridDim.x * laneThreads;
	if (!kCount && is_helper(helpers)) {
		__builtin_amdgcn_s_setprio(3);
		for (uint32_t slot = handoff_claim(handoff); slot != kInvalid; slot = handoff_claim(handoff)) {
			uint32_t batch;
			solo_run<kAnyHit, kClock>(S, solo_start_slot(handoff, slot, ldsStacks, spill, stride, depth,
			                                     batch));
		}
		retire_wave(work);
		return;
	}

	LaneStack st;
	st.lds = ldsStacks + (threadIdx.x >> 6) * (uint32_t(depth) + 1u) * 64u;
	st.depth = depth;
	st.stride = stride;
	st.spill = spill + (blockIdx.x * laneThreads + threadIdx.x);
	st.lane = __lane_id();
	st.sp = 0;

	const PipeArgs P = {nodes, tris, rays, hits, inLight, n, work, refillMin,
	                    kTreelet ? topCount : 0u, finiteBounds, treelet, handoff, window,
	                    kClock ? counters : nullptr};
	WaveQueue q;
	Tracer L;
	L.c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
	L.s = {0u, 0u, 0u, 0u, -1};
	L.rayIdx = kInvalid;
	L.doneIdx = kInvalid;
	L.kind = kRecIdle;
	L.fastSlab = false;
	L.r = dev::make_ray(make_float4(0.f, 0.f, 0.f, 0.f), make_float4(1.f, 1.f, 1.f, 0.f));
	const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
	Rec A = {z, z, z, z, z, z}, B = A;
	Counts k;
	int next = kStepGo;
	while (next == kStepGo) {
		pipe_step<kAnyHit, kCount, kTreelet, false, kClock>(P, L, st, q, A, B, k);
		next = pipe_step<kAnyHit, kCount, kTreelet, true, kClock>(P, L, st, q, B, A, k);
	}
	if (L.doneIdx != kInvalid) {
		store_result<kAnyHit>(P, L.doneIdx, L.c);
		if (kClock) P.clock[L.doneIdx] = wall_clock64();
	}
	if (!kCount) {
		// Solo: keep the first busy lane's ray, offer the others to the block, then trace offered
		// rays until none is left (the loads still in flight land in A, which is not read again).
		const uint64_t busy = next == kStepSolo ? __ballot(L.kind != kRecIdle) : 0ull;
		handoff_offer(handoff, busy & (busy - 1u), L, st, 0u);
		if (busy != 0)
			solo_run<kAnyHit, kClock>(S, solo_start_lane(L, st, uint32_t(__builtin_ctzll(busy))));
		for (uint32_t slot = handoff_claim(handoff); slot != kInvalid; slot = handoff_claim(handoff)) {
			uint32_t batch;
			solo_run<kAnyHit, kClock>(S, solo_start_slot(handoff, slot, ldsStacks, spill, st.stride, depth,
			                                     batch));
		}
	}
	if (kCount) add_counts(counters, k);
	retire_wave(work);
}
template <bool kCount, bool kTreelet, bool kClock = false>
__global__ __launch_bounds__(kMaxBlock) void k_closest_persistent(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t n,
	uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	uint32_t finiteBounds, uint32_t topCount, uint32_t helpers, uint32_t depth, uint32_t window,
	unsigned long long* __restrict__ counters)
{
	persistent_trace<false, kCount, kTreelet, kClock>(nodes, tris, rays, hits, nullptr, n, work,
	                                                  spill, refillMin, finiteBounds != 0,
	                                                  topCount, helpers, int(depth), window,
	                                                  counters);
}

template <bool kTreelet>
__global__ __launch_bounds__(kMaxBlock) void k_anyhit_persistent(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint8_t* __restrict__ inLight, uint32_t n,
	uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	uint32_t finiteBounds, uint32_t topCount, uint32_t helpers, uint32_t depth, uint32_t window)
{
	persistent_trace<true, false, kTreelet>(nodes, tris, rays, nullptr, inLight, n, work, spill,
	                                        refillMin, finiteBounds != 0, topCount, helpers,
	                                        int(depth), window, nullptr);
}

// ---- fused multi-batch launch ---------------------------------------------------------------------
//
// One persistent launch over many ray batches (closest-hit or any-hit, on one or several scenes):
// the batches' rays form one index space [0, total) pulled from the shared counter as above; a
// lane maps its ray index to (batch, local index) by binary search over the batch offsets held in
// LDS. The treelet, if any, belongs to one scene (`treeletNodes`); lanes on other scenes read
// every node from global memory. With completion tracking on, the lanes completing rays in the
// block count completions and the latest completion time per batch in LDS, and every block adds
// them to the batch table with one global atomic per batch when it retires, so each batch's
// latency inside the launch is known (the clock only grows, so completion times need no reset
// between launches). The last wave to retire resets the counters for the next launch.

constexpr uint32_t kMaxBatches = 1024;

__device__ __forceinline__ uint32_t find_batch(const uint32_t* offsets, uint32_t numBatches,
                                               uint32_t g)
{
	// largest b with offsets[b] <= g (offsets[0] = 0, offsets[numBatches] = total > g)
	uint32_t lo = 0, hi = numBatches;
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (offsets[mid] <= g)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// The batch of each new ray of a refill, without a search per lane: a wave's windows come in
// increasing order, so the wave keeps the batch of its current window's first ray (`cur`) and of the
// window before it (`prev`, whose last rays a refill may still hand out). A lane starts from its
// window's batch and steps over the batch ends below its index -- usually none or one, since a
// window (at most 4096 rays) is mostly shorter than a batch; after four steps it searches.
struct BatchCursor {
	uint32_t start = 0;  // first ray of the window
	uint32_t batch = 0;  // its batch
};

__device__ __forceinline__ uint32_t batch_from(const uint32_t* offsets, uint32_t numBatches,
                                               uint32_t from, uint32_t g)
{
	uint32_t b = from;
	for (int k = 0; k < 4; k++) {
		if (offsets[b + 1u] > g) return b;
		b++;
	}
	return find_batch(offsets, numBatches, g);
}

// Launch-wide arguments of the fused kernel.
struct MultiArgs {
	const BatchDesc* descs;
	const uint32_t* offsets;  // LDS copy: offsets[b] = first global ray index of batch b
	uint32_t numBatches;
	uint32_t total;
	uint32_t* work;
	uint32_t refillMin;
	uint32_t topCount;
	const float4* treeletNodes;  // the scene whose top nodes are in the LDS treelet
	const float4* treelet;       // LDS
	unsigned long long* blockFinish;  // LDS, per batch (completion tracking)
	uint32_t* blockDone;              // LDS, per batch
	const uint32_t* handoff;          // the block's handoff area (LDS)
	uint32_t window;                  // rays per fetch
};

// Per-lane state of the fused kernel: the pipelined loop's, plus the scene, output and kind of
// the lane's batch.
struct MultiTracer {
	Tracer t;
	const float4* nodes;
	const float4* tris;
	void* out;
	uint32_t batch;
	bool anyHit;
	bool finiteBounds;
	bool inTreelet;
};

__device__ __forceinline__ void track_done(const MultiArgs& P, uint32_t batch)
{
	// Block-local LDS atomics; flushed to the batch table when the block retires.
	atomicAdd(&P.blockDone[batch], 1u);
	atomicMax(&P.blockFinish[batch], wall_clock64());
}

// pipe_step for the fused kernel: the same iteration, with the scene, output and kind taken from
// each lane's batch (looked up in the batch table when the lane takes a new ray). A lane's result
// is written with its own batch's output before a new ray's batch replaces it.
template <bool kTreelet, bool kTrack, bool kExitCheck>
__device__ __forceinline__ int pipe_step_multi(const MultiArgs& P, MultiTracer& M, LaneStack& st,
                                               WaveQueue& q, BatchCursor& cur, BatchCursor& prev,
                                               const Rec& A, Rec& B)
{
	__builtin_amdgcn_s_waitcnt(kWaitVmcntZero);
	Tracer& L = M.t;
	Counts none;
	const uint32_t kind = L.kind;
	uint32_t triFirst = 0u, triCnt = 0u;
	bool anyHit = false;
	if (kind == kRecTri) {
		triFirst = L.s.triCur;
		const uint32_t left = L.s.triEnd - triFirst;
		triCnt = left < kTriBatch ? left : kTriBatch;
		const bool rangeDone = left <= kTriBatch;
		L.s.triCur = rangeDone ? L.s.pendCur : triFirst + kTriBatch;
		L.s.triEnd = rangeDone ? L.s.pendEnd : L.s.triEnd;
		L.s.pendCur = rangeDone ? 0u : L.s.pendCur;
		L.s.pendEnd = rangeDone ? 0u : L.s.pendEnd;
		if (M.anyHit) anyHit = test_pair_any(L.r, A, triCnt);
	} else if (kind != kRecIdle) {
		if (kind == kRecRay) {
			L.r = dev::make_ray(A.w4, A.w5);
			L.fastSlab = M.finiteBounds && L.r.nanFree;
			L.c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			st.sp = 0;
			L.s = {0u, 0u, 0u, 0u, -1};
		}
		NodeWords nd;
		nd.w0 = A.w0;
		nd.w1 = A.w1;
		nd.w2 = A.w2;
		nd.w3 = make_int4(__float_as_int(A.w3.x), __float_as_int(A.w3.y), __float_as_int(A.w3.z),
		                  __float_as_int(A.w3.w));
		const float tCurrMax = M.anyHit ? L.r.tMax : dev::min_first(L.r.tMax, L.c.t);
		bool visitL, visitR;
		decide(L.r, L.fastSlab, nd, tCurrMax, visitL, visitR);
		expand_pipe(nd, visitL, visitR, L.s, st);
	}

	const bool more = kind != kRecIdle && !anyHit && (L.s.triCur < L.s.triEnd || L.s.next >= 0);
	uint32_t fresh = kInvalid;
	if (!q.exhausted) {
		const uint64_t freeLanes = __ballot(!more);
		if (freeLanes != 0 && (uint32_t(__popcll(freeLanes)) >= P.refillMin || freeLanes == ~0ull)) {
			const uint32_t start = q.start;
			fresh = refill(q, freeLanes, P.work, P.total, !more, P.window);
			if (q.start != start && !q.exhausted) {
				// a new window (uniform): its batch, searched from the current one onwards
				prev = cur;
				cur.start = q.start;
				cur.batch = batch_from(P.offsets, P.numBatches, cur.batch, cur.start);
			}
		}
	}
	const bool tri = more && L.s.triCur < L.s.triEnd;
	const bool node = more && !tri;
	const bool ray = fresh != kInvalid;
	// A new ray's batch: the largest b with offsets[b] <= its global index.
	uint32_t nb = 0u, local = 0u;
	BatchDesc nd{};
	if (ray) {
		nb = batch_from(P.offsets, P.numBatches, fresh >= cur.start ? cur.batch : prev.batch, fresh);
		local = fresh - P.offsets[nb];
		nd = P.descs[nb];
	}
	const uint32_t nextKind = tri ? kRecTri : node ? kRecNode : ray ? kRecRay : kRecIdle;
	// The same six loads for every lane: a node, a triangle pair, or a new ray's root node and
	// 32 B; idle lanes and words 4-5 of a node read their scene's root.
	const float4* nodes = ray ? nd.nodes : M.nodes;
	const uint32_t idx = tri ? L.s.triCur : node ? uint32_t(L.s.next) : 0u;
	const uint32_t word = tri ? idx + (idx << 1) : idx << 2;
	const float4* rec = (tri ? M.tris : nodes) + word;
	const float4* src45 = tri ? rec + 4 : ray ? nd.rays + 2u * size_t(local) : nodes;
	const bool inTreelet = ray ? nd.nodes == P.treeletNodes : M.inTreelet;
	if (kTreelet && !tri && inTreelet && idx < P.topCount) {
		const LdsF4* t = as_lds(P.treelet) + 4u * idx;
		B.w0 = t[0];
		B.w1 = t[1];
		B.w2 = t[2];
		B.w3 = t[3];
	} else {
		B.w0 = as_global(rec)[0];
		B.w1 = as_global(rec)[1];
		B.w2 = as_global(rec)[2];
		B.w3 = as_global(rec)[3];
	}
	B.w4 = as_global(src45)[0];
	B.w5 = as_global(src45)[1];

	if (kind == kRecTri && !M.anyHit) test_pair_closest<false>(L.r, A, triCnt, triFirst, L.c, none);

	if (kind != kRecIdle && !more) {
		if (M.anyHit)
			static_cast<uint8_t*>(M.out)[L.rayIdx] = anyHit ? 0 : 1;
		else
			static_cast<uint4*>(M.out)[L.rayIdx] = make_uint4(
				L.c.tri, __float_as_uint(L.c.t), __float_as_uint(L.c.u), __float_as_uint(L.c.v));
		if (kTrack) track_done(P, M.batch);
	}
	if (ray) {
		M.nodes = nd.nodes;
		M.tris = nd.tris;
		M.out = nd.out;
		M.batch = nb;
		M.anyHit = nd.kind != 0u;
		M.finiteBounds = nd.finiteBounds != 0u;
		M.inTreelet = inTreelet;
	}
	if (!more) L.rayIdx = ray ? local : kInvalid;
	L.kind = nextKind;
	if (!kExitCheck || !q.exhausted) return kStepGo;
	const uint64_t busy = __ballot(nextKind != kRecIdle);
	if (busy == 0) return kStepDone;
	return uint32_t(__popcll(busy)) <= solo_threshold(P.handoff) ? kStepSolo : kStepGo;
}

template <bool kTreelet, bool kTrack>
__device__ __forceinline__ void trace_batches(const BatchDesc* __restrict__ descs,
                                              const uint32_t* __restrict__ offsetsGlobal,
                                              uint32_t numBatches, uint32_t total,
                                              uint32_t* __restrict__ work,
                                              uint32_t* __restrict__ spill, uint32_t refillMin,
                                              const float4* __restrict__ treeletNodes,
                                              uint32_t topCount, uint32_t helpers, int depth,
                                              uint32_t window, BatchClock* __restrict__ clock)
{
	extern __shared__ float4 ldsRaw[];
	const uint32_t laneWaves = blockDim.x / 64u - helpers;
	const uint32_t laneThreads = laneWaves * 64u;
	uint32_t* ldsStacks = reinterpret_cast<uint32_t*>(ldsRaw);
	float4* treelet = ldsRaw + laneWaves * (uint32_t(depth) + 1u) * 64u / 4u;
	uint32_t* handoff = reinterpret_cast<uint32_t*>(treelet + 4u * topCount);
	unsigned long long* blockFinish =
		reinterpret_cast<unsigned long long*>(handoff + handoff_bytes(laneThreads) / 4u);
	uint32_t* blockDone = reinterpret_cast<uint32_t*>(blockFinish + (kTrack ? numBatches : 0u));
	uint32_t* offsets = blockDone + (kTrack ? numBatches : 0u);
	if (kTrack && threadIdx.x == 0) atomicMin(&clock->start, wall_clock64());
	for (uint32_t i = threadIdx.x; i < 4u * topCount; i += blockDim.x) treelet[i] = treeletNodes[i];
	for (uint32_t i = threadIdx.x; i <= numBatches; i += blockDim.x) offsets[i] = offsetsGlobal[i];
	if (kTrack)
		for (uint32_t i = threadIdx.x; i < numBatches; i += blockDim.x) {
			blockFinish[i] = 0ull;
			blockDone[i] = 0u;
		}
	handoff_init(handoff, laneWaves);
	__syncthreads();

	const MultiArgs P = {descs, offsets, numBatches, total, work, refillMin,
	                     kTreelet ? topCount : 0u, treeletNodes, treelet, blockFinish, blockDone,
	                     handoff, window};
	const uint32_t stride = gridDim.x * laneThreads;
	// Traces the ray of a claimed slot solo.
	auto runSlot = [&](uint32_t slot) {
		uint32_t b;
		const SoloStart S0 = solo_start_slot(handoff, slot, ldsStacks, spill, stride, depth, b);
		const BatchDesc d = descs[b];
		const SoloScene S = {d.nodes, d.tris, d.rays, d.out};
		if (d.kind != 0u)
			solo_run<true>(S, S0);
		else
			solo_run<false>(S, S0);
		if (kTrack && __lane_id() == 0u) track_done(P, b);
	};
	const bool helper = is_helper(helpers);
	if (helper) {
		__builtin_amdgcn_s_setprio(3);
		for (uint32_t slot = handoff_claim(handoff); slot != kInvalid; slot = handoff_claim(handoff))
			runSlot(slot);
	}

	LaneStack st;
	st.lds = ldsStacks + (threadIdx.x >> 6) * (uint32_t(depth) + 1u) * 64u;
	st.depth = depth;
	st.stride = stride;
	st.spill = spill + (blockIdx.x * laneThreads + threadIdx.x);
	st.lane = __lane_id();
	st.sp = 0;
	WaveQueue q;
	BatchCursor cur, prev;
	MultiTracer M;
	Tracer& L = M.t;
	L.c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
	L.s = {0u, 0u, 0u, 0u, -1};
	L.rayIdx = kInvalid;
	L.doneIdx = kInvalid;
	L.kind = kRecIdle;
	L.fastSlab = false;
	L.r = dev::make_ray(make_float4(0.f, 0.f, 0.f, 0.f), make_float4(1.f, 1.f, 1.f, 0.f));
	// Until a lane takes its first ray, it reads batch 0's root node.
	M.nodes = descs[0].nodes;
	M.tris = descs[0].tris;
	M.out = descs[0].out;
	M.batch = 0u;
	M.anyHit = false;
	M.finiteBounds = false;
	M.inTreelet = false;
	const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
	Rec A = {z, z, z, z, z, z}, B = A;
	int next = helper ? kStepDone : kStepGo;
	while (next == kStepGo) {
		pipe_step_multi<kTreelet, kTrack, false>(P, M, st, q, cur, prev, A, B);
		next = pipe_step_multi<kTreelet, kTrack, true>(P, M, st, q, cur, prev, B, A);
	}
	if (!helper) {
		// Solo: keep the first busy lane's ray, offer the others to the block, then trace offered
		// rays until none is left.
		const uint64_t busy = next == kStepSolo ? __ballot(L.kind != kRecIdle) : 0ull;
		handoff_offer(handoff, busy & (busy - 1u), L, st, M.batch);
		if (busy != 0) {
			const uint32_t lsel = uint32_t(__builtin_ctzll(busy));
			const uint32_t b = lane_u(M.batch, lsel);
			const SoloStart S0 = solo_start_lane(L, st, lsel);
			const BatchDesc d = descs[b];
			const SoloScene S = {d.nodes, d.tris, d.rays, d.out};
			if (d.kind != 0u)
				solo_run<true>(S, S0);
			else
				solo_run<false>(S, S0);
			if (kTrack && __lane_id() == 0u) track_done(P, b);
		}
		for (uint32_t slot = handoff_claim(handoff); slot != kInvalid; slot = handoff_claim(handoff))
			runSlot(slot);
	}
	if (kTrack) {
		__syncthreads();
		for (uint32_t b = threadIdx.x; b < numBatches; b += blockDim.x)
			if (blockDone[b]) {
				atomicAdd(&clock->done[b], blockDone[b]);
				atomicMax(&clock->finish[b], blockFinish[b]);
			}
	}
	// The last wave resets the shared counters (and, with tracking, the batch counters).
	if (__lane_id() == 0) {
		__threadfence();
		const uint32_t totalWaves = gridDim.x * (blockDim.x / 64u);
		const bool last = atomicAdd(&work[1], 1u) == totalWaves - 1u;
		if (last) {
			atomicExch(&work[0], 0u);
			atomicExch(&work[1], 0u);
			if (kTrack) {
				for (uint32_t b = 0; b < numBatches; b++) atomicExch(&clock->done[b], 0u);
				clock->lastStart = clock->start;
				atomicExch(&clock->start, ~0ull);
			}
		}
	}
}

template <bool kTreelet, bool kTrack>
__global__ __launch_bounds__(kMaxBlock) void k_trace_batches(
	const BatchDesc* __restrict__ descs, const uint32_t* __restrict__ offsetsGlobal,
	uint32_t numBatches, uint32_t total, uint32_t* __restrict__ work,
	uint32_t* __restrict__ spill, uint32_t refillMin, const float4* __restrict__ treeletNodes,
	uint32_t topCount, uint32_t helpers, uint32_t depth, uint32_t window,
	BatchClock* __restrict__ clock)
{
	trace_batches<kTreelet, kTrack>(descs, offsetsGlobal, numBatches, total, work, spill, refillMin,
	                                treeletNodes, topCount, helpers, int(depth), window, clock);
}

// ---- device self-check of the fast reciprocal -----------------------------------------------------

constexpr uint64_t kBinade = 1ull << 23;           // significands per exponent
constexpr uint64_t kNormalBinades = 252;            // exponent fields 1 .. 252: 2^-126 <= |a| < 2^126
constexpr uint32_t kSampleExponents[4] = {1u, 126u, 127u, 252u};

__global__ __launch_bounds__(kBlock) void k_check_reciprocal(bool exhaustive, uint64_t total,
                                                             unsigned long long* mismatches)
{
	const uint64_t perSign = total / 2u;
	unsigned long long bad = 0;
	for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < total;
	     i += uint64_t(gridDim.x) * blockDim.x) {
		const uint32_t sign = i >= perSign ? 0x80000000u : 0u;
		const uint64_t j = i % perSign;
		const uint32_t binade = uint32_t(j / kBinade), mant = uint32_t(j % kBinade);
		const uint32_t expField = exhaustive ? 1u + binade : kSampleExponents[binade];
		const float a = __uint_as_float(sign | (expField << 23) | mant);
		if (__float_as_uint(dev::recip_newton(a)) != __float_as_uint(1.0f / a)) bad++;
	}
	if (bad) atomicAdd(mismatches, bad);
}

uint32_t gridFor(uint32_t n, uint32_t block) { return n / block + (n % block != 0u); }

// Persistent launch shape: grid capped by the work and by the spill area's lane count (lane-mode
// threads only); the block is the lane-mode threads plus the helper waves.
void persistentShape(const TraceLaunch& L, const SceneView& s, uint32_t n, uint32_t& grid,
                     uint32_t& block, uint32_t& top, uint32_t& ldsBytes)
{
	block = L.blockThreads + 64u * L.helperWaves;
	top = L.topNodes < s.topCount ? L.topNodes : s.topCount;
	grid = L.gridBlocks < gridFor(n, L.blockThreads) ? L.gridBlocks : gridFor(n, L.blockThreads);
	const uint32_t maxGrid = L.spillLanes / L.blockThreads;
	if (grid > maxGrid) grid = maxGrid;
	if (grid == 0) grid = 1;
	ldsBytes = persistent_lds_bytes(L.blockThreads, top, L.ldsDepth);
}

}  // namespace

// Device scratch of the allocation-free closest-hit cast (phx_cast_closest_noalloc): the ray
// counter pair of its persistent launch in the first kNoAllocCounterBytes, then the stack spill
// area, kNoAllocSpillWords words. A file-scope device array: every GPU the library runs on holds
// its own copy from the module load until the process ends. Host code must not pass `scratch`
// itself to a kernel (on the host it is only a stand-in for the symbol); it uses the address
// hipGetSymbolAddress returns for the current device, looked up once per device.
static __device__ __align__(256) uint8_t
	scratch[kNoAllocCounterBytes + size_t(kNoAllocSpillWords) * 4u];
static_assert(kPersistentStack == int(kPersistentStackEntries), "stack entries per lane");

namespace {

std::mutex gScratchMu;
void* gScratchPtr[kNoAllocMaxDevices] = {};

// Zeroes `n` words (n <= 64); a kernel rather than a memset, so a caller may capture it in a graph.
__global__ void k_zero_words(uint32_t* p, uint32_t n)
{
	if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

}  // namespace

hipError_t noAllocScratchInit(int device)
{
	if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> g(gScratchMu);
	if (gScratchPtr[device]) return hipSuccess;
	void* ptr = nullptr;
	const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
	if (e != hipSuccess) return e;
	gScratchPtr[device] = ptr;
	return hipSuccess;
}

void* noAllocScratch(int device)
{
	if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
	std::lock_guard<std::mutex> g(gScratchMu);
	return gScratchPtr[device];
}

hipError_t launchZeroWords(uint32_t* p, uint32_t n, hipStream_t stream)
{
	hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, stream, p, n);
	return hipGetLastError();
}

uint32_t persistentSpillWords(uint32_t lanes)
{
	return uint32_t(kPersistentStack - kMinLdsDepth) * lanes;
}

bool validLdsDepth(uint32_t depth)
{
	return depth >= uint32_t(kMinLdsDepth) && depth <= uint32_t(kMaxLdsDepth);
}

uint32_t persistentLdsBytes(uint32_t blockThreads, uint32_t topNodes, uint32_t ldsDepth)
{
	return persistent_lds_bytes(blockThreads, topNodes, ldsDepth);
}

hipError_t persistentBlocksPerCU(uint32_t blockThreads, uint32_t topNodes, uint32_t helperWaves,
                                 uint32_t ldsDepth, uint32_t* blocks)
{
	int a = 0;
	const void* fn = topNodes ? reinterpret_cast<const void*>(k_closest_persistent<false, true>)
	                          : reinterpret_cast<const void*>(k_closest_persistent<false, false>);
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
		&a, fn, int(blockThreads + 64u * helperWaves),
		persistent_lds_bytes(blockThreads, topNodes, ldsDepth));
	*blocks = uint32_t(a > 0 ? a : 0);
	return e;
}

hipError_t launchCheckReciprocal(bool exhaustive, unsigned long long* mismatches, uint64_t* tested,
                                 hipStream_t stream)
{
	const uint64_t total = 2u * kBinade * (exhaustive ? kNormalBinades : 4u);
	*tested = total;
	hipLaunchKernelGGL(k_check_reciprocal, dim3(2048), dim3(kBlock), 0, stream, exhaustive, total,
	                   mismatches);
	return hipGetLastError();
}

hipError_t launchClosest(const SceneView& s, const PhxRayIn* rays, uint32_t n, PhxRayHit* hits,
                         const TraceLaunch& L, unsigned long long* counters, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const float4* nodes = reinterpret_cast<const float4*>(s.nodes);
	const float4* tris = reinterpret_cast<const float4*>(s.trisPre);
	const float4* r4 = reinterpret_cast<const float4*>(rays);
	uint4* h4 = reinterpret_cast<uint4*>(hits);
	if (L.kernel == TraceKernel::Packet && !counters) {
		hipLaunchKernelGGL(k_packet<false>, dim3(gridFor(n, kPacketBlock)), dim3(kPacketBlock), 0,
		                   stream, nodes, tris, r4, static_cast<void*>(h4), n, s.finiteBounds);
		return hipGetLastError();
	}
	if (L.kernel != TraceKernel::Persistent || n > 0x7FFFFFFFu) {
		if (counters)
			hipLaunchKernelGGL(k_closest_simple<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0,
			                   stream, nodes, tris, r4, h4, n, counters);
		else
			hipLaunchKernelGGL(k_closest_simple<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0,
			                   stream, nodes, tris, r4, h4, n, counters);
		return hipGetLastError();
	}
	uint32_t grid, block, top, lds;
	TraceLaunch shape = L;
	if (counters) shape.helperWaves = 0;  // the counting variant traces every ray in lane mode
	persistentShape(shape, s, n, grid, block, top, lds);
	auto kernel = counters ? (top ? k_closest_persistent<true, true> : k_closest_persistent<true, false>)
	                       : (top ? k_closest_persistent<false, true> : k_closest_persistent<false, false>);
	hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, nodes, tris, r4, h4, n,
	                   L.work, L.spill, L.refillMin, s.finiteBounds, top, shape.helperWaves,
	                   L.ldsDepth, L.window, counters);
	return hipGetLastError();
}

hipError_t launchClosestClock(const SceneView& s, const PhxRayIn* rays, uint32_t n, PhxRayHit* hits,
                              const TraceLaunch& L, unsigned long long* clock, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	if (n > 0x7FFFFFFFu) return hipErrorInvalidValue;
	uint32_t grid, block, top, lds;
	persistentShape(L, s, n, grid, block, top, lds);
	auto kernel = top ? k_closest_persistent<false, true, true> : k_closest_persistent<false, false, true>;
	hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream,
	                   reinterpret_cast<const float4*>(s.nodes),
	                   reinterpret_cast<const float4*>(s.trisPre),
	                   reinterpret_cast<const float4*>(rays), reinterpret_cast<uint4*>(hits), n,
	                   L.work, L.spill, L.refillMin, s.finiteBounds, top, L.helperWaves, L.ldsDepth,
	                   L.window, clock);
	return hipGetLastError();
}

hipError_t launchAnyHitCounted(const SceneView& s, const PhxRayIn* rays, uint32_t n,
                               uint8_t* inLight, unsigned long long* counters, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	hipLaunchKernelGGL(k_anyhit_simple<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, stream,
	                   reinterpret_cast<const float4*>(s.nodes),
	                   reinterpret_cast<const float4*>(s.trisPre),
	                   reinterpret_cast<const float4*>(rays), inLight, n, counters);
	return hipGetLastError();
}

hipError_t launchInstances(const PhxBVHNode* top, const DevInstance* inst, uint32_t numInst,
                          const PhxRayIn* rays, uint32_t n, PhxRayHit* hits, uint32_t* instOut,
                          uint8_t* inLight, bool seeded, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const float4* t4 = reinterpret_cast<const float4*>(top);
	const float4* r4 = reinterpret_cast<const float4*>(rays);
	// Rays per wave: enough waves to fill the device (about 5,120: five per SIMD, the most the
	// kernel's registers allow), at least one ray per lane.
	const uint32_t perWave = std::max<uint32_t>(64u, std::min<uint32_t>(1024u, (n + 5119u) / 5120u));
	const uint32_t waves = (n + perWave - 1u) / perWave;
	const dim3 grid((waves + kBlock / 64u - 1u) / (kBlock / 64u));
	if (inLight)
		hipLaunchKernelGGL(k_trace_instances<true>, grid, dim3(kBlock), 0, stream, t4, inst,
		                   numInst, r4, nullptr, nullptr, inLight, n, uint32_t(seeded), perWave);
	else
		hipLaunchKernelGGL(k_trace_instances<false>, grid, dim3(kBlock), 0, stream, t4, inst,
		                   numInst, r4, reinterpret_cast<uint4*>(hits), instOut, nullptr, n,
		                   uint32_t(seeded), perWave);
	return hipGetLastError();
}

hipError_t launchAnyHit(const SceneView& s, const PhxRayIn* rays, uint32_t n, uint8_t* inLight,
                        const TraceLaunch& L, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const float4* nodes = reinterpret_cast<const float4*>(s.nodes);
	const float4* tris = reinterpret_cast<const float4*>(s.trisPre);
	const float4* r4 = reinterpret_cast<const float4*>(rays);
	if (L.kernel == TraceKernel::Packet) {
		hipLaunchKernelGGL(k_packet<true>, dim3(gridFor(n, kPacketBlock)), dim3(kPacketBlock), 0,
		                   stream, nodes, tris, r4, static_cast<void*>(inLight), n, s.finiteBounds);
		return hipGetLastError();
	}
	if (L.kernel == TraceKernel::Simple || n > 0x7FFFFFFFu) {
		hipLaunchKernelGGL(k_anyhit_simple<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0,
		                   stream, nodes, tris, r4, inLight, n, nullptr);
		return hipGetLastError();
	}
	uint32_t grid, block, top, lds;
	persistentShape(L, s, n, grid, block, top, lds);
	hipLaunchKernelGGL(top ? k_anyhit_persistent<true> : k_anyhit_persistent<false>, dim3(grid),
	                   dim3(block), lds, stream, nodes, tris, r4, inLight, n, L.work, L.spill,
	                   L.refillMin, s.finiteBounds, top, L.helperWaves, L.ldsDepth, L.window);
	return hipGetLastError();
}

uint32_t batchLdsBytes(uint32_t blockThreads, uint32_t topNodes, uint32_t numBatches,
                       bool tracked, uint32_t ldsDepth)
{
	return persistent_lds_bytes(blockThreads, topNodes, ldsDepth) + (numBatches + 1u) * 4u +
	       (tracked ? numBatches * 12u : 0u);
}

hipError_t batchBlocksPerCU(uint32_t blockThreads, uint32_t topNodes, uint32_t helperWaves,
                            uint32_t ldsDepth, uint32_t numBatches, bool tracked,
                            uint32_t* blocks)
{
	int a = 0;
	const void* fn =
		tracked ? (topNodes ? reinterpret_cast<const void*>(k_trace_batches<true, true>)
		                    : reinterpret_cast<const void*>(k_trace_batches<false, true>))
		        : (topNodes ? reinterpret_cast<const void*>(k_trace_batches<true, false>)
		                    : reinterpret_cast<const void*>(k_trace_batches<false, false>));
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
		&a, fn, int(blockThreads + 64u * helperWaves),
		batchLdsBytes(blockThreads, topNodes, numBatches, tracked, ldsDepth));
	*blocks = uint32_t(a > 0 ? a : 0);
	return e;
}

hipError_t launchBatches(const BatchDesc* descs, const uint32_t* offsets, uint32_t numBatches,
                         uint32_t total, const TraceLaunch& L, const float4* treeletNodes,
                         uint32_t treeletCount, BatchClock* clock, hipStream_t stream)
{
	if (total == 0) return hipSuccess;
	const uint32_t block = L.blockThreads;  // lane-mode threads; helper waves come on top
	const uint32_t top = L.topNodes < treeletCount ? L.topNodes : treeletCount;
	uint32_t grid = L.gridBlocks < gridFor(total, block) ? L.gridBlocks : gridFor(total, block);
	if (grid > L.spillLanes / block) grid = L.spillLanes / block;
	if (grid == 0) grid = 1;
	auto kernel = clock ? (top ? k_trace_batches<true, true> : k_trace_batches<false, true>)
	                    : (top ? k_trace_batches<true, false> : k_trace_batches<false, false>);
	hipLaunchKernelGGL(kernel, dim3(grid), dim3(block + 64u * L.helperWaves),
	                   batchLdsBytes(block, top, numBatches, clock != nullptr, L.ldsDepth), stream,
	                   descs, offsets, numBatches, total, L.work, L.spill, L.refillMin, treeletNodes,
	                   top, L.helperWaves, L.ldsDepth, L.window, clock);
	return hipGetLastError();
}

}  // namespace phx
