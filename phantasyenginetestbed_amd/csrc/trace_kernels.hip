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

#include "kernels.hpp"
#include "trace_device.hpp"

namespace phx {
namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

// Address-space views: records are loaded through pointers the compiler knows to be global (or
// LDS), so the loads are global_load / ds_read rather than flat loads, whose latency is higher
// and which wait on both the vector-memory and LDS counters.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const float4 __attribute__((address_space(1))) GlobalF4;
typedef const float4 __attribute__((address_space(3))) LdsF4;
#else  // the host pass only needs the declarations to parse
typedef const float4 GlobalF4;
typedef const float4 LdsF4;
#endif
__device__ __forceinline__ GlobalF4* as_global(const float4* p) { return (GlobalF4*)(p); }
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

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_closest_simple(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t n,
	unsigned long long* __restrict__ counters)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);
	uint32_t stack[kSimpleStack];
	stack[0] = 0u;
	int sp = 1;
	Closest c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
	Counts k;
	while (sp > 0) {
		sp--;
		const NodeWords nd = load_node(nodes, stack[sp]);
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
				if (sp < kSimpleStack) stack[sp++] = child_index(nd.w3.x);
			} else {
				closest_leaf<kCount>(r, tris, child_index(nd.w3.x), nd.w3.z, c, k);
			}
		}
		if (visitR) {
			if (kCount) k.boxHits++;
			if (nd.w3.w == 0) {
				if (sp < kSimpleStack) stack[sp++] = child_index(nd.w3.y);
			} else {
				closest_leaf<kCount>(r, tris, child_index(nd.w3.y), nd.w3.w, c, k);
			}
		}
	}
	hits[i] = make_uint4(c.tri, __float_as_uint(c.t), __float_as_uint(c.u), __float_as_uint(c.v));
	if (kCount) add_counts(counters, k);
}

__global__ __launch_bounds__(kBlock) void k_anyhit_simple(const float4* __restrict__ nodes,
                                                          const float4* __restrict__ tris,
                                                          const float4* __restrict__ rays,
                                                          uint8_t* __restrict__ inLight, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const dev::Ray r = dev::make_ray(rays[2 * size_t(i)], rays[2 * size_t(i) + 1]);
	uint32_t stack[kSimpleStack];
	stack[0] = 0u;
	int sp = 1;
	bool occluded = false;
	while (sp > 0 && !occluded) {
		sp--;
		const NodeWords nd = load_node(nodes, stack[sp]);
		float inL, outL, inR, outR;
		dev::slab(r, nd.w0.x, nd.w0.y, nd.w0.z, nd.w0.w, nd.w1.x, nd.w1.y, inL, outL);
		dev::slab(r, nd.w1.z, nd.w1.w, nd.w2.x, nd.w2.y, nd.w2.z, nd.w2.w, inR, outR);
		const bool visitL = (outL >= inL) && outL > r.tMin && inL < r.tMax;
		const bool visitR = (outR >= inR) && outR > r.tMin && inR < r.tMax;
		if (visitL) {
			if (nd.w3.z == 0) {
				if (sp < kSimpleStack) stack[sp++] = child_index(nd.w3.x);
			} else {
				occluded = anyhit_leaf(r, tris, child_index(nd.w3.x), nd.w3.z);
			}
		}
		if (visitR && !occluded) {
			if (nd.w3.w == 0) {
				if (sp < kSimpleStack) stack[sp++] = child_index(nd.w3.y);
			} else {
				occluded = anyhit_leaf(r, tris, child_index(nd.w3.y), nd.w3.w);
			}
		}
	}
	inLight[i] = occluded ? 0 : 1;
}

// ---- persistent wavefronts ----------------------------------------------------------------------
//
// A fixed grid of waves pulls rays from a shared counter, one atomic per wavefront for a window
// of kWindow rays. Every iteration, lanes without a ray are refilled from the window when at least
// `refillMin` of the wave's 64 lanes are idle (always when all are). The last wave to finish resets
// the counter, so every launch starts from zero without a memset.
//
// LDS (dynamic) holds, per block:
//   - castRay's pending stack, [wave][entry][lane] (one 4-B column per lane, so each push/pop is a
//     conflict-free 256-B row), kLdsDepth entries deep; deeper entries spill to a global array laid
//     out [entry][lane] (coalesced). Trees are at most 64 levels deep (checked at upload), so the
//     pending stack never exceeds kPersistentStack entries;
//   - the top treelet: nodes [0, topCount) of the device node array, which the upload lays out in
//     breadth-first order so they are the top levels of the tree -- where most node visits of
//     incoherent rays happen. They are copied once per block and read from LDS afterwards.

constexpr int kLdsDepth = 16;
constexpr int kLdsRows = kLdsDepth + 1;  // + one row that lanes without a push write to
constexpr int kPersistentStack = 72;
constexpr uint32_t kWindow = 64;
constexpr int kMaxBlock = 1024;

struct WaveQueue {
	uint32_t base = 0;   // next unclaimed ray of this wave's window
	uint32_t left = 0;   // rays left in the window
	bool exhausted = false;
};

// Assigns ray indices to the idle lanes of the wave (`idle` = ballot of idle lanes). Returns the
// lane's new ray index, or kInvalid. Wave-uniform state lives in `q`.
__device__ __forceinline__ uint32_t refill(WaveQueue& q, uint64_t idle, uint32_t* work,
                                           uint32_t n, bool isIdle)
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
		if (int(__lane_id()) == leader) b = atomicAdd(work, kWindow);
		b = __shfl(b, leader);
		if (b >= n) {
			q.exhausted = true;
			q.left = 0;
		} else {
			const uint32_t avail = (n - b) < kWindow ? (n - b) : kWindow;
			take2 = (need - take1) < avail ? (need - take1) : avail;
			base2 = b;
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
	uint32_t* lds;    // this wave's [entry][lane] block
	uint32_t* spill;  // this lane's spill column
	uint32_t stride;  // spill row stride (lanes in the grid)
	uint32_t lane;
	int sp;

	__device__ __forceinline__ void push(uint32_t v)
	{
		if (sp < kLdsDepth)
			lds[sp * 64 + lane] = v;
		else if (sp < kPersistentStack)
			spill[size_t(sp - kLdsDepth) * stride] = v;
		sp++;
	}
	// Pushes v if doPush. Lanes that do not push (or whose entry spills) store to the extra row, so
	// the LDS store needs no branch; only a spilling entry takes one.
	__device__ __forceinline__ void push_if(bool doPush, uint32_t v)
	{
		const bool toLds = doPush && sp < kLdsDepth;
		lds[(toLds ? sp : kLdsDepth) * 64 + lane] = v;
		if (doPush && sp >= kLdsDepth && sp < kPersistentStack)
			spill[size_t(sp - kLdsDepth) * stride] = v;
		sp += doPush ? 1 : 0;
	}
	__device__ __forceinline__ uint32_t pop()
	{
		sp--;
		if (sp < kLdsDepth) return lds[sp * 64 + lane];
		if (sp < kPersistentStack) return spill[size_t(sp - kLdsDepth) * stride];
		return 0u;
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

// A triangle step tests up to kTriBatch consecutive triangles of the current leaf range: the first
// triangle's words come through the step's unified loads, the others are loaded right after them
// so all are in flight together, and the tests then run in castRay's order (BVHTraversal.hpp:
// 78-93), so every acceptance is the one castRay makes. Leaves on the paths of costly incoherent
// rays hold about five triangles, so this shortens their chains of dependent steps.
constexpr uint32_t kTriBatch = 2;

struct TriExtra {
	float4 w[3 * kTriBatch - 4];  // words 4 .. 3*kTriBatch-1 of the range (word 3 is the 4th load)
};

__device__ __forceinline__ uint32_t load_tri_extra(const float4* src, uint32_t avail, TriExtra& x)
{
	// The words past a leaf's last triangle lie in the next triangle or in the 64 B of padding
	// after the array, so they are loaded whether or not the leaf has a second triangle.
	static_assert(kTriBatch == 2, "the padding covers the words of one more triangle");
	x.w[0] = as_global(src)[4];
	x.w[1] = as_global(src)[5];
	return avail < kTriBatch ? avail : kTriBatch;
}

// Tests triangles first .. first+cnt-1 (words a0..a3 then x); any hit stops at the first hit.
template <bool kAnyHit, bool kCount>
__device__ __forceinline__ void test_tris(const dev::Ray& r, float4 a0, float4 a1, float4 a2,
                                          float4 a3, const TriExtra& x, uint32_t cnt,
                                          uint32_t first, Closest& c, Counts& k, bool& done)
{
#pragma unroll
	for (uint32_t j = 0; j < kTriBatch; j++) {
		if (j >= cnt || done) break;
		const float4 p0 = j == 0 ? a0 : j == 1 ? a3 : x.w[3 * j - 4];
		const float4 e1 = j == 0 ? a1 : x.w[3 * j - 3];
		const float4 e2 = j == 0 ? a2 : x.w[3 * j - 2];
		float t, u, v;
		const bool hit = dev::triangle_pre(r, p0, e1, e2, t, u, v);
		if (kAnyHit) {
			done = hit && r.tMin <= t && t < r.tMax;
			if (done) c.tri = 0u;
		} else {
			if (hit && t < c.t && r.tMin <= t && t <= r.tMax) {
				c.tri = first + j;
				c.t = t;
				c.u = u;
				c.v = v;
				if (kCount) k.accepted++;
			}
			if (kCount) k.triTests++;
		}
	}
}

// Per-lane traversal state of the persistent kernels. Every iteration a lane does one unit of
// work in castRay's per-ray order: expand a node (decide both children, queue the triangles of
// visited leaf children -- left, then right -- and pick the next node), or test one queued
// triangle. Nodes are expanded only once the queued triangles are done, so each box decision
// sees the closest.t that castRay would see.
struct LaneTrace {
	uint32_t triCur, triEnd;    // triangles being tested
	uint32_t pendCur, pendEnd;  // right-leaf triangles queued behind them
	int32_t next;               // node castRay pops next, or -1
};

__device__ __forceinline__ void expand(const NodeWords& nd, bool visitL, bool visitR,
                                       LaneTrace& s, LaneStack& st)
{
	// Selects instead of branches: a node step starts with an empty triangle range, visited leaf
	// children become the range (left first, the right one queued behind it), the right inner
	// child is castRay's next pop, a left inner child with it is pushed, and with no inner child
	// the stack is popped.
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
	s.next = innerR ? int32_t(rFirst) : innerL ? int32_t(lFirst) : -1;
	if (!innerL && !innerR && st.sp > 0) s.next = int32_t(st.pop());
}

// LDS bytes of a persistent launch: the stacks of every wave plus the treelet.
__host__ __device__ __forceinline__ uint32_t persistent_lds_bytes(uint32_t blockThreads,
                                                                  uint32_t topCount)
{
	return (blockThreads / 64u) * uint32_t(kLdsRows) * 64u * 4u + topCount * 64u;
}

// Shared skeleton of the persistent closest-hit (kAnyHit = false) and any-hit kernels. Each
// iteration every busy lane fetches one record -- its next node (64 B, from the LDS treelet or
// global memory), its next triangle (48 B) or, if it was just refilled, its ray (32 B) -- through
// the same 128-bit load instructions, so the fetches of a wave go out together. Without a treelet
// (kTreelet false) every record comes from global memory and no LDS path is compiled.
template <bool kAnyHit, bool kCount, bool kTreelet>
__device__ __forceinline__ void persistent_trace(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint8_t* __restrict__ inLight,
	uint32_t n, uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	bool finiteBounds, uint32_t topCount, unsigned long long* __restrict__ counters)
{
	extern __shared__ float4 ldsRaw[];
	uint32_t* ldsStacks = reinterpret_cast<uint32_t*>(ldsRaw);
	float4* treelet = ldsRaw + (blockDim.x / 64u) * uint32_t(kLdsRows) * 64u / 4u;
	if (kTreelet) {
		for (uint32_t i = threadIdx.x; i < 4u * topCount; i += blockDim.x) treelet[i] = nodes[i];
		__syncthreads();
	}

	const uint32_t lane = __lane_id();
	LaneStack st;
	st.lds = ldsStacks + (threadIdx.x >> 6) * uint32_t(kLdsRows) * 64u;
	st.stride = gridDim.x * blockDim.x;
	st.spill = spill + (blockIdx.x * blockDim.x + threadIdx.x);
	st.lane = lane;
	st.sp = 0;

	WaveQueue q;
	uint32_t rayIdx = kInvalid;
	LaneTrace s = {0u, 0u, 0u, 0u, -1};
	dev::Ray r;
	bool fastSlab = false;
	Closest c;
	Counts k;

	while (true) {
		const uint64_t idle = __ballot(rayIdx == kInvalid);
		bool fresh = false;
		if (!q.exhausted && idle != 0 &&
		    (uint32_t(__popcll(idle)) >= refillMin || idle == ~0ull)) {
			const uint32_t got = refill(q, idle, work, n, rayIdx == kInvalid);
			if (got != kInvalid) {
				rayIdx = got;
				fresh = true;
			}
		}
		if (q.exhausted && __ballot(rayIdx != kInvalid) == 0) break;
		if (rayIdx == kInvalid) continue;

		// A refilled lane only loads and prepares its ray this iteration; its root node then comes
		// through the same path as every other node. Words 2-3 of a refilled lane come from the root
		// node instead of past its 32-B ray, so the global loads need no mask and issue back to
		// back; a triangle's word 3 is the next triangle's first word (the triangle array carries
		// 64 B of padding). Treelet nodes are read from LDS instead.
		const bool doTri = !fresh && s.triCur < s.triEnd;
		const bool inLds = kTreelet && !fresh && !doTri && uint32_t(s.next) < topCount;
		// Record address by integer selects (branches here cost a lone lane more than the math).
		const uint64_t base = fresh ? uint64_t(rays) : doTri ? uint64_t(tris) : uint64_t(nodes);
		const uint32_t words = fresh ? 2u : doTri ? 3u : 4u;
		const uint32_t index = fresh ? rayIdx : doTri ? s.triCur : uint32_t(s.next);
		const float4* src = reinterpret_cast<const float4*>(base + uint64_t(index * words) * 16u);
		const float4* src23 = fresh ? nodes : src;
		float4 a0, a1, a2, a3;
		if (inLds) {
			const LdsF4* t = as_lds(treelet) + 4u * uint32_t(s.next);
			a0 = t[0];
			a1 = t[1];
			a2 = t[2];
			a3 = t[3];
		} else {
			a0 = as_global(src)[0];
			a1 = as_global(src)[1];
			a2 = as_global(src23)[2];
			a3 = as_global(src23)[3];
		}
		TriExtra x;
		uint32_t cnt = 0;
		if (doTri) cnt = load_tri_extra(src, s.triEnd - s.triCur, x);

		bool done = false;
		if (fresh) {
			r = dev::make_ray(a0, a1);
			fastSlab = finiteBounds && r.nanFree;
			c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			st.sp = 0;
			s = {0u, 0u, 0u, 0u, 0};
		} else if (doTri) {
			test_tris<kAnyHit, kCount>(r, a0, a1, a2, a3, x, cnt, s.triCur, c, k, done);
			s.triCur += cnt;
			if (s.triCur == s.triEnd) {
				s.triCur = s.pendCur;
				s.triEnd = s.pendEnd;
				s.pendCur = s.pendEnd = 0u;
			}
		} else {
			NodeWords nd;
			nd.w0 = a0;
			nd.w1 = a1;
			nd.w2 = a2;
			nd.w3 = make_int4(__float_as_int(a3.x), __float_as_int(a3.y), __float_as_int(a3.z),
			                  __float_as_int(a3.w));
			const float tCurrMax = kAnyHit ? r.tMax : dev::min_first(r.tMax, c.t);
			bool visitL, visitR;
			decide(r, fastSlab, nd, tCurrMax, visitL, visitR);
			if (kCount) {
				k.nodes++;
				k.boxHits += uint32_t(visitL) + uint32_t(visitR);
			}
			expand(nd, visitL, visitR, s, st);
		}
		if (!done) done = s.triCur >= s.triEnd && s.next < 0;
		if (done) {
			if (kAnyHit)
				inLight[rayIdx] = (c.tri == 0u) ? 0 : 1;
			else
				hits[rayIdx] = make_uint4(c.tri, __float_as_uint(c.t), __float_as_uint(c.u),
				                          __float_as_uint(c.v));
			rayIdx = kInvalid;
		}
	}
	if (kCount) add_counts(counters, k);
	retire_wave(work);
}

template <bool kCount, bool kTreelet>
__global__ __launch_bounds__(kMaxBlock) void k_closest_persistent(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint4* __restrict__ hits, uint32_t n,
	uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	uint32_t finiteBounds, uint32_t topCount, unsigned long long* __restrict__ counters)
{
	persistent_trace<false, kCount, kTreelet>(nodes, tris, rays, hits, nullptr, n, work, spill,
	                                          refillMin, finiteBounds != 0, topCount, counters);
}

template <bool kTreelet>
__global__ __launch_bounds__(kMaxBlock) void k_anyhit_persistent(
	const float4* __restrict__ nodes, const float4* __restrict__ tris,
	const float4* __restrict__ rays, uint8_t* __restrict__ inLight, uint32_t n,
	uint32_t* __restrict__ work, uint32_t* __restrict__ spill, uint32_t refillMin,
	uint32_t finiteBounds, uint32_t topCount)
{
	persistent_trace<true, false, kTreelet>(nodes, tris, rays, nullptr, inLight, n, work, spill,
	                                        refillMin, finiteBounds != 0, topCount, nullptr);
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

template <bool kTreelet>
__global__ __launch_bounds__(kMaxBlock) void k_trace_batches(
	const BatchDesc* __restrict__ descs, const uint32_t* __restrict__ offsetsGlobal,
	uint32_t numBatches, uint32_t total, uint32_t* __restrict__ work,
	uint32_t* __restrict__ spill, uint32_t refillMin, const float4* __restrict__ treeletNodes,
	uint32_t topCount, BatchClock* __restrict__ clock)
{
	extern __shared__ float4 ldsRaw[];
	uint32_t* ldsStacks = reinterpret_cast<uint32_t*>(ldsRaw);
	float4* treelet = ldsRaw + (blockDim.x / 64u) * uint32_t(kLdsRows) * 64u / 4u;
	unsigned long long* blockFinish = reinterpret_cast<unsigned long long*>(treelet + 4u * topCount);
	uint32_t* blockDone = reinterpret_cast<uint32_t*>(blockFinish + (clock ? numBatches : 0u));
	uint32_t* offsets = blockDone + (clock ? numBatches : 0u);
	if (clock && threadIdx.x == 0) atomicMin(&clock->start, wall_clock64());
	for (uint32_t i = threadIdx.x; i < 4u * topCount; i += blockDim.x) treelet[i] = treeletNodes[i];
	for (uint32_t i = threadIdx.x; i <= numBatches; i += blockDim.x) offsets[i] = offsetsGlobal[i];
	if (clock)
		for (uint32_t i = threadIdx.x; i < numBatches; i += blockDim.x) {
			blockFinish[i] = 0ull;
			blockDone[i] = 0u;
		}
	__syncthreads();

	LaneStack st;
	st.lds = ldsStacks + (threadIdx.x >> 6) * uint32_t(kLdsRows) * 64u;
	st.stride = gridDim.x * blockDim.x;
	st.spill = spill + (blockIdx.x * blockDim.x + threadIdx.x);
	st.lane = __lane_id();
	st.sp = 0;

	WaveQueue q;
	uint32_t rayIdx = kInvalid;  // index within the batch
	uint32_t batch = 0;
	LaneTrace s = {0u, 0u, 0u, 0u, -1};
	dev::Ray r;
	bool fastSlab = false, anyHit = false, useTreelet = false;
	const float4* nodes = nullptr;
	const float4* tris = nullptr;
	void* out = nullptr;
	Closest c;

	while (true) {
		const uint64_t idle = __ballot(rayIdx == kInvalid);
		bool fresh = false;
		uint32_t g = kInvalid;
		if (!q.exhausted && idle != 0 &&
		    (uint32_t(__popcll(idle)) >= refillMin || idle == ~0ull)) {
			g = refill(q, idle, work, total, rayIdx == kInvalid);
			fresh = g != kInvalid;
		}
		if (q.exhausted && __ballot(rayIdx != kInvalid || fresh) == 0) break;

		const float4* src;
		const float4* src23;
		if (fresh) {
			batch = find_batch(offsets, numBatches, g);
			rayIdx = g - offsets[batch];
			const BatchDesc d = descs[batch];
			nodes = d.nodes;
			tris = d.tris;
			out = d.out;
			anyHit = d.kind != 0;
			useTreelet = d.nodes == treeletNodes;
			fastSlab = d.finiteBounds != 0;
			src = d.rays + 2 * size_t(rayIdx);
			src23 = nodes;
		} else {
			src = nullptr;
			src23 = nullptr;
		}
		if (rayIdx == kInvalid) continue;

		const bool doTri = !fresh && s.triCur < s.triEnd;
		const bool inLds = kTreelet && !fresh && !doTri && useTreelet && uint32_t(s.next) < topCount;
		if (!fresh) {
			src = doTri ? tris + 3u * s.triCur : nodes + 4u * uint32_t(s.next);
			src23 = src;
		}
		float4 a0, a1, a2, a3;
		if (inLds) {
			const LdsF4* t = as_lds(treelet) + 4u * uint32_t(s.next);
			a0 = t[0];
			a1 = t[1];
			a2 = t[2];
			a3 = t[3];
		} else {
			a0 = as_global(src)[0];
			a1 = as_global(src)[1];
			a2 = as_global(src23)[2];
			a3 = as_global(src23)[3];
		}
		TriExtra x;
		uint32_t cnt = 0;
		if (doTri) cnt = load_tri_extra(src, s.triEnd - s.triCur, x);

		bool done = false;
		if (fresh) {
			r = dev::make_ray(a0, a1);
			fastSlab = fastSlab && r.nanFree;
			c = {kInvalid, FLT_MAX, FLT_MAX, FLT_MAX};
			st.sp = 0;
			s = {0u, 0u, 0u, 0u, 0};
		} else if (doTri) {
			Counts none;
			if (anyHit)
				test_tris<true, false>(r, a0, a1, a2, a3, x, cnt, s.triCur, c, none, done);
			else
				test_tris<false, false>(r, a0, a1, a2, a3, x, cnt, s.triCur, c, none, done);
			s.triCur += cnt;
			if (s.triCur == s.triEnd) {
				s.triCur = s.pendCur;
				s.triEnd = s.pendEnd;
				s.pendCur = s.pendEnd = 0u;
			}
		} else {
			NodeWords nd;
			nd.w0 = a0;
			nd.w1 = a1;
			nd.w2 = a2;
			nd.w3 = make_int4(__float_as_int(a3.x), __float_as_int(a3.y), __float_as_int(a3.z),
			                  __float_as_int(a3.w));
			const float tCurrMax = anyHit ? r.tMax : dev::min_first(r.tMax, c.t);
			bool visitL, visitR;
			decide(r, fastSlab, nd, tCurrMax, visitL, visitR);
			expand(nd, visitL, visitR, s, st);
		}
		if (!done) done = s.triCur >= s.triEnd && s.next < 0;
		if (done) {
			if (anyHit)
				static_cast<uint8_t*>(out)[rayIdx] = (c.tri == 0u) ? 0 : 1;
			else
				static_cast<uint4*>(out)[rayIdx] = make_uint4(
					c.tri, __float_as_uint(c.t), __float_as_uint(c.u), __float_as_uint(c.v));
			rayIdx = kInvalid;
		}
		if (clock && done) {
			// Block-local LDS atomics; flushed to the batch table when the block retires.
			atomicAdd(&blockDone[batch], 1u);
			atomicMax(&blockFinish[batch], wall_clock64());
		}
	}
	if (clock) {
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
			if (clock) {
				for (uint32_t b = 0; b < numBatches; b++) atomicExch(&clock->done[b], 0u);
				clock->lastStart = clock->start;
				atomicExch(&clock->start, ~0ull);
			}
		}
	}
}

uint32_t gridFor(uint32_t n, uint32_t block) { return n / block + (n % block != 0u); }

// Persistent launch shape: grid capped by the work and by the spill area's lane count.
void persistentShape(const TraceLaunch& L, const SceneView& s, uint32_t n, uint32_t& grid,
                     uint32_t& block, uint32_t& top, uint32_t& ldsBytes)
{
	block = L.blockThreads;
	top = L.topNodes < s.topCount ? L.topNodes : s.topCount;
	grid = L.gridBlocks < gridFor(n, block) ? L.gridBlocks : gridFor(n, block);
	const uint32_t maxGrid = L.spillLanes / block;
	if (grid > maxGrid) grid = maxGrid;
	if (grid == 0) grid = 1;
	ldsBytes = persistent_lds_bytes(block, top);
}

}  // namespace

uint32_t persistentSpillWords(uint32_t lanes)
{
	return uint32_t(kPersistentStack - kLdsDepth) * lanes;
}

uint32_t persistentLdsBytes(uint32_t blockThreads, uint32_t topNodes)
{
	return persistent_lds_bytes(blockThreads, topNodes);
}

hipError_t persistentBlocksPerCU(uint32_t blockThreads, uint32_t topNodes, uint32_t* blocks)
{
	int a = 0;
	const void* fn = topNodes ? reinterpret_cast<const void*>(k_closest_persistent<false, true>)
	                          : reinterpret_cast<const void*>(k_closest_persistent<false, false>);
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
		&a, fn, int(blockThreads), persistent_lds_bytes(blockThreads, topNodes));
	*blocks = uint32_t(a > 0 ? a : 0);
	return e;
}

hipError_t launchClosest(const SceneView& s, const PhxRayIn* rays, uint32_t n, PhxRayHit* hits,
                         const TraceLaunch& L, unsigned long long* counters, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const float4* nodes = reinterpret_cast<const float4*>(s.nodes);
	const float4* tris = reinterpret_cast<const float4*>(s.trisPre);
	const float4* r4 = reinterpret_cast<const float4*>(rays);
	uint4* h4 = reinterpret_cast<uint4*>(hits);
	if (L.kernel == TraceKernel::Simple || n > 0x7FFFFFFFu) {
		if (counters)
			hipLaunchKernelGGL(k_closest_simple<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0,
			                   stream, nodes, tris, r4, h4, n, counters);
		else
			hipLaunchKernelGGL(k_closest_simple<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0,
			                   stream, nodes, tris, r4, h4, n, counters);
		return hipGetLastError();
	}
	uint32_t grid, block, top, lds;
	persistentShape(L, s, n, grid, block, top, lds);
	auto kernel = counters ? (top ? k_closest_persistent<true, true> : k_closest_persistent<true, false>)
	                       : (top ? k_closest_persistent<false, true> : k_closest_persistent<false, false>);
	hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, nodes, tris, r4, h4, n,
	                   L.work, L.spill, L.refillMin, s.finiteBounds, top, counters);
	return hipGetLastError();
}

hipError_t launchAnyHit(const SceneView& s, const PhxRayIn* rays, uint32_t n, uint8_t* inLight,
                        const TraceLaunch& L, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const float4* nodes = reinterpret_cast<const float4*>(s.nodes);
	const float4* tris = reinterpret_cast<const float4*>(s.trisPre);
	const float4* r4 = reinterpret_cast<const float4*>(rays);
	if (L.kernel == TraceKernel::Simple || n > 0x7FFFFFFFu) {
		hipLaunchKernelGGL(k_anyhit_simple, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, stream,
		                   nodes, tris, r4, inLight, n);
		return hipGetLastError();
	}
	uint32_t grid, block, top, lds;
	persistentShape(L, s, n, grid, block, top, lds);
	hipLaunchKernelGGL(top ? k_anyhit_persistent<true> : k_anyhit_persistent<false>, dim3(grid),
	                   dim3(block), lds, stream, nodes, tris, r4,
	                   inLight, n, L.work, L.spill, L.refillMin, s.finiteBounds, top);
	return hipGetLastError();
}

uint32_t batchLdsBytes(uint32_t blockThreads, uint32_t topNodes, uint32_t numBatches,
                       bool tracked)
{
	return persistent_lds_bytes(blockThreads, topNodes) + (numBatches + 1u) * 4u +
	       (tracked ? numBatches * 12u : 0u);
}

hipError_t launchBatches(const BatchDesc* descs, const uint32_t* offsets, uint32_t numBatches,
                         uint32_t total, const TraceLaunch& L, const float4* treeletNodes,
                         uint32_t treeletCount, BatchClock* clock, hipStream_t stream)
{
	if (total == 0) return hipSuccess;
	const uint32_t block = L.blockThreads;
	const uint32_t top = L.topNodes < treeletCount ? L.topNodes : treeletCount;
	uint32_t grid = L.gridBlocks < gridFor(total, block) ? L.gridBlocks : gridFor(total, block);
	if (grid > L.spillLanes / block) grid = L.spillLanes / block;
	if (grid == 0) grid = 1;
	hipLaunchKernelGGL(top ? k_trace_batches<true> : k_trace_batches<false>, dim3(grid), dim3(block),
	                   batchLdsBytes(block, top, numBatches, clock != nullptr), stream, descs,
	                   offsets, numBatches,
	                   total, L.work, L.spill, L.refillMin, treeletNodes, top, clock);
	return hipGetLastError();
}

}  // namespace phx
