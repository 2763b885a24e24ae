// Split-BVH (SBVH, Stich et al. 2009) builder producing the reference's sanitized node/triangle
// arrays bit for bit.
//
// Semantics followed (reference paths relative to its repository root):
//   - build parameters: source/phantasy_engine/sbvh-stuff/BVH.h:68-81 (splitAlpha 1e-5),
//     Util.h:71-109 (unit node/triangle SAH costs, min leaf 1, max leaf 0x7FFFFFF),
//     SplitBVHBuilder.h:36-41 (MaxDepth 64, MaxSpatialDepth 48, 32 spatial bins);
//   - node recursion, object/spatial split search and partition: SplitBVHBuilder.cpp:55-488;
//   - sort key (centroid sum along the axis, then triangle index): SplitBVHBuilder.cpp:104-113;
//   - AABB arithmetic (grow/intersect/area with std::min/std::max operand order):
//     sbvh-stuff/BVHNode.h:43-67 and sfzCore Vector.inl:425-470;
//   - conversion (right child first) and sanitising (left-first DFS node order, negated leaf
//     indices, end markers): ray_tracer_common/StaticBVHBuilder.cpp:19-130, 213-260.
//
// Differences in structure, not in output:
//   - Sorting uses 64-bit keys and std::sort. Within one node every reference has a distinct
//     triangle index (a spatial split sends the two halves of a triangle to different children),
//     so the comparator is a strict total order and every correct sort yields the same
//     permutation as the reference's quicksort (Sort.cpp:40-146). -0.0 and +0.0 compare equal in
//     the reference's `<`; the key canonicalises -0.0 to +0.0 so they tie and fall through to the
//     triangle index, as there. Non-finite vertices are rejected up front (no NaN keys).
//   - A subtree only reads its own slice of the reference stack, so large right subtrees are built
//     on other threads from a copy of their slice; leaf order is restored by concatenation.
//   - The tree is flattened straight into the sanitized layout instead of convert-then-sanitise.
//   - A root that the reference would make a leaf (<= 1 triangle, or a leaf SAH that wins at the
//     root) crashes the reference's conversion (getChildNode on a leaf); here it becomes one node
//     whose two children are the same leaf.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <thread>

#include "phx_internal.hpp"

namespace phx {
namespace {

constexpr float kF32Max = 3.402823466e+38f;  // FW_F32_MAX, sbvh-stuff/Util.h:34
constexpr int kMaxDepth = 64;
constexpr int kMaxSpatialDepth = 48;
constexpr int kNumBins = 32;
constexpr float kSplitAlpha = 1.0e-5f;
constexpr int kMinLeafSize = 1;
constexpr int kMaxLeafSize = 0x7FFFFFF;

// std::min / std::max: the first operand is returned when the comparison is false (NaN, ties).
inline float fmin_first(float a, float b) { return (b < a) ? b : a; }
inline float fmax_first(float a, float b) { return (a < b) ? b : a; }

// float -> int32 as x86 cvttss2si does it: truncation, 0x80000000 for NaN and out-of-range input.
// The reference's vec3 -> vec3i conversion (static_cast<int32_t>) compiles to exactly this.
inline int32_t cvt_trunc_i32(float f)
{
	if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
	return static_cast<int32_t>(f);
}

struct Box {
	float mn[3] = {kF32Max, kF32Max, kF32Max};
	float mx[3] = {-kF32Max, -kF32Max, -kF32Max};

	void growPoint(const float* p)
	{
		for (int k = 0; k < 3; k++) {
			mn[k] = fmin_first(mn[k], p[k]);
			mx[k] = fmax_first(mx[k], p[k]);
		}
	}
	void growBox(const Box& b)
	{
		growPoint(b.mn);
		growPoint(b.mx);
	}
	void intersect(const Box& b)
	{
		for (int k = 0; k < 3; k++) {
			mn[k] = fmax_first(mn[k], b.mn[k]);
			mx[k] = fmin_first(mx[k], b.mx[k]);
		}
	}
	bool valid() const { return mn[0] <= mx[0] && mn[1] <= mx[1] && mn[2] <= mx[2]; }
	float area() const
	{
		if (!valid()) return 0.0f;
		float dx = mx[0] - mn[0];
		float dy = mx[1] - mn[1];
		float dz = mx[2] - mn[2];
		float s = dx * dy;
		s = s + dy * dz;
		s = s + dz * dx;
		return s * 2.0f;
	}
};

struct Ref {
	int32_t tri;
	Box bounds;
};

// Tree produced by the recursion (the reference's nv::InnerNode / nv::LeafNode).
struct TNode {
	Box bounds;
	TNode* child[2] = {nullptr, nullptr};  // [0] = left, [1] = right; null for leaves
	std::vector<int32_t> tris;             // leaf only, in leaf-creation order
	bool leaf() const { return child[0] == nullptr; }
};

struct NodeSpec {
	int numRef = 0;
	Box bounds;
};

struct ObjectSplit {
	float sah = kF32Max;
	int sortDim = 0;
	int numLeft = 0;
	Box leftBounds;
	Box rightBounds;
};

struct SpatialSplit {
	float sah = kF32Max;
	int dim = 0;
	float pos = 0.0f;
};

// True for a default-constructed box: any growBox leaves at least one bound changed.
inline bool untouched(const Box& b)
{
	for (int k = 0; k < 3; k++)
		if (b.mn[k] != kF32Max || b.mx[k] != -kF32Max) return false;
	return true;
}

struct SpatialBin {
	Box bounds;
	int enter = 0;
	int exit = 0;
};

struct KeyPos {
	uint64_t key;
	uint32_t pos;
};

// Shared, read-only inputs of one build.
struct Shared {
	const PhxTriangleVertices* tris = nullptr;
	float minOverlap = 0.0f;
	unsigned maxThreads = 1;
	std::atomic<unsigned> activeThreads{1};
	std::atomic<bool> outOfMemory{false};
};

constexpr int kParallelMinRefs = 8192;
// Nodes at least this large search their splits with several threads (the three axes of the object
// split in parallel, the spatial binning in chunks). The results are those of the sequential
// search: sort keys are unique within a node, and bin bounds and counts reduce exactly.
constexpr int kWideNodeRefs = 1 << 16;

// Per-thread build state: the reference's m_refStack plus scratch arrays.
class Worker {
public:
	Worker(Shared& shared, std::vector<Ref>&& stack) : mShared(shared), mStack(std::move(stack)) {}

	TNode* build(const NodeSpec& spec, int level) { return buildNode(spec, level); }

private:
	Shared& mShared;
	std::vector<Ref> mStack;
	std::vector<Box> mRightBounds;
	std::vector<KeyPos> mKeys;
	std::vector<Ref> mTmp;
	SpatialBin mBins[3][kNumBins];

	const float* vertex(int32_t tri, int k) const
	{
		const PhxTriangleVertices& t = mShared.tris[tri];
		return k == 0 ? t.v0 : (k == 1 ? t.v1 : t.v2);
	}

	static float ntf(int n) { return static_cast<float>(n); }  // getTriangleCost(n), unit cost

	// Sort keys of [start, end) of the stack along dim: (min[dim] + max[dim], tri), sorted.
	void sortedKeys(int start, int end, int dim, std::vector<KeyPos>& keys) const
	{
		int n = end - start;
		keys.resize(n);
		for (int i = 0; i < n; i++) {
			const Ref& r = mStack[start + i];
			float c = r.bounds.mn[dim] + r.bounds.mx[dim];
			c = c + 0.0f;  // -0.0 -> +0.0: the reference's `<` treats them as equal
			uint32_t bits;
			std::memcpy(&bits, &c, 4);
			bits = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
			keys[i].key = (uint64_t(bits) << 32) | uint32_t(r.tri);
			keys[i].pos = uint32_t(start + i);
		}
		std::sort(keys.begin(), keys.end(),
		          [](const KeyPos& a, const KeyPos& b) { return a.key < b.key; });
	}

	void applyOrder(int start, const std::vector<KeyPos>& keys)
	{
		const int n = int(keys.size());
		mTmp.resize(n);
		for (int i = 0; i < n; i++) mTmp[i] = mStack[keys[i].pos];
		std::copy(mTmp.begin(), mTmp.end(), mStack.begin() + start);
	}

	// Sorts [start, end) of the stack by (min[dim] + max[dim], tri).
	void sortRange(int start, int end, int dim)
	{
		if (end - start < 2) return;
		sortedKeys(start, end, dim, mKeys);
		applyOrder(start, mKeys);
	}

	// Best object split along one axis over refs in `keys` order: the first split index with the
	// smallest SAH (SplitBVHBuilder.cpp:211-231).
	void sweepAxis(int numRef, float nodeSAH, int dim, const std::vector<KeyPos>& keys,
	               std::vector<Box>& rightBounds, ObjectSplit& best) const
	{
		rightBounds.resize(numRef);
		Box right;
		for (int i = numRef - 1; i > 0; i--) {
			right.growBox(mStack[keys[i].pos].bounds);
			rightBounds[i - 1] = right;
		}
		Box left;
		for (int i = 1; i < numRef; i++) {
			left.growBox(mStack[keys[i - 1].pos].bounds);
			float sah = nodeSAH + left.area() * ntf(i);
			sah = sah + rightBounds[i - 1].area() * ntf(numRef - i);
			if (sah < best.sah) {
				best.sah = sah;
				best.sortDim = dim;
				best.numLeft = i;
				best.leftBounds = left;
				best.rightBounds = rightBounds[i - 1];
			}
		}
	}

	// findObjectSplit for wide nodes: the three axes sorted and swept by three threads, the best
	// chosen as the sequential loop would (lowest axis among equal SAHs), and the stack left sorted
	// along the last axis, as the sequential search leaves it.
	ObjectSplit findObjectSplitWide(const NodeSpec& spec, float nodeSAH)
	{
		const int start = int(mStack.size()) - spec.numRef;
		const int end = int(mStack.size());
		std::vector<KeyPos> keys[3];
		std::vector<Box> rb[3];
		ObjectSplit best[3];
		auto axis = [&](int dim) {
			sortedKeys(start, end, dim, keys[dim]);
			sweepAxis(spec.numRef, nodeSAH, dim, keys[dim], rb[dim], best[dim]);
		};
		std::thread t1(axis, 1), t2(axis, 2);
		axis(0);
		t1.join();
		t2.join();
		ObjectSplit split;
		for (int dim = 0; dim < 3; dim++)
			if (best[dim].sah < split.sah) split = best[dim];
		applyOrder(start, keys[2]);
		return split;
	}

	TNode* createLeaf(const NodeSpec& spec)
	{
		TNode* node = new TNode();
		node->bounds = spec.bounds;
		node->tris.reserve(spec.numRef);
		for (int i = 0; i < spec.numRef; i++) {
			node->tris.push_back(mStack.back().tri);
			mStack.pop_back();
		}
		return node;
	}

	ObjectSplit findObjectSplit(const NodeSpec& spec, float nodeSAH)
	{
		if (spec.numRef >= kWideNodeRefs) return findObjectSplitWide(spec, nodeSAH);
		ObjectSplit split;
		int start = int(mStack.size()) - spec.numRef;
		int end = int(mStack.size());
		if (int(mRightBounds.size()) < spec.numRef) mRightBounds.resize(spec.numRef);
		for (int dim = 0; dim < 3; dim++) {
			sortRange(start, end, dim);
			const Ref* refs = mStack.data() + start;
			Box right;
			for (int i = spec.numRef - 1; i > 0; i--) {
				right.growBox(refs[i].bounds);
				mRightBounds[i - 1] = right;
			}
			Box left;
			for (int i = 1; i < spec.numRef; i++) {
				left.growBox(refs[i - 1].bounds);
				float sah = nodeSAH + left.area() * ntf(i);
				sah = sah + mRightBounds[i - 1].area() * ntf(spec.numRef - i);
				if (sah < split.sah) {
					split.sah = sah;
					split.sortDim = dim;
					split.numLeft = i;
					split.leftBounds = left;
					split.rightBounds = mRightBounds[i - 1];
				}
			}
		}
		return split;
	}

	void performObjectSplit(NodeSpec& left, NodeSpec& right, const NodeSpec& spec,
	                        const ObjectSplit& split)
	{
		int start = int(mStack.size()) - spec.numRef;
		sortRange(start, int(mStack.size()), split.sortDim);
		left.numRef = split.numLeft;
		left.bounds = split.leftBounds;
		right.numRef = spec.numRef - split.numLeft;
		right.bounds = split.rightBounds;
	}

	// Clips `ref` at the plane x[dim] = pos (SplitBVHBuilder.cpp:445-488).
	void splitReference(Ref& left, Ref& right, const Ref& ref, int dim, float pos) const
	{
		left.tri = right.tri = ref.tri;
		left.bounds = Box();
		right.bounds = Box();
		const float* v1 = vertex(ref.tri, 2);
		for (int i = 0; i < 3; i++) {
			const float* v0 = v1;
			v1 = vertex(ref.tri, i);
			float v0p = v0[dim];
			float v1p = v1[dim];
			if (v0p <= pos) left.bounds.growPoint(v0);
			if (v0p >= pos) right.bounds.growPoint(v0);
			if ((v0p < pos && v1p > pos) || (v0p > pos && v1p < pos)) {
				float f = (pos - v0p) / (v1p - v0p);
				f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
				float omf = 1.0f - f;
				float p[3];
				for (int k = 0; k < 3; k++) {
					float a = v0[k] * omf;
					float b = v1[k] * f;
					p[k] = a + b;
				}
				left.bounds.growPoint(p);
				right.bounds.growPoint(p);
			}
		}
		left.bounds.mx[dim] = pos;
		right.bounds.mn[dim] = pos;
		left.bounds.intersect(ref.bounds);
		right.bounds.intersect(ref.bounds);
	}

	SpatialSplit findSpatialSplit(const NodeSpec& spec, float nodeSAH)
	{
		float origin[3], binSize[3], invBinSize[3];
		for (int k = 0; k < 3; k++) {
			origin[k] = spec.bounds.mn[k];
			float d = spec.bounds.mx[k] - origin[k];
			binSize[k] = d * (1.0f / float(kNumBins));
		}
		for (int k = 0; k < 3; k++) invBinSize[k] = 1.0f / binSize[k];
		for (int dim = 0; dim < 3; dim++)
			for (int i = 0; i < kNumBins; i++) mBins[dim][i] = SpatialBin();

		const int size = int(mStack.size());
		auto binRange = [&](int from, int to, SpatialBin (*bins)[kNumBins]) {
			for (int refIdx = from; refIdx < to; refIdx++) {
				const Ref ref = mStack[refIdx];
				int first[3], last[3];
				for (int k = 0; k < 3; k++) {
					int f = cvt_trunc_i32((ref.bounds.mn[k] - origin[k]) * invBinSize[k]);
					f = f < 0 ? 0 : (f > kNumBins - 1 ? kNumBins - 1 : f);
					int l = cvt_trunc_i32((ref.bounds.mx[k] - origin[k]) * invBinSize[k]);
					l = l < f ? f : (l > kNumBins - 1 ? kNumBins - 1 : l);
					first[k] = f;
					last[k] = l;
				}
				for (int dim = 0; dim < 3; dim++) {
					Ref curr = ref;
					for (int i = first[dim]; i < last[dim]; i++) {
						Ref l, r;
						splitReference(l, r, curr, dim, origin[dim] + binSize[dim] * float(i + 1));
						bins[dim][i].bounds.growBox(l.bounds);
						curr = r;
					}
					bins[dim][last[dim]].bounds.growBox(curr.bounds);
					bins[dim][first[dim]].enter++;
					bins[dim][last[dim]].exit++;
				}
			}
		};
		const int from = size - spec.numRef;
		const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
		const int chunks = spec.numRef >= kWideNodeRefs ? int(std::min(hw, 16u)) : 1;
		if (chunks <= 1) {
			binRange(from, size, mBins);
		} else {
			// Bins per chunk, merged after: bounds grow by min/max and counts add, both exact and
			// independent of order. A chunk bin nothing was added to is skipped: growing a box by
			// an empty box makes it unbounded (the reference's AABB::grow semantics), which the
			// sequential binning never does for a bin without references.
			std::vector<SpatialBin> part(size_t(chunks) * 3 * kNumBins);
			std::vector<std::thread> threads;
			const int per = (spec.numRef + chunks - 1) / chunks;
			for (int c = 0; c < chunks; c++) {
				const int a = from + c * per, b = std::min(size, a + per);
				auto* bins = reinterpret_cast<SpatialBin(*)[kNumBins]>(&part[size_t(c) * 3 * kNumBins]);
				if (a >= b) continue;
				if (c == 0)
					continue;  // chunk 0 runs on this thread below
				threads.emplace_back(binRange, a, b, bins);
			}
			binRange(from, std::min(size, from + per),
			         reinterpret_cast<SpatialBin(*)[kNumBins]>(&part[0]));
			for (std::thread& t : threads) t.join();
			for (int c = 0; c < chunks; c++)
				for (int dim = 0; dim < 3; dim++)
					for (int i = 0; i < kNumBins; i++) {
						const SpatialBin& b = part[(size_t(c) * 3 + dim) * kNumBins + i];
						if (!untouched(b.bounds)) mBins[dim][i].bounds.growBox(b.bounds);
						mBins[dim][i].enter += b.enter;
						mBins[dim][i].exit += b.exit;
					}
		}

		SpatialSplit split;
		if (int(mRightBounds.size()) < kNumBins) mRightBounds.resize(kNumBins);
		for (int dim = 0; dim < 3; dim++) {
			Box right;
			for (int i = kNumBins - 1; i > 0; i--) {
				right.growBox(mBins[dim][i].bounds);
				mRightBounds[i - 1] = right;
			}
			Box left;
			int leftNum = 0;
			int rightNum = spec.numRef;
			for (int i = 1; i < kNumBins; i++) {
				left.growBox(mBins[dim][i - 1].bounds);
				leftNum += mBins[dim][i - 1].enter;
				rightNum -= mBins[dim][i - 1].exit;
				float sah = nodeSAH + left.area() * ntf(leftNum);
				sah = sah + mRightBounds[i - 1].area() * ntf(rightNum);
				if (sah < split.sah) {
					split.sah = sah;
					split.dim = dim;
					split.pos = origin[dim] + binSize[dim] * float(i);
				}
			}
		}
		return split;
	}

	// SAH of two children: area * count of one plus area * count of the other, in that order.
	static float sahPair(float areaA, float countA, float areaB, float countB)
	{
		const float a = areaA * countA;
		return a + areaB * countB;
	}

	// Spatial split of the node's references (the last spec.numRef entries of mStack) at
	// split.pos on axis split.dim. First the references are sorted in place into three runs:
	// [base, lo) entirely below the plane, [lo, hi) straddling it, [hi, end) entirely above (a
	// reference moved to the upper run is examined again at the same index). Then each straddler,
	// in stack order, either stays whole on one side or is clipped into both halves (the upper
	// half appended to the stack), whichever gives the children the smallest SAH -- Stich et al.
	// 2009's reference unsplitting; ties keep it below, then above. Every swap and append happens
	// in a fixed order, so the output is deterministic.
	void performSpatialSplit(NodeSpec& left, NodeSpec& right, const NodeSpec& spec,
	                         const SpatialSplit& split)
	{
		std::vector<Ref>& stack = mStack;
		const int axis = split.dim;
		const float plane = split.pos;
		const int base = int(stack.size()) - spec.numRef;
		int lo = base;
		int hi = int(stack.size());
		Box below, above;

		int i = lo;
		while (i < hi) {
			const Box& b = stack[i].bounds;
			if (b.mx[axis] <= plane) {
				below.growBox(b);
				std::swap(stack[i], stack[lo]);
				lo++;
				i++;
			} else if (b.mn[axis] >= plane) {
				above.growBox(b);
				hi--;
				std::swap(stack[i], stack[hi]);
			} else {
				i++;
			}
		}

		while (lo < hi) {
			Ref half[2];
			splitReference(half[0], half[1], stack[lo], axis, plane);
			Box belowWhole = below, aboveWhole = above, belowHalf = below, aboveHalf = above;
			belowWhole.growBox(stack[lo].bounds);
			aboveWhole.growBox(stack[lo].bounds);
			belowHalf.growBox(half[0].bounds);
			aboveHalf.growBox(half[1].bounds);
			const int nBelow = lo - base;
			const int nAbove = int(stack.size()) - hi;
			const float keepBelow =
				sahPair(belowWhole.area(), ntf(nBelow + 1), above.area(), ntf(nAbove));
			const float keepAbove =
				sahPair(below.area(), ntf(nBelow), aboveWhole.area(), ntf(nAbove + 1));
			const float clip =
				sahPair(belowHalf.area(), ntf(nBelow + 1), aboveHalf.area(), ntf(nAbove + 1));
			const float best = fmin_first(fmin_first(keepBelow, keepAbove), clip);
			if (best == keepBelow) {
				below = belowWhole;
				lo++;
			} else if (best == keepAbove) {
				above = aboveWhole;
				hi--;
				std::swap(stack[lo], stack[hi]);
			} else {
				below = belowHalf;
				above = aboveHalf;
				stack[lo] = half[0];
				lo++;
				stack.push_back(half[1]);
			}
		}
		left.bounds = below;
		right.bounds = above;
		left.numRef = lo - base;
		right.numRef = int(stack.size()) - hi;
	}

	TNode* buildNode(const NodeSpec& spec, int level)
	{
		if (spec.numRef <= kMinLeafSize || level >= kMaxDepth) return createLeaf(spec);

		float area = spec.bounds.area();
		float leafSAH = area * ntf(spec.numRef);
		float nodeSAH = area * 2.0f;
		ObjectSplit object = findObjectSplit(spec, nodeSAH);

		SpatialSplit spatial;
		if (level < kMaxSpatialDepth) {
			Box overlap = object.leftBounds;
			overlap.intersect(object.rightBounds);
			if (overlap.area() >= mShared.minOverlap) spatial = findSpatialSplit(spec, nodeSAH);
		}

		float minSAH = fmin_first(fmin_first(leafSAH, object.sah), spatial.sah);
		if (minSAH == leafSAH && spec.numRef <= kMaxLeafSize) return createLeaf(spec);

		NodeSpec left, right;
		if (minSAH == spatial.sah) performSpatialSplit(left, right, spec, spatial);
		if (!left.numRef || !right.numRef) performObjectSplit(left, right, spec, object);

		// The right child is built first; its slice is on top of the stack.
		TNode* rightNode = nullptr;
		std::future<TNode*> rightFuture;
		bool spawned = false;
		if (right.numRef >= kParallelMinRefs && left.numRef >= kParallelMinRefs) {
			unsigned active = mShared.activeThreads.load();
			while (active < mShared.maxThreads &&
			       !mShared.activeThreads.compare_exchange_weak(active, active + 1)) {}
			if (active < mShared.maxThreads) {
				int rs = int(mStack.size()) - right.numRef;
				std::vector<Ref> slice(mStack.begin() + rs, mStack.end());
				mStack.resize(rs);
				Shared* sh = &mShared;
				int lvl = level + 1;
				try {
					rightFuture = std::async(std::launch::async,
					                         [sh, lvl, right, s = std::move(slice)]() mutable {
						                         Worker w(*sh, std::move(s));
						                         TNode* r = w.build(right, lvl);
						                         sh->activeThreads.fetch_sub(1);
						                         return r;
					                         });
					spawned = true;
				} catch (...) {
					mShared.activeThreads.fetch_sub(1);
					throw;
				}
			}
		}
		if (!spawned) rightNode = buildNode(right, level + 1);
		TNode* leftNode = buildNode(left, level + 1);
		if (spawned) rightNode = rightFuture.get();

		TNode* node = new TNode();
		node->bounds = spec.bounds;
		node->child[0] = leftNode;
		node->child[1] = rightNode;
		return node;
	}
};

void deleteTree(TNode* n)
{
	if (!n) return;
	if (!n->leaf()) {
		deleteTree(n->child[0]);
		deleteTree(n->child[1]);
	}
	delete n;
}

// Leaf triangle ranges follow leaf-creation order, i.e. a DFS that visits the right child first
// (StaticBVHBuilder.cpp:115-128 assigns them with a running counter in that order).
void assignTriangles(TNode* n, std::vector<int32_t>& triOrder,
                     std::vector<std::pair<TNode*, int32_t>>& leafStart)
{
	if (n->leaf()) {
		leafStart.emplace_back(n, int32_t(triOrder.size()));
		triOrder.insert(triOrder.end(), n->tris.begin(), n->tris.end());
		return;
	}
	assignTriangles(n->child[1], triOrder, leafStart);
	assignTriangles(n->child[0], triOrder, leafStart);
}

void setChildBox(PhxBVHNode& node, int side, const Box& b)
{
	if (side == 0) {
		node.fData[0] = b.mn[0];
		node.fData[1] = b.mn[1];
		node.fData[2] = b.mn[2];
		node.fData[3] = b.mx[0];
		node.fData[4] = b.mx[1];
		node.fData[5] = b.mx[2];
	} else {
		node.fData[6] = b.mn[0];
		node.fData[7] = b.mn[1];
		node.fData[8] = b.mn[2];
		node.fData[9] = b.mx[0];
		node.fData[10] = b.mx[1];
		node.fData[11] = b.mx[2];
	}
}

struct Flattener {
	std::vector<PhxBVHNode>& out;
	const std::vector<std::pair<TNode*, int32_t>>& leafStart;
	std::vector<std::pair<TNode*, int32_t>> sorted;

	Flattener(std::vector<PhxBVHNode>& o, const std::vector<std::pair<TNode*, int32_t>>& ls)
		: out(o), leafStart(ls), sorted(ls)
	{
		std::sort(sorted.begin(), sorted.end());
	}

	int32_t firstTri(TNode* leaf) const
	{
		auto it = std::lower_bound(sorted.begin(), sorted.end(), std::make_pair(leaf, INT32_MIN));
		return it->second;
	}

	// Pre-order, left child first (sanitizeInternal, StaticBVHBuilder.cpp:19-78).
	int32_t emit(TNode* n)
	{
		int32_t idx = int32_t(out.size());
		out.push_back(PhxBVHNode{});
		for (int side = 0; side < 2; side++) {
			TNode* c = n->child[side];
			setChildBox(out[idx], side, c->bounds);
			if (c->leaf()) {
				out[idx].iData[side] = ~firstTri(c);
				out[idx].iData[2 + side] = int32_t(c->tris.size());
			} else {
				int32_t ci = emit(c);
				out[idx].iData[side] = ci;
				out[idx].iData[2 + side] = 0;
			}
		}
		return idx;
	}
};

}  // namespace

phx_status buildSBVH(const PhxTriangleVertices* tris, const PhxTriangleData* datas, uint32_t n,
                     uint32_t numThreads, BuiltBVH& out)
{
	if (!tris) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_bvh_build: triangles is NULL");
	if (n == 0) return fail(PHX_ERR_INVALID_SCENE, "phx_bvh_build: scene has no triangles");
	if (n > 0x3FFFFFFFu) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_bvh_build: too many triangles");
	for (uint32_t i = 0; i < n; i++) {
		const float* v[3] = {tris[i].v0, tris[i].v1, tris[i].v2};
		for (int k = 0; k < 3; k++)
			for (int c = 0; c < 3; c++)
				if (!std::isfinite(v[k][c]))
					return fail(PHX_ERR_INVALID_SCENE,
					            "phx_bvh_build: triangle %u has a non-finite vertex", i);
	}

	Shared shared;
	shared.tris = tris;
	unsigned hw = std::thread::hardware_concurrency();
	shared.maxThreads = numThreads ? numThreads : (hw ? hw : 1);

	// Root references and bounds (SplitBVHBuilder.cpp:62-85).
	std::vector<Ref> stack(n);
	NodeSpec root;
	root.numRef = int(n);
	for (uint32_t i = 0; i < n; i++) {
		stack[i].tri = int32_t(i);
		stack[i].bounds.growPoint(tris[i].v0);
		stack[i].bounds.growPoint(tris[i].v1);
		stack[i].bounds.growPoint(tris[i].v2);
		root.bounds.growBox(stack[i].bounds);
	}
	shared.minOverlap = root.bounds.area() * kSplitAlpha;

	TNode* tree = nullptr;
	try {
		Worker w(shared, std::move(stack));
		tree = w.build(root, 0);
	} catch (const std::bad_alloc&) {
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_bvh_build: out of host memory");
	} catch (const std::exception& e) {
		return fail(PHX_ERR_INVALID_SCENE, "phx_bvh_build: %s", e.what());
	}

	if (tree->leaf()) {
		TNode* wrap = new TNode();
		wrap->bounds = tree->bounds;
		wrap->child[0] = tree;
		wrap->child[1] = tree;
		tree = wrap;
	}

	std::vector<int32_t> triOrder;
	std::vector<std::pair<TNode*, int32_t>> leafStart;
	triOrder.reserve(n + n / 4);
	if (tree->child[0] == tree->child[1]) {
		leafStart.emplace_back(tree->child[0], 0);
		triOrder = tree->child[0]->tris;
	} else {
		assignTriangles(tree, triOrder, leafStart);
	}

	out.nodes.clear();
	out.nodes.reserve(leafStart.size());
	Flattener(out.nodes, leafStart).emit(tree);

	out.triangles.resize(triOrder.size());
	out.datas.assign(triOrder.size(), PhxTriangleData{});
	for (size_t i = 0; i < triOrder.size(); i++) {
		out.triangles[i] = tris[triOrder[i]];
		if (datas) out.datas[i] = datas[triOrder[i]];
	}
	// End markers: w = 0 except the last triangle of each leaf, where w = -1.
	for (auto& t : out.triangles) t.v0[3] = t.v1[3] = t.v2[3] = 0.0f;
	for (const auto& ls : leafStart) {
		int32_t last = ls.second + int32_t(ls.first->tris.size()) - 1;
		out.triangles[last].v0[3] = out.triangles[last].v1[3] = out.triangles[last].v2[3] = -1.0f;
	}

	if (tree->child[0] == tree->child[1]) {
		delete tree->child[0];
		delete tree;
	} else {
		deleteTree(tree);
	}
	return ok();
}

phx_status validateBVH(const PhxBVHNode* nodes, uint32_t numNodes,
                       const PhxTriangleVertices* tris, uint32_t numTris, uint32_t* depthOut)
{
	if (!nodes || numNodes == 0) return fail(PHX_ERR_MALFORMED_BVH, "BVH has no nodes");
	if (!tris || numTris == 0) return fail(PHX_ERR_MALFORMED_BVH, "BVH has no triangles");
	std::vector<uint8_t> seen(numNodes, 0);
	struct Item {
		uint32_t node;
		uint32_t depth;
	};
	std::vector<Item> stack;
	stack.push_back({0, 1});
	seen[0] = 1;
	uint32_t maxDepth = 0;
	uint32_t reached = 1;
	while (!stack.empty()) {
		Item it = stack.back();
		stack.pop_back();
		if (it.depth > 64)
			return fail(PHX_ERR_MALFORMED_BVH, "BVH deeper than 64 levels (node %u)", it.node);
		maxDepth = std::max(maxDepth, it.depth);
		const PhxBVHNode& nd = nodes[it.node];
		for (int side = 0; side < 2; side++) {
			int32_t idx = nd.iData[side];
			int32_t cnt = nd.iData[2 + side];
			if (cnt < 0)
				return fail(PHX_ERR_MALFORMED_BVH, "node %u: negative triangle count", it.node);
			if (cnt == 0) {
				if (idx < 0 || uint32_t(idx) >= numNodes)
					return fail(PHX_ERR_MALFORMED_BVH, "node %u: child index %d out of range",
					            it.node, idx);
				if (seen[idx])
					return fail(PHX_ERR_MALFORMED_BVH, "node %d reached twice (cycle/DAG)", idx);
				seen[idx] = 1;
				reached++;
				stack.push_back({uint32_t(idx), it.depth + 1});
			} else {
				if (idx >= 0)
					return fail(PHX_ERR_MALFORMED_BVH, "node %u: leaf index %d is not negated",
					            it.node, idx);
				int64_t first = ~int64_t(idx);
				if (first + cnt > int64_t(numTris))
					return fail(PHX_ERR_MALFORMED_BVH, "node %u: leaf range [%lld,+%d) out of range",
					            it.node, (long long)first, cnt);
				for (int64_t t = first; t < first + cnt; t++) {
					bool isLast = t == first + cnt - 1;
					bool marked = tris[t].v0[3] < 0.0f;
					if (marked != isLast)
						return fail(PHX_ERR_MALFORMED_BVH,
						            "node %u: leaf [%lld,+%d) end marker wrong at triangle %lld",
						            it.node, (long long)first, cnt, (long long)t);
				}
			}
		}
	}
	if (reached != numNodes)
		return fail(PHX_ERR_MALFORMED_BVH, "%u of %u nodes unreachable from the root",
		            numNodes - reached, numNodes);
	if (depthOut) *depthOut = maxDepth;
	return ok();
}

}  // namespace phx
