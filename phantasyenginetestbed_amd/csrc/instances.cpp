// Two-level scenes: instances of uploaded meshes (each an object-space SBVH, a PhxScene) placed by
// affine transforms, under a top tree over their world boxes. This is what the reference's
// dynamic-mesh path sets out to do (DynamicBVH.hpp:13-26, DynamicBVH.cpp:16-329: per-mesh BVHs
// under an OuterBVH of their root boxes, for DynObject {meshIndex, transform}); its CUDA backend
// never traced them (CudaTracerRenderer.cpp:317-327, :412-415). Here meshes keep their SBVH in
// object space and only the top tree is rebuilt when instances move.
//
// Semantics (what the kernels and the oracle compute): an instance is traced iff the ray enters
// its world box (the castRay box test with tIn <= min(tMax, best t)); it runs castRay on its own
// SBVH with the ray moved into object space (o' = A o + b, d' = A d, same tMin/tMax), from a miss;
// the closest instance result wins, the lowest instance index on equal t; a static scene traced
// with the instances counts as instance -1 (it wins ties).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "runtime.hpp"

using namespace phx;

struct PhxInstanceSet {
	PhxContext* ctx = nullptr;
	uint32_t count = 0;
	PhxBVHNode* top = nullptr;    // device: top tree, one instance per leaf
	DevInstance* inst = nullptr;  // device: instances in top-tree leaf order
	DevInstance* byIndex = nullptr;  // device: instances in the caller's order (hit records)
	std::vector<const PhxScene*> meshes;  // distinct meshes, held until release
	uint64_t stamp = 0;
};

namespace {

struct WorldBox {
	float mn[3], mx[3];
};

// world -> object rows (A | b) of the object -> world rows (R | t): A = R^-1 by the adjugate over
// the determinant, b = -A t, in double, each rounded once to float. False if R is singular or a
// value is not finite.
bool invertAffine(const float m[12], float out[12])
{
	double r[3][3], t[3];
	for (int i = 0; i < 3; i++) {
		for (int j = 0; j < 3; j++) r[i][j] = m[4 * i + j];
		t[i] = m[4 * i + 3];
	}
	const double c00 = r[1][1] * r[2][2] - r[1][2] * r[2][1];
	const double c01 = r[1][2] * r[2][0] - r[1][0] * r[2][2];
	const double c02 = r[1][0] * r[2][1] - r[1][1] * r[2][0];
	const double det = (r[0][0] * c00 + r[0][1] * c01) + r[0][2] * c02;
	if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
	double a[3][3];
	a[0][0] = c00 / det;
	a[1][0] = c01 / det;
	a[2][0] = c02 / det;
	a[0][1] = (r[0][2] * r[2][1] - r[0][1] * r[2][2]) / det;
	a[1][1] = (r[0][0] * r[2][2] - r[0][2] * r[2][0]) / det;
	a[2][1] = (r[0][1] * r[2][0] - r[0][0] * r[2][1]) / det;
	a[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) / det;
	a[1][2] = (r[0][2] * r[1][0] - r[0][0] * r[1][2]) / det;
	a[2][2] = (r[0][0] * r[1][1] - r[0][1] * r[1][0]) / det;
	for (int i = 0; i < 3; i++) {
		const double b = -((a[i][0] * t[0] + a[i][1] * t[1]) + a[i][2] * t[2]);
		for (int j = 0; j < 3; j++) out[4 * i + j] = float(a[i][j]);
		out[4 * i + 3] = float(b);
	}
	for (int k = 0; k < 12; k++)
		if (!std::isfinite(out[k])) return false;
	return true;
}

// World box of the object box [mn, mx] under (R | t): the 8 corners mapped in double, their
// bounds widened by 1e-5 of the box's largest extent plus 1e-6 of its largest coordinate (so
// float rounding in the traversal never puts a triangle hit outside), rounded outwards to float.
bool worldBox(const float m[12], const float mn[3], const float mx[3], WorldBox& out)
{
	double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
	for (int c = 0; c < 8; c++) {
		const double p[3] = {(c & 1) ? mx[0] : mn[0], (c & 2) ? mx[1] : mn[1],
		                     (c & 4) ? mx[2] : mn[2]};
		for (int i = 0; i < 3; i++) {
			const double w = ((double(m[4 * i]) * p[0] + double(m[4 * i + 1]) * p[1]) +
			                  double(m[4 * i + 2]) * p[2]) + double(m[4 * i + 3]);
			lo[i] = std::min(lo[i], w);
			hi[i] = std::max(hi[i], w);
		}
	}
	double ext = 0.0, mag = 0.0;
	for (int i = 0; i < 3; i++) {
		ext = std::max(ext, hi[i] - lo[i]);
		mag = std::max(mag, std::max(std::fabs(lo[i]), std::fabs(hi[i])));
	}
	const double pad = 1e-5 * ext + 1e-6 * mag;
	for (int i = 0; i < 3; i++) {
		const float a = float(lo[i] - pad), b = float(hi[i] + pad);
		out.mn[i] = double(a) > lo[i] - pad ? std::nextafter(a, -HUGE_VALF) : a;
		out.mx[i] = double(b) < hi[i] + pad ? std::nextafter(b, HUGE_VALF) : b;
		if (!std::isfinite(out.mn[i]) || !std::isfinite(out.mx[i])) return false;
	}
	return true;
}

void setChild(PhxBVHNode& n, int side, const WorldBox& b, int32_t index, int32_t count)
{
	float* f = n.fData + 6 * side;
	for (int k = 0; k < 3; k++) {
		f[k] = b.mn[k];
		f[3 + k] = b.mx[k];
	}
	n.iData[side] = index;
	n.iData[2 + side] = count;
}

WorldBox unite(const std::vector<WorldBox>& boxes, const uint32_t* ids, size_t n)
{
	WorldBox u = boxes[ids[0]];
	for (size_t i = 1; i < n; i++)
		for (int k = 0; k < 3; k++) {
			u.mn[k] = std::min(u.mn[k], boxes[ids[i]].mn[k]);
			u.mx[k] = std::max(u.mx[k], boxes[ids[i]].mx[k]);
		}
	return u;
}

// Top tree: median split of the instances (by box centre along the widest axis of the centres)
// until one instance per leaf. Leaf children hold (~slot, 1), slots in leaf order; a single
// instance becomes both children of the root.
struct TopBuilder {
	const std::vector<WorldBox>& boxes;
	std::vector<PhxBVHNode> nodes;
	std::vector<uint32_t> order;  // slot -> instance

	int32_t child(uint32_t* ids, size_t n)
	{
		if (n == 1) {
			order.push_back(ids[0]);
			return ~int32_t(order.size() - 1);
		}
		return int32_t(build(ids, n));
	}

	uint32_t build(uint32_t* ids, size_t n)
	{
		const uint32_t at = uint32_t(nodes.size());
		nodes.push_back(PhxBVHNode{});
		double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
		for (size_t i = 0; i < n; i++)
			for (int k = 0; k < 3; k++) {
				const double c = 0.5 * (double(boxes[ids[i]].mn[k]) + double(boxes[ids[i]].mx[k]));
				lo[k] = std::min(lo[k], c);
				hi[k] = std::max(hi[k], c);
			}
		int axis = 0;
		for (int k = 1; k < 3; k++)
			if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
		const size_t half = n / 2;
		std::nth_element(ids, ids + half, ids + n, [&](uint32_t a, uint32_t b) {
			const float ca = boxes[a].mn[axis] + boxes[a].mx[axis];
			const float cb = boxes[b].mn[axis] + boxes[b].mx[axis];
			return ca < cb || (ca == cb && a < b);
		});
		PhxBVHNode node{};
		const WorldBox lb = unite(boxes, ids, half), rb = unite(boxes, ids + half, n - half);
		const int32_t l = child(ids, half);
		const int32_t r = child(ids + half, n - half);
		setChild(node, 0, lb, l, half == 1 ? 1 : 0);
		setChild(node, 1, rb, r, n - half == 1 ? 1 : 0);
		nodes[at] = node;
		return at;
	}
};

phx_status checkSet(PhxContext* ctx, const PhxInstanceSet* set)
{
	if (!set || !registry().has(set)) return fail(PHX_ERR_INVALID_HANDLE, "invalid instance set");
	if (set->ctx != ctx) return fail(PHX_ERR_INVALID_HANDLE, "instance set of another context");
	return PHX_OK;
}

void freeSet(PhxInstanceSet* s)
{
	hipFree(s->top);
	hipFree(s->inst);
	hipFree(s->byIndex);
	for (const PhxScene* m : s->meshes) dropScene(m);
	delete s;
}

}  // namespace

namespace phx {

const PhxBVHNode* instanceTop(const PhxInstanceSet* s) { return s->top; }
const DevInstance* instanceArray(const PhxInstanceSet* s) { return s->inst; }
const DevInstance* instanceByIndex(const PhxInstanceSet* s) { return s->byIndex; }
uint32_t instanceCount(const PhxInstanceSet* s) { return s->count; }
uint64_t instanceStamp(const PhxInstanceSet* s) { return s->stamp; }
bool instanceSetValid(PhxContext* ctx, const PhxInstanceSet* s) { return checkSet(ctx, s) == PHX_OK; }

}  // namespace phx

extern "C" {

static phx_status phx_instances_build_impl(PhxContext* ctx, const PhxInstance* instances,
                                           uint32_t count, PhxInstanceSet** outSet)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outSet) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_instances_build: NULL output");
	*outSet = nullptr;
	if (count && !instances) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_instances_build: NULL");
	if (count > (1u << 24))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_instances_build: more than 2^24 instances");
	std::vector<DevInstance> dev(count);
	std::vector<WorldBox> boxes(count);
	for (uint32_t i = 0; i < count; i++) {
		const PhxInstance& in = instances[i];
		st = checkScene(ctx, in.mesh);
		if (st != PHX_OK) return fail(st, "phx_instances_build: instance %u: %s", i, phx_last_error());
		DevInstance& d = dev[i];
		std::memset(&d, 0, sizeof(d));
		if (!invertAffine(in.objectToWorld, d.worldToObject) ||
		    !worldBox(in.objectToWorld, in.mesh->rootMin, in.mesh->rootMax, boxes[i]))
			return fail(PHX_ERR_INVALID_ARGUMENT, "phx_instances_build: instance %u has a singular "
			            "or non-finite transform", i);
		d.nodes = reinterpret_cast<const float4*>(in.mesh->nodes);
		d.tris = reinterpret_cast<const float4*>(in.mesh->tris);
		d.datas = in.mesh->datas;
		d.index = i;
		d.numTris = in.mesh->numTris;
	}
	PhxInstanceSet* s = new (std::nothrow) PhxInstanceSet();
	if (!s) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_instances_build: out of host memory");
	s->ctx = ctx;
	s->count = count;
	s->stamp = nextSceneStamp();
	if (count) {
		TopBuilder tb{boxes, {}, {}};
		std::vector<uint32_t> ids(count);
		std::iota(ids.begin(), ids.end(), 0u);
		if (count == 1) {
			PhxBVHNode root{};
			setChild(root, 0, boxes[0], ~0, 1);
			setChild(root, 1, boxes[0], ~0, 1);
			tb.nodes.push_back(root);
			tb.order.push_back(0);
		} else {
			tb.build(ids.data(), count);
		}
		std::vector<DevInstance> laid(count);
		for (size_t slot = 0; slot < tb.order.size(); slot++) laid[slot] = dev[tb.order[slot]];
		size_t bytes = 0;
		if (hipMalloc(reinterpret_cast<void**>(&s->top), tb.nodes.size() * sizeof(PhxBVHNode)) !=
		        hipSuccess ||
		    hipMalloc(reinterpret_cast<void**>(&s->inst), laid.size() * sizeof(DevInstance)) !=
		        hipSuccess ||
		    hipMalloc(reinterpret_cast<void**>(&s->byIndex), dev.size() * sizeof(DevInstance)) !=
		        hipSuccess) {
			freeSet(s);
			return fail(PHX_ERR_OUT_OF_MEMORY, "phx_instances_build: device allocation");
		}
		bytes = tb.nodes.size() * sizeof(PhxBVHNode);
		hipError_t e = hipMemcpyAsync(s->top, tb.nodes.data(), bytes, hipMemcpyHostToDevice,
		                              ctx->stream);
		if (e == hipSuccess)
			e = hipMemcpyAsync(s->inst, laid.data(), laid.size() * sizeof(DevInstance),
			                   hipMemcpyHostToDevice, ctx->stream);
		if (e == hipSuccess)
			e = hipMemcpyAsync(s->byIndex, dev.data(), dev.size() * sizeof(DevInstance),
			                   hipMemcpyHostToDevice, ctx->stream);
		if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
		if (e != hipSuccess) {
			freeSet(s);
			return hipFail(e, "phx_instances_build");
		}
	}
	for (uint32_t i = 0; i < count; i++)
		if (std::find(s->meshes.begin(), s->meshes.end(), instances[i].mesh) == s->meshes.end()) {
			s->meshes.push_back(instances[i].mesh);
			holdScene(instances[i].mesh);
		}
	holdContext(ctx);
	registry().add(s);
	*outSet = s;
	return ok();
}

phx_status phx_instances_build(PhxContext* ctx, const PhxInstance* instances, uint32_t count,
                               PhxInstanceSet** outSet)
{
	return guarded("phx_instances_build", [&] {
		return phx_instances_build_impl(ctx, instances, count, outSet);
	});
}

void phx_instances_release(PhxInstanceSet* set)
{
	if (!set || !registry().remove(set)) return;
	PhxContext* ctx = set->ctx;
	hipSetDevice(ctx->device);
	hipStreamSynchronize(ctx->stream);
	freeSet(set);
	dropContext(ctx);
}

namespace {

// Static scene (if any) then the instances, seeded with its results, on device buffers.
hipError_t enqueueInstanceCast(PhxContext* ctx, const PhxScene* scene, const PhxInstanceSet* set,
                               const PhxRayIn* rays, uint32_t n, PhxRayHit* hits, uint32_t* ids,
                               uint8_t* inLight)
{
	hipError_t e = hipSuccess;
	if (scene) {
		const TraceLaunch L = traceLaunch(ctx, 0, inLight != nullptr);
		e = inLight ? launchAnyHit(view(scene), rays, n, inLight, L, ctx->stream)
		            : launchClosest(view(scene), rays, n, hits, L, nullptr, ctx->stream);
	}
	if (e == hipSuccess)
		e = launchInstances(set->top, set->inst, set->count, rays, n, hits, ids, inLight,
		                    scene != nullptr, ctx->stream);
	return e;
}

phx_status castInstances(PhxContext* ctx, const PhxScene* scene, const PhxInstanceSet* set,
                         const PhxRayIn* rays, uint32_t n, PhxRayHit* hits, uint32_t* ids,
                         uint8_t* inLight, uint32_t flags, const char* what)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (scene && (st = checkScene(ctx, scene)) != PHX_OK) return st;
	if ((st = checkSet(ctx, set)) != PHX_OK) return st;
	if (flags & ~(PHX_MEM_DEVICE | PHX_ASYNC))
		return fail(PHX_ERR_INVALID_ARGUMENT, "%s: flags 0x%x", what, flags);
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	if ((flags & PHX_ASYNC) && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "%s: PHX_ASYNC needs PHX_MEM_DEVICE", what);
	if (n == 0) return ok();
	if (!rays || (!hits && !inLight)) return fail(PHX_ERR_INVALID_ARGUMENT, "%s: NULL", what);
	if (n > 0x7FFFFFFFu) return fail(PHX_ERR_INVALID_ARGUMENT, "%s: more than 2^31 - 1 rays", what);
	if (device) {
		PHX_HIP(enqueueInstanceCast(ctx, scene, set, rays, n, hits, ids, inLight), what);
		if (!(flags & PHX_ASYNC)) PHX_HIP(hipStreamSynchronize(ctx->stream), what);
		return ok();
	}
	const size_t rb = size_t(n) * sizeof(PhxRayIn);
	const size_t ob = inLight ? size_t(n) : size_t(n) * sizeof(PhxRayHit);
	const size_t ib = (ids && !inLight) ? size_t(n) * 4u : 0u;
	char* tmp = nullptr;
	if (hipMalloc(reinterpret_cast<void**>(&tmp), rb + ob + ib) != hipSuccess)
		return fail(PHX_ERR_OUT_OF_MEMORY, "%s: staging of %zu bytes", what, rb + ob + ib);
	PhxRayIn* dr = reinterpret_cast<PhxRayIn*>(tmp);
	void* dout = tmp + rb;
	uint32_t* dids = ib ? reinterpret_cast<uint32_t*>(tmp + rb + ob) : nullptr;
	hipError_t e = hipMemcpyAsync(dr, rays, rb, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess)
		e = enqueueInstanceCast(ctx, scene, set, dr, n,
		                        inLight ? nullptr : static_cast<PhxRayHit*>(dout), dids,
		                        inLight ? static_cast<uint8_t*>(dout) : nullptr);
	if (e == hipSuccess)
		e = hipMemcpyAsync(inLight ? static_cast<void*>(inLight) : static_cast<void*>(hits), dout,
		                   ob, hipMemcpyDeviceToHost, ctx->stream);
	if (e == hipSuccess && dids)
		e = hipMemcpyAsync(ids, dids, ib, hipMemcpyDeviceToHost, ctx->stream);
	const hipError_t e2 = hipStreamSynchronize(ctx->stream);
	hipFree(tmp);
	if (e != hipSuccess) return hipFail(e, what);
	if (e2 != hipSuccess) return hipFail(e2, what);
	return ok();
}

}  // namespace

phx_status phx_cast_closest_instances(PhxContext* ctx, const PhxScene* scene,
                                      const PhxInstanceSet* set, const PhxRayIn* rays,
                                      uint32_t numRays, PhxRayHit* hits, uint32_t* instanceIds,
                                      uint32_t flags)
{
	if (!hits && numRays) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_instances: NULL");
	return castInstances(ctx, scene, set, rays, numRays, hits, instanceIds, nullptr, flags,
	                     "phx_cast_closest_instances");
}

phx_status phx_cast_anyhit_instances(PhxContext* ctx, const PhxScene* scene,
                                     const PhxInstanceSet* set, const PhxRayIn* rays,
                                     uint32_t numRays, uint8_t* inLight, uint32_t flags)
{
	if (!inLight && numRays) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_anyhit_instances: NULL");
	return castInstances(ctx, scene, set, rays, numRays, nullptr, nullptr, inLight, flags,
	                     "phx_cast_anyhit_instances");
}

}  // extern "C"
