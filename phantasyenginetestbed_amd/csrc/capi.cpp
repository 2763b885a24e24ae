// The C ABI (include/phx.h): contexts, device memory, scene upload, cast-rays and ray generation.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_set>

#include "runtime.hpp"

using namespace phx;

namespace {

phx_status growStage(void*& buf, size_t& have, size_t need)
{
	if (have >= need) return PHX_OK;
	if (buf) hipFree(buf);
	buf = nullptr;
	have = 0;
	size_t sz = std::max<size_t>(need, 1u << 20);
	hipError_t e = hipMalloc(&buf, sz);
	if (e != hipSuccess) {
		buf = nullptr;
		return fail(PHX_ERR_OUT_OF_MEMORY, "staging allocation of %zu bytes failed", sz);
	}
	have = sz;
	return PHX_OK;
}

// Device copy of `count` records followed by `padBytes` zero bytes (kernels that read whole
// 64-B words may touch up to 64 B past the last triangle).
template <class T>
phx_status uploadArray(PhxContext* ctx, T*& dst, const T* src, size_t count, size_t& total,
                       size_t padBytes = 0)
{
	dst = nullptr;
	if (count == 0) return PHX_OK;
	size_t bytes = count * sizeof(T);
	hipError_t e = hipMalloc(reinterpret_cast<void**>(&dst), bytes + padBytes);
	if (e == hipSuccess && padBytes)
		e = hipMemsetAsync(reinterpret_cast<char*>(dst) + bytes, 0, padBytes, ctx->stream);
	if (e != hipSuccess) {
		dst = nullptr;
		return fail(PHX_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes failed", bytes);
	}
	total += bytes;
	e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
	if (e != hipSuccess) return hipFail(e, "hipMemcpyAsync(scene upload)");
	return PHX_OK;
}

void freeScene(PhxScene* s)
{
	hipFree(s->nodes);
	hipFree(s->tris);
	hipFree(s->datas);
	hipFree(s->materials);
	hipFree(s->lights);
	hipFree(s->texArena);
	delete s;
}

void freeContext(PhxContext* ctx)
{
	hipSetDevice(ctx->device);
	hipStreamSynchronize(ctx->stream);
	hipFree(ctx->stageIn);
	hipFree(ctx->stageOut);
	hipFree(ctx->work);
	hipFree(ctx->spill);
	hipStreamDestroy(ctx->stream);
	delete ctx;
}

std::mutex& sceneHoldMutex()
{
	static std::mutex mu;
	return mu;
}

constexpr uint32_t kMaxSceneElements = 1u << 28;

// Runs `launch(raysDev, outDev)` with host/device handling of the ray and output arrays.
template <class Launch>
phx_status castCommon(PhxContext* ctx, const PhxRayIn* rays, uint32_t n, void* out,
                      size_t outBytesPerRay, uint32_t flags, Launch launch, const char* what)
{
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	const bool async = (flags & PHX_ASYNC) != 0;
	if (async && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "%s: PHX_ASYNC needs PHX_MEM_DEVICE", what);
	if (n == 0) return ok();
	if (!rays || !out) return fail(PHX_ERR_INVALID_ARGUMENT, "%s: null ray/output pointer", what);
	if (device) {
		PHX_HIP(launch(rays, out), what);
		if (!async) PHX_HIP(hipStreamSynchronize(ctx->stream), what);
		return ok();
	}
	size_t inBytes = size_t(n) * sizeof(PhxRayIn);
	size_t outBytes = size_t(n) * outBytesPerRay;
	phx_status st = growStage(ctx->stageIn, ctx->stageInBytes, inBytes);
	if (st != PHX_OK) return st;
	st = growStage(ctx->stageOut, ctx->stageOutBytes, outBytes);
	if (st != PHX_OK) return st;
	PHX_HIP(hipMemcpyAsync(ctx->stageIn, rays, inBytes, hipMemcpyHostToDevice, ctx->stream), what);
	PHX_HIP(launch(static_cast<const PhxRayIn*>(ctx->stageIn), ctx->stageOut), what);
	PHX_HIP(hipMemcpyAsync(out, ctx->stageOut, outBytes, hipMemcpyDeviceToHost, ctx->stream), what);
	PHX_HIP(hipStreamSynchronize(ctx->stream), what);
	return ok();
}

// Most node visits of incoherent rays happen in the top levels of the tree, which the persistent
// kernels keep in LDS. The device copy therefore starts with the first kTopCap nodes in
// breadth-first order (the root stays node 0) followed by the remaining nodes in their original
// depth-first order; inner-child indices are remapped, leaf (triangle) indices are unchanged, so
// every traversal decision -- and every hit -- is the same.
constexpr uint32_t kTopCap = 2048;

std::vector<PhxBVHNode> relayoutTop(const PhxBVHNode* nodes, uint32_t n, uint32_t* topCount)
{
	std::vector<uint32_t> bfs;
	bfs.reserve(std::min(n, kTopCap));
	bfs.push_back(0);
	for (size_t h = 0; h < bfs.size() && bfs.size() < kTopCap; h++)
		for (int side = 0; side < 2 && bfs.size() < kTopCap; side++)
			if (nodes[bfs[h]].iData[2 + side] == 0) bfs.push_back(uint32_t(nodes[bfs[h]].iData[side]));
	std::vector<uint32_t> newIndex(n, 0xFFFFFFFFu);
	for (uint32_t i = 0; i < bfs.size(); i++) newIndex[bfs[i]] = i;
	uint32_t next = uint32_t(bfs.size());
	for (uint32_t i = 0; i < n; i++)
		if (newIndex[i] == 0xFFFFFFFFu) newIndex[i] = next++;
	std::vector<PhxBVHNode> out(n);
	for (uint32_t i = 0; i < n; i++) {
		PhxBVHNode nd = nodes[i];
		for (int side = 0; side < 2; side++)
			if (nd.iData[2 + side] == 0) nd.iData[side] = int32_t(newIndex[uint32_t(nd.iData[side])]);
		out[newIndex[i]] = nd;
	}
	*topCount = uint32_t(bfs.size());
	return out;
}

// Triangles as (p0, e1 = p1 - p0, e2 = p2 - p0); p0.w keeps the end marker. The subtractions are
// the IEEE float operations Intersection.hpp:33-34 performs per test, done once here.
std::vector<PhxTriangleVertices> pretransform(const PhxTriangleVertices* t, uint32_t n)
{
	std::vector<PhxTriangleVertices> out(n);
	for (uint32_t i = 0; i < n; i++) {
		for (int k = 0; k < 4; k++) out[i].v0[k] = t[i].v0[k];
		for (int k = 0; k < 3; k++) {
			out[i].v1[k] = t[i].v1[k] - t[i].v0[k];
			out[i].v2[k] = t[i].v2[k] - t[i].v0[k];
		}
		out[i].v1[3] = 0.0f;
		out[i].v2[3] = 0.0f;
	}
	return out;
}

}  // namespace

namespace phx {

uint64_t nextSceneStamp()
{
	static std::atomic<uint64_t> counter{0};
	return ++counter;
}

void holdScene(const PhxScene* s)
{
	std::lock_guard<std::mutex> g(sceneHoldMutex());
	const_cast<PhxScene*>(s)->holders++;
}

void dropScene(const PhxScene* cs)
{
	PhxScene* s = const_cast<PhxScene*>(cs);
	{
		std::lock_guard<std::mutex> g(sceneHoldMutex());
		if (--s->holders != 0 || !s->released) return;
	}
	PhxContext* ctx = s->ctx;
	hipSetDevice(ctx->device);
	hipStreamSynchronize(ctx->stream);
	freeScene(s);
	dropContext(ctx);
}

void holdContext(PhxContext* ctx)
{
	std::lock_guard<std::mutex> g(sceneHoldMutex());
	ctx->holders++;
}

void dropContext(PhxContext* ctx)
{
	{
		std::lock_guard<std::mutex> g(sceneHoldMutex());
		if (--ctx->holders != 0 || !ctx->released) return;
	}
	freeContext(ctx);
}

}  // namespace phx

extern "C" {

const char* phx_version(void) { return "phx 0.1.0 gfx950"; }

phx_status phx_bvh_build(const PhxTriangleVertices* triangles, const PhxTriangleData* datas,
                         uint32_t numTriangles, uint32_t numThreads, PhxBVH** outBvh)
{
	if (!outBvh) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_bvh_build: outBvh is NULL");
	*outBvh = nullptr;
	PhxBVH* b = new (std::nothrow) PhxBVH();
	if (!b) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_bvh_build: out of host memory");
	phx_status st;
	try {
		st = buildSBVH(triangles, datas, numTriangles, numThreads, b->bvh);
	} catch (const std::bad_alloc&) {
		st = fail(PHX_ERR_OUT_OF_MEMORY, "phx_bvh_build: out of host memory");
	}
	if (st != PHX_OK) {
		delete b;
		return st;
	}
	registry().add(b);
	*outBvh = b;
	return ok();
}

phx_status phx_bvh_arrays(const PhxBVH* bvh, const PhxBVHNode** nodes, uint32_t* numNodes,
                          const PhxTriangleVertices** triangles, const PhxTriangleData** datas,
                          uint32_t* numTriangles)
{
	if (!bvh || !registry().has(bvh)) return fail(PHX_ERR_INVALID_HANDLE, "invalid BVH handle");
	if (nodes) *nodes = bvh->bvh.nodes.data();
	if (numNodes) *numNodes = uint32_t(bvh->bvh.nodes.size());
	if (triangles) *triangles = bvh->bvh.triangles.data();
	if (datas) *datas = bvh->bvh.datas.data();
	if (numTriangles) *numTriangles = uint32_t(bvh->bvh.triangles.size());
	return ok();
}

void phx_bvh_release(PhxBVH* bvh)
{
	if (bvh && registry().remove(bvh)) delete bvh;
}

phx_status phx_bvh_validate(const PhxBVHNode* nodes, uint32_t numNodes,
                            const PhxTriangleVertices* triangles, uint32_t numTriangles,
                            uint32_t* outDepth)
{
	return validateBVH(nodes, numNodes, triangles, numTriangles, outDepth);
}

phx_status phx_create(int device, PhxContext** outCtx)
{
	if (!outCtx) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_create: outCtx is NULL");
	*outCtx = nullptr;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count == 0) return fail(PHX_ERR_NO_DEVICE, "phx_create: no HIP device");
	if (device < 0 || device >= count)
		return fail(PHX_ERR_NO_DEVICE, "phx_create: device %d out of range [0,%d)", device, count);
	PHX_HIP(hipSetDevice(device), "hipSetDevice");
	hipDeviceProp_t props;
	PHX_HIP(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties");
	PhxContext* ctx = new (std::nothrow) PhxContext();
	if (!ctx) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_create: out of host memory");
	ctx->device = device;
	ctx->numCUs = uint32_t(props.multiProcessorCount);
	e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete ctx;
		return hipFail(e, "hipStreamCreate");
	}
	// Default shape: one block of 512 threads per CU, no treelet, refill at 8 idle lanes -- the
	// fastest or close to it for incoherent closest-hit and any-hit casts on MI355X;
	// phx_autotune refines it per scene and ray mix.
	uint32_t bpc = 0;
	e = persistentBlocksPerCU(ctx->blockThreads, ctx->topNodes, 0, ctx->ldsDepth, &bpc);
	if (e == hipSuccess && bpc == 0) e = hipErrorInvalidConfiguration;
	ctx->gridClosest = ctx->numCUs;
	ctx->gridAnyhit = ctx->numCUs;
	// The spill area covers every lane the device can hold (2048 per CU).
	ctx->spillLanes = ctx->numCUs * 2048u;
	if (e == hipSuccess) e = hipMalloc(&ctx->work, 64);
	if (e == hipSuccess) e = hipMemset(ctx->work, 0, 64);
	if (e == hipSuccess)
		e = hipMalloc(&ctx->spill, size_t(persistentSpillWords(ctx->spillLanes)) * sizeof(uint32_t));
	if (e != hipSuccess) {
		hipFree(ctx->work);
		hipFree(ctx->spill);
		hipStreamDestroy(ctx->stream);
		delete ctx;
		return hipFail(e, "phx_create: persistent-kernel state");
	}
	// The allocation-free cast's scratch: its address on this device, looked up once.
	if (e == hipSuccess) e = noAllocScratchInit(device);
	if (e != hipSuccess) {
		hipFree(ctx->work);
		hipFree(ctx->spill);
		hipStreamDestroy(ctx->stream);
		delete ctx;
		return hipFail(e, "phx_create: scratch of phx_cast_closest_noalloc");
	}
	registry().add(ctx);
	// The traversal's reciprocal must equal IEEE division on this device (see phx_check_reciprocal).
	uint64_t counts[2] = {0, 0};
	const phx_status st = phx_check_reciprocal(ctx, 0, counts);
	if (st != PHX_OK || counts[1] != 0) {
		registry().remove(ctx);
		hipFree(ctx->work);
		hipFree(ctx->spill);
		hipStreamDestroy(ctx->stream);
		delete ctx;
		if (st != PHX_OK) return st;
		return fail(PHX_ERR_DEVICE,
		            "phx_create: the device's reciprocal differs from IEEE division for %llu of %llu "
		            "inputs; traversal results would not match the reference",
		            (unsigned long long)counts[1], (unsigned long long)counts[0]);
	}
	*outCtx = ctx;
	return ok();
}

phx_status phx_check_reciprocal(PhxContext* ctx, uint32_t flags, uint64_t counts[2])
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!counts) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_check_reciprocal: counts is NULL");
	if (flags & ~PHX_CHECK_EXHAUSTIVE)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_check_reciprocal: unknown flags 0x%x", flags);
	PHX_HIP(hipSetDevice(ctx->device), "hipSetDevice");
	unsigned long long* bad = nullptr;
	PHX_HIP(hipMalloc(&bad, sizeof(*bad)), "hipMalloc");
	uint64_t tested = 0;
	unsigned long long host = 0;
	hipError_t e = hipMemsetAsync(bad, 0, sizeof(*bad), ctx->stream);
	if (e == hipSuccess)
		e = launchCheckReciprocal((flags & PHX_CHECK_EXHAUSTIVE) != 0, bad, &tested, ctx->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(&host, bad, sizeof(host), hipMemcpyDeviceToHost, ctx->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
	hipFree(bad);
	if (e != hipSuccess) return hipFail(e, "phx_check_reciprocal");
	counts[0] = tested;
	counts[1] = host;
	return ok();
}

void phx_destroy(PhxContext* ctx)
{
	if (!ctx || !registry().remove(ctx)) return;
	{
		std::lock_guard<std::mutex> g(sceneHoldMutex());
		if (ctx->holders) {  // freed when the last object made from it is released
			ctx->released = true;
			return;
		}
	}
	freeContext(ctx);
}

void* phx_context_stream(PhxContext* ctx)
{
	if (!ctx || !registry().has(ctx)) return nullptr;
	return ctx->stream;
}

uint32_t phx_context_num_cus(PhxContext* ctx)
{
	if (!ctx || !registry().has(ctx)) return 0;
	return ctx->numCUs;
}

phx_status phx_get_launch_params(PhxContext* ctx, PhxLaunchParams* out)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_get_launch_params: out is NULL");
	out->gridBlocksClosest = ctx->gridClosest;
	out->gridBlocksAnyhit = ctx->gridAnyhit;
	out->refillMin = ctx->refillMin;
	out->blockThreads = ctx->blockThreads;
	out->topNodes = ctx->topNodes;
	uint32_t bpc = 0;
	persistentBlocksPerCU(ctx->blockThreads, ctx->topNodes, ctx->helperWaves, ctx->ldsDepth, &bpc);
	out->residentGridBlocks = ctx->numCUs * bpc;
	out->helperWaves = ctx->helperWaves;
	out->raysPerFetch = ctx->window;
	out->stackLdsEntries = ctx->ldsDepth;
	return ok();
}

phx_status phx_set_launch_params(PhxContext* ctx, const PhxLaunchParams* p)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!p) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: params is NULL");
	const uint32_t block = p->blockThreads ? p->blockThreads : ctx->blockThreads;
	const uint32_t top = p->topNodes != 0xFFFFFFFFu ? p->topNodes : ctx->topNodes;
	const uint32_t helpers = p->helperWaves != 0xFFFFFFFFu ? p->helperWaves : ctx->helperWaves;
	const uint32_t depth = p->stackLdsEntries ? p->stackLdsEntries : ctx->ldsDepth;
	const uint32_t window = p->raysPerFetch ? p->raysPerFetch : ctx->window;
	if (block % 64 != 0 || block > 1024)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: blockThreads %u", block);
	if (top > kTopCap)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: topNodes above %u", kTopCap);
	if (helpers > 16u || block + 64u * helpers > 1024u)
		return fail(PHX_ERR_INVALID_ARGUMENT,
		            "phx_set_launch_params: %u threads and %u helper waves exceed 1024 threads",
		            block, helpers);
	if (!validLdsDepth(depth))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: stackLdsEntries %u outside 4..32",
		            depth);
	if (window < 8u || window > 4096u)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: raysPerFetch %u outside 8..4096",
		            window);
	uint32_t bpc = 0;
	persistentBlocksPerCU(block, top, helpers, depth, &bpc);
	if (bpc == 0)
		return fail(PHX_ERR_INVALID_ARGUMENT,
		            "phx_set_launch_params: %u threads (+%u helper waves) with a %u-node treelet and "
		            "%u LDS stack entries do not fit a CU", block, helpers, top, depth);
	const uint32_t maxGrid = ctx->spillLanes / block;
	if (p->gridBlocksClosest > maxGrid || p->gridBlocksAnyhit > maxGrid)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: grid above %u blocks",
		            maxGrid);
	if (p->refillMin > 64)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_set_launch_params: refillMin above 64");
	// A new block or treelet size without an explicit grid gets the resident grid of that shape.
	const bool reshaped = block != ctx->blockThreads || top != ctx->topNodes ||
	                      helpers != ctx->helperWaves || depth != ctx->ldsDepth;
	const uint32_t resident = std::min(ctx->numCUs * bpc, maxGrid);
	ctx->blockThreads = block;
	ctx->topNodes = top;
	ctx->helperWaves = helpers;
	ctx->ldsDepth = depth;
	ctx->window = window;
	if (p->gridBlocksClosest)
		ctx->gridClosest = p->gridBlocksClosest;
	else if (reshaped)
		ctx->gridClosest = resident;
	if (p->gridBlocksAnyhit)
		ctx->gridAnyhit = p->gridBlocksAnyhit;
	else if (reshaped)
		ctx->gridAnyhit = resident;
	if (p->refillMin) ctx->refillMin = p->refillMin;
	return ok();
}

phx_status phx_device_alloc(PhxContext* ctx, size_t bytes, void** outPtr)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outPtr || bytes == 0) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_device_alloc: bad args");
	*outPtr = nullptr;
	hipError_t e = hipMalloc(outPtr, bytes);
	if (e != hipSuccess) {
		*outPtr = nullptr;
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_device_alloc(%zu): %s", bytes, hipGetErrorString(e));
	}
	return ok();
}

phx_status phx_device_free(PhxContext* ctx, void* ptr)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!ptr) return ok();
	PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_device_free");
	PHX_HIP(hipFree(ptr), "hipFree");
	return ok();
}

phx_status phx_copy_to_device(PhxContext* ctx, void* dst, const void* src, size_t bytes)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (bytes == 0) return ok();
	if (!dst || !src) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_copy_to_device: null pointer");
	PHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "copy H2D");
	PHX_HIP(hipStreamSynchronize(ctx->stream), "copy H2D");
	return ok();
}

phx_status phx_copy_to_host(PhxContext* ctx, void* dst, const void* src, size_t bytes)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (bytes == 0) return ok();
	if (!dst || !src) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_copy_to_host: null pointer");
	PHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy D2H");
	PHX_HIP(hipStreamSynchronize(ctx->stream), "copy D2H");
	return ok();
}

phx_status phx_copy_on_device(PhxContext* ctx, void* dst, const void* src, size_t bytes)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (bytes == 0) return ok();
	if (!dst || !src) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_copy_on_device: null pointer");
	PHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "copy D2D");
	PHX_HIP(hipStreamSynchronize(ctx->stream), "copy D2D");
	return ok();
}

phx_status phx_synchronize(PhxContext* ctx)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_synchronize");
	return ok();
}

static phx_status phx_scene_upload_impl(PhxContext* ctx, const PhxBVHNode* nodes,
                                        uint32_t numNodes, const PhxTriangleVertices* triangles,
                                        const PhxTriangleData* datas, uint32_t numTriangles,
                                        const PhxMaterial* materials, uint32_t numMaterials,
                                        const PhxSphereLight* lights, uint32_t numLights,
                                        PhxScene** outScene)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outScene) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_upload: outScene is NULL");
	*outScene = nullptr;
	if (numNodes >= kMaxSceneElements || numTriangles >= kMaxSceneElements)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_upload: more than 2^28 nodes/triangles");
	if ((numMaterials && !materials) || (numLights && !lights))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_upload: null material/light array");
	uint32_t depth = 0;
	st = validateBVH(nodes, numNodes, triangles, numTriangles, &depth);
	if (st != PHX_OK) return st;

	PhxScene* s = new (std::nothrow) PhxScene();
	if (!s) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_scene_upload: out of host memory");
	s->ctx = ctx;
	s->numNodes = numNodes;
	s->numTris = numTriangles;
	s->numMaterials = numMaterials;
	s->numLights = numLights;
	s->depth = depth;
	s->finiteBounds = 1;
	for (int k = 0; k < 3; k++) {  // root box: union of node 0's child boxes (fData layout)
		const float* f = nodes[0].fData;
		const float lmn[3] = {f[0], f[1], f[2]}, lmx[3] = {f[3], f[4], f[5]};
		const float rmn[3] = {f[6], f[7], f[8]}, rmx[3] = {f[9], f[10], f[11]};
		s->rootMin[k] = std::min(lmn[k], rmn[k]);
		s->rootMax[k] = std::max(lmx[k], rmx[k]);
	}
	// Bounds within +-2^126 (false for inf and NaN) keep every b - o of a ray whose origin is in the
	// same range finite: the condition of the kernels' NaN-free slab test.
	for (uint32_t i = 0; i < numNodes && s->finiteBounds; i++)
		for (int k = 0; k < 12; k++)
			if (!(std::fabs(nodes[i].fData[k]) <= 0x1p126f)) {
				s->finiteBounds = 0;
				break;
			}
	{
		std::vector<PhxBVHNode> laid = relayoutTop(nodes, numNodes, &s->topCount);
		st = uploadArray(ctx, s->nodes, laid.data(), numNodes, s->bytes);
		if (st == PHX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
			st = fail(PHX_ERR_HIP, "phx_scene_upload: node upload failed");
	}
	if (st == PHX_OK) {
		std::vector<PhxTriangleVertices> pre = pretransform(triangles, numTriangles);
		st = uploadArray(ctx, s->tris, pre.data(), numTriangles, s->bytes, 64);
		if (st == PHX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
			st = fail(PHX_ERR_HIP, "phx_scene_upload: triangle upload failed");
	}
	if (st == PHX_OK && datas) st = uploadArray(ctx, s->datas, datas, numTriangles, s->bytes);
	if (st == PHX_OK) st = uploadArray(ctx, s->materials, materials, numMaterials, s->bytes);
	if (st == PHX_OK) st = uploadArray(ctx, s->lights, lights, numLights, s->bytes);
	if (st == PHX_OK) {
		hipError_t e = hipStreamSynchronize(ctx->stream);
		if (e != hipSuccess) st = hipFail(e, "phx_scene_upload");
	}
	if (st != PHX_OK) {
		hipStreamSynchronize(ctx->stream);
		freeScene(s);
		return st;
	}
	s->version = nextSceneStamp();
	holdContext(ctx);
	registry().add(s);
	*outScene = s;
	return ok();
}

phx_status phx_scene_upload(PhxContext* ctx, const PhxBVHNode* nodes, uint32_t numNodes,
                            const PhxTriangleVertices* triangles, const PhxTriangleData* datas,
                            uint32_t numTriangles, const PhxMaterial* materials,
                            uint32_t numMaterials, const PhxSphereLight* lights,
                            uint32_t numLights, PhxScene** outScene)
{
	return guarded("phx_scene_upload", [&] {
		return phx_scene_upload_impl(ctx, nodes, numNodes, triangles, datas, numTriangles,
		                             materials, numMaterials, lights, numLights, outScene);
	});
}

void phx_scene_release(PhxScene* scene)
{
	if (!scene || !registry().remove(scene)) return;
	{
		std::lock_guard<std::mutex> g(sceneHoldMutex());
		if (scene->holders) {  // freed by the last plan or server that holds it
			scene->released = true;
			return;
		}
	}
	PhxContext* ctx = scene->ctx;
	hipSetDevice(ctx->device);
	hipStreamSynchronize(ctx->stream);
	freeScene(scene);
	dropContext(ctx);
}


size_t phx_scene_device_bytes(const PhxScene* scene)
{
	if (!scene || !registry().has(scene)) return 0;
	return scene->bytes;
}

static phx_status phx_scene_set_textures_impl(PhxContext* ctx, PhxScene* scene,
                                              const PhxTexture* textures, uint32_t count)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (count && !textures)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_set_textures: textures is NULL");
	if (count > (1u << 20))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_set_textures: more than 2^20 textures");
	// Arena: linearisation table (256 floats), descriptors, then each image's packed rows.
	const size_t tableBytes = 256 * sizeof(float);
	const size_t descBytes = size_t(count) * sizeof(TexDesc);
	std::vector<TexDesc> descs(count);
	size_t texelBytes = 0;
	for (uint32_t i = 0; i < count; i++) {
		const PhxTexture& t = textures[i];
		const uint32_t bpp = t.bytesPerPixel;
		if (!t.data || t.width == 0 || t.height == 0)
			return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_set_textures: texture %u is empty", i);
		if (bpp != 1 && bpp != 2 && bpp != 4)
			return fail(PHX_ERR_INVALID_ARGUMENT,
			            "phx_scene_set_textures: texture %u has %u bytes per texel (1, 2 or 4)", i,
			            bpp);
		// Texel coordinates are computed in float: dimensions up to 2^16 stay exact with room.
		if (t.width > 65536u || t.height > 65536u)
			return fail(PHX_ERR_INVALID_ARGUMENT,
			            "phx_scene_set_textures: texture %u is larger than 65536 texels a side", i);
		if (t.pitch < t.width * bpp)
			return fail(PHX_ERR_INVALID_ARGUMENT,
			            "phx_scene_set_textures: texture %u pitch %u < width * bytes per texel", i,
			            t.pitch);
		const uint32_t rowBytes = t.width * bpp;
		descs[i] = TexDesc{uint64_t(tableBytes + descBytes + texelBytes), t.width, t.height,
		                   rowBytes, bpp, {0u, 0u}};
		texelBytes += (size_t(rowBytes) * t.height + 15u) & ~size_t(15);
	}
	void* arena = nullptr;
	const size_t total = tableBytes + descBytes + texelBytes;
	if (count) {
		// Texel offsets are relative to the arena base (the `texels` pointer of the set).
		std::vector<uint8_t> host(total);
		float* table = reinterpret_cast<float*>(host.data());
		for (uint32_t c = 0; c < 256; c++) {
			// powf(c / 255.0f, 2.2f) (InterpretRayHitKernel.cu:39-45), evaluated in double and
			// rounded once to float.
			const float x = float(c) / 255.0f;
			table[c] = float(std::pow(double(x), double(2.2f)));
		}
		std::memcpy(host.data() + tableBytes, descs.data(), descBytes);
		for (uint32_t i = 0; i < count; i++) {
			const PhxTexture& t = textures[i];
			const uint32_t rowBytes = descs[i].pitch;
			for (uint32_t y = 0; y < t.height; y++)
				std::memcpy(host.data() + descs[i].offset + size_t(y) * rowBytes,
				            t.data + size_t(y) * t.pitch, rowBytes);
		}
		if (hipMalloc(&arena, total) != hipSuccess)
			return fail(PHX_ERR_OUT_OF_MEMORY, "phx_scene_set_textures: device allocation of %zu "
			            "bytes failed", total);
		hipError_t e = hipMemcpy(arena, host.data(), total, hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			hipFree(arena);
			return hipFail(e, "phx_scene_set_textures");
		}
	}
	// Work queued on the context stream may still read the old arena.
	hipError_t e = hipStreamSynchronize(ctx->stream);
	if (e != hipSuccess) {
		hipFree(arena);
		return hipFail(e, "phx_scene_set_textures");
	}
	hipFree(scene->texArena);
	scene->bytes -= scene->texBytes;
	scene->texArena = arena;
	scene->texBytes = count ? total : 0;
	scene->bytes += scene->texBytes;
	const uint8_t* base = static_cast<const uint8_t*>(arena);
	scene->tex.linear = count ? reinterpret_cast<const float*>(base) : nullptr;
	scene->tex.descs = count ? reinterpret_cast<const TexDesc*>(base + tableBytes) : nullptr;
	scene->tex.texels = count ? base : nullptr;
	scene->tex.count = count;
	scene->version = nextSceneStamp();
	return ok();
}

phx_status phx_scene_set_textures(PhxContext* ctx, PhxScene* scene, const PhxTexture* textures,
                                  uint32_t count)
{
	return guarded("phx_scene_set_textures", [&] {
		return phx_scene_set_textures_impl(ctx, scene, textures, count);
	});
}

phx_status phx_cast_closest(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                            uint32_t numRays, PhxRayHit* hits, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (flags & PHX_TRAVERSAL_MASK)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest: reserved flag bits 0x%x",
		            flags & PHX_TRAVERSAL_MASK);
	if ((flags & PHX_KERNEL_ONE_RAY_PER_LANE) && (flags & PHX_KERNEL_PACKET))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest: two launch models requested");
	SceneView sv = view(scene);
	hipStream_t stream = ctx->stream;
	const TraceLaunch L = traceLaunch(ctx, flags, false);
	return castCommon(
		ctx, rays, numRays, hits, sizeof(PhxRayHit), flags,
		[&](const PhxRayIn* r, void* out) {
			return launchClosest(sv, r, numRays, static_cast<PhxRayHit*>(out), L, nullptr, stream);
		},
		"phx_cast_closest");
}

phx_status phx_cast_closest_timeline(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                                     uint32_t numRays, PhxRayHit* hits, float* msSinceStart,
                                     uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (flags != PHX_MEM_DEVICE)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_timeline: flags 0x%x (device rays "
		            "and hits, PHX_MEM_DEVICE only)", flags);
	if (numRays > 0x7FFFFFFFu)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_timeline: %u rays (at most "
		            "2^31 - 1)", numRays);
	if (numRays == 0) return ok();
	if (!rays || !hits || !msSinceStart)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_timeline: NULL");
	int khz = 0;
	PHX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device),
	        "phx_cast_closest_timeline: clock rate");
	if (khz <= 0) return fail(PHX_ERR_DEVICE, "phx_cast_closest_timeline: no device clock rate");
	const size_t words = size_t(numRays) + 1u;
	std::unique_ptr<unsigned long long[]> host(new (std::nothrow) unsigned long long[words]);
	if (!host) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_cast_closest_timeline: host memory");
	unsigned long long* clock = nullptr;
	if (hipMalloc(reinterpret_cast<void**>(&clock), words * 8u) != hipSuccess)
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_cast_closest_timeline: %zu bytes", words * 8u);
	const TraceLaunch L = traceLaunch(ctx, 0, false);
	hipError_t e = hipMemsetAsync(clock, 0xFF, words * 8u, ctx->stream);
	if (e == hipSuccess) e = launchClosestClock(view(scene), rays, numRays, hits, L, clock, ctx->stream);
	if (e == hipSuccess)
		e = hipMemcpyAsync(host.get(), clock, words * 8u, hipMemcpyDeviceToHost, ctx->stream);
	const hipError_t e2 = hipStreamSynchronize(ctx->stream);
	hipFree(clock);
	if (e != hipSuccess) return hipFail(e, "phx_cast_closest_timeline");
	if (e2 != hipSuccess) return hipFail(e2, "phx_cast_closest_timeline");
	const unsigned long long start = host[numRays];
	for (uint32_t i = 0; i < numRays; i++)
		msSinceStart[i] = host[i] == ~0ull ? -1.0f : float(double(host[i] - start) / double(khz));
	return ok();
}

phx_status phx_cast_closest_noalloc(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                                    uint32_t numRays, PhxRayHit* hits, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!(flags & PHX_MEM_DEVICE) || (flags & ~(PHX_MEM_DEVICE | PHX_ASYNC)))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_noalloc: flags 0x%x (device "
		            "pointers required; only PHX_MEM_DEVICE and PHX_ASYNC allowed)", flags);
	if (numRays > 0x7FFFFFFFu)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_noalloc: %u rays (at most "
		            "2^31 - 1)", numRays);
	if (numRays == 0) return ok();
	if (!rays || !hits) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_noalloc: NULL");
	char* base = static_cast<char*>(noAllocScratch(ctx->device));
	if (!base) return fail(PHX_ERR_HIP, "phx_cast_closest_noalloc: scratch not set up");
	// The context's shape, with the largest LDS stack split that fits (the fewest spilled
	// entries) and the grid capped so its spill columns fit the scratch.
	TraceLaunch L = traceLaunch(ctx, 0, false);
	uint32_t bpc = 0;
	if (persistentBlocksPerCU(L.blockThreads, L.topNodes, L.helperWaves, 32, &bpc) ==
	        hipSuccess && bpc > 0)
		L.ldsDepth = 32;
	const uint32_t perLane = kPersistentStackEntries - L.ldsDepth;
	const uint32_t maxGrid = kNoAllocSpillWords / (perLane * L.blockThreads);
	if (maxGrid == 0)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_closest_noalloc: a block of %u threads "
		            "needs more stack spill than the %u-word scratch holds", L.blockThreads,
		            kNoAllocSpillWords);
	L.gridBlocks = std::min(L.gridBlocks, maxGrid);
	L.spillLanes = L.gridBlocks * L.blockThreads;
	L.work = reinterpret_cast<uint32_t*>(base);
	L.spill = reinterpret_cast<uint32_t*>(base + kNoAllocCounterBytes);
	PHX_HIP(launchZeroWords(L.work, 2u, ctx->stream), "phx_cast_closest_noalloc");
	PHX_HIP(launchClosest(view(scene), rays, numRays, hits, L, nullptr, ctx->stream),
	        "phx_cast_closest_noalloc");
	if (!(flags & PHX_ASYNC))
		PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_cast_closest_noalloc");
	return ok();
}

phx_status phx_cast_anyhit(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                           uint32_t numRays, uint8_t* inLight, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if ((flags & PHX_KERNEL_ONE_RAY_PER_LANE) && (flags & PHX_KERNEL_PACKET))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_cast_anyhit: two launch models requested");
	SceneView sv = view(scene);
	hipStream_t stream = ctx->stream;
	const TraceLaunch L = traceLaunch(ctx, flags, true);
	return castCommon(
		ctx, rays, numRays, inLight, 1, flags,
		[&](const PhxRayIn* r, void* out) {
			return launchAnyHit(sv, r, numRays, static_cast<uint8_t*>(out), L, stream);
		},
		"phx_cast_anyhit");
}

phx_status phx_traversal_stats(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                               uint32_t numRays, uint32_t flags, uint64_t counters[5])
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!counters) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_traversal_stats: counters is NULL");
	if (flags & PHX_TRAVERSAL_MASK)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_traversal_stats: reserved flag bits");
	if (flags & PHX_ASYNC) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_traversal_stats: no PHX_ASYNC");
	const bool anyhit = (flags & PHX_STATS_ANYHIT) != 0;
	for (int k = 0; k < 5; k++) counters[k] = 0;
	if (numRays == 0) return ok();
	unsigned long long* dCounters = nullptr;
	PHX_HIP(hipMalloc(&dCounters, 5 * sizeof(unsigned long long)), "phx_traversal_stats");
	hipError_t e = hipMemsetAsync(dCounters, 0, 5 * sizeof(unsigned long long), ctx->stream);
	PhxRayHit* dHits = nullptr;
	if (e == hipSuccess) e = hipMalloc(&dHits, size_t(numRays) * sizeof(PhxRayHit));
	SceneView sv = view(scene);
	const TraceLaunch L = traceLaunch(ctx, flags & ~PHX_STATS_ANYHIT, false);
	if (e == hipSuccess) {
		st = castCommon(
			ctx, rays, numRays, dHits, sizeof(PhxRayHit), (flags & PHX_MEM_DEVICE),
			[&](const PhxRayIn* r, void*) {
				return anyhit ? launchAnyHitCounted(sv, r, numRays, reinterpret_cast<uint8_t*>(dHits),
				                                    dCounters, ctx->stream)
				              : launchClosest(sv, r, numRays, dHits, L, dCounters, ctx->stream);
			},
			"phx_traversal_stats");
	}
	unsigned long long host[5] = {0, 0, 0, 0, 0};
	if (e == hipSuccess && st == PHX_OK)
		e = hipMemcpy(host, dCounters, sizeof(host), hipMemcpyDeviceToHost);
	hipFree(dHits);
	hipFree(dCounters);
	if (e != hipSuccess) return hipFail(e, "phx_traversal_stats");
	if (st != PHX_OK) return st;
	for (int k = 0; k < 5; k++) counters[k] = host[k];
	return ok();
}

phx_status phx_gen_bounce_rays(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                               const PhxRayHit* hits, uint32_t numRays, uint64_t seed,
                               uint32_t frame, PhxRayIn* outRays, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!scene->datas)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_bounce_rays: scene has no TriangleData");
	if (numRays == 0) return ok();
	if (!rays || !hits || !outRays)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_bounce_rays: null pointer");
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	const bool async = (flags & PHX_ASYNC) != 0;
	if (async && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_bounce_rays: PHX_ASYNC needs PHX_MEM_DEVICE");
	if (device) {
		PHX_HIP(launchGenBounceRays(rays, hits, scene->datas, scene->numTris, numRays, seed, frame,
		                            outRays, ctx->stream),
		        "phx_gen_bounce_rays");
		if (!async) PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_gen_bounce_rays");
		return ok();
	}
	const size_t rb = size_t(numRays) * sizeof(PhxRayIn), hb = size_t(numRays) * sizeof(PhxRayHit);
	void* tmp = nullptr;
	PHX_HIP(hipMalloc(&tmp, 2 * rb + hb), "phx_gen_bounce_rays");
	PhxRayIn* dr = static_cast<PhxRayIn*>(tmp);
	PhxRayIn* dout = dr + numRays;
	PhxRayHit* dh = reinterpret_cast<PhxRayHit*>(dout + numRays);
	hipError_t e = hipMemcpyAsync(dr, rays, rb, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(dh, hits, hb, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess)
		e = launchGenBounceRays(dr, dh, scene->datas, scene->numTris, numRays, seed, frame, dout,
		                        ctx->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(outRays, dout, rb, hipMemcpyDeviceToHost, ctx->stream);
	hipError_t e2 = hipStreamSynchronize(ctx->stream);
	hipFree(tmp);
	if (e != hipSuccess) return hipFail(e, "phx_gen_bounce_rays");
	if (e2 != hipSuccess) return hipFail(e2, "phx_gen_bounce_rays");
	return ok();
}

phx_status phx_interpret_hits(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                              const PhxRayHit* hits, uint32_t numRays, PhxRayHitInfo* out,
                              uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!scene->datas)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_interpret_hits: scene has no TriangleData");
	if (numRays == 0) return ok();
	if (!rays || !hits || !out)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_interpret_hits: null pointer");
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	const bool async = (flags & PHX_ASYNC) != 0;
	if (async && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_interpret_hits: PHX_ASYNC needs PHX_MEM_DEVICE");
	if (device) {
		PHX_HIP(launchInterpret(rays, hits, numRays, scene->datas, scene->numTris,
		                        scene->materials, scene->numMaterials, scene->tex,
		                        reinterpret_cast<float4*>(out), ctx->stream),
		        "phx_interpret_hits");
		if (!async) PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_interpret_hits");
		return ok();
	}
	const size_t rb = size_t(numRays) * sizeof(PhxRayIn), hb = size_t(numRays) * sizeof(PhxRayHit);
	const size_t ob = size_t(numRays) * sizeof(PhxRayHitInfo);
	void* tmp = nullptr;
	PHX_HIP(hipMalloc(&tmp, rb + hb + ob), "phx_interpret_hits");
	PhxRayIn* dr = static_cast<PhxRayIn*>(tmp);
	PhxRayHit* dh = reinterpret_cast<PhxRayHit*>(dr + numRays);
	float4* dout = reinterpret_cast<float4*>(dh + numRays);
	hipError_t e = hipMemcpyAsync(dr, rays, rb, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(dh, hits, hb, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess)
		e = launchInterpret(dr, dh, numRays, scene->datas, scene->numTris, scene->materials,
		                    scene->numMaterials, scene->tex, dout, ctx->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream);
	hipError_t e2 = hipStreamSynchronize(ctx->stream);
	hipFree(tmp);
	if (e != hipSuccess) return hipFail(e, "phx_interpret_hits");
	if (e2 != hipSuccess) return hipFail(e2, "phx_interpret_hits");
	return ok();
}

static phx_status phx_plan_create_impl(PhxContext* ctx, const PhxBatch* batches,
                                       uint32_t numBatches, uint32_t flags, PhxPlan** outPlan)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outPlan) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_create: outPlan is NULL");
	*outPlan = nullptr;
	if (!batches || numBatches == 0 || numBatches > kMaxFusedBatches)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_create: 1..%u batches", kMaxFusedBatches);
	std::vector<BatchDesc> descs(numBatches);
	std::vector<uint32_t> offsets(numBatches + 1, 0);
	uint64_t total = 0;
	const PhxScene* firstScene = batches[0].scene;
	bool oneScene = true;
	for (uint32_t b = 0; b < numBatches; b++) {
		const PhxBatch& in = batches[b];
		st = checkScene(ctx, in.scene);
		if (st != PHX_OK) return st;
		if (in.kind != PHX_BATCH_CLOSEST && in.kind != PHX_BATCH_ANYHIT)
			return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_create: batch %u kind %u", b, in.kind);
		if (in.numRays && (!in.rays || !in.out))
			return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_create: batch %u null pointer", b);
		descs[b] = BatchDesc{reinterpret_cast<const float4*>(in.scene->nodes),
		                     reinterpret_cast<const float4*>(in.scene->tris),
		                     reinterpret_cast<const float4*>(in.rays), in.out, in.numRays, in.kind,
		                     in.scene->finiteBounds, 0u};
		offsets[b] = uint32_t(total);
		total += in.numRays;
		if (total > 0x7FFFFFFFull)
			return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_create: more than 2^31 rays in a plan");
		oneScene = oneScene && in.scene == firstScene;
	}
	offsets[numBatches] = uint32_t(total);
	PhxPlan* p = new (std::nothrow) PhxPlan();
	if (!p) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_plan_create: out of host memory");
	p->ctx = ctx;
	p->numBatches = numBatches;
	p->total = uint32_t(total);
	// A fused launch is throughput bound: default to the resident grid of the fused kernel without
	// a treelet or helper waves.
	p->launch = traceLaunch(ctx, 0, false);
	p->launch.topNodes = 0;
	uint32_t bpc = 0;
	p->launch.helperWaves = 0;
	batchBlocksPerCU(p->launch.blockThreads, 0, 0, p->launch.ldsDepth, numBatches,
	                 (flags & PHX_PLAN_TRACK_COMPLETION) != 0, &bpc);
	if (bpc == 0) {  // the context's stack split does not fit this kernel: use the default split
		p->launch.ldsDepth = 16;
		batchBlocksPerCU(p->launch.blockThreads, 0, 0, 16, numBatches,
		                 (flags & PHX_PLAN_TRACK_COMPLETION) != 0, &bpc);
	}
	p->launch.gridBlocks = std::max(1u, std::min(ctx->numCUs * std::max(bpc, 1u),
	                                             ctx->spillLanes / p->launch.blockThreads));
	p->hostDescs = descs;
	if (oneScene) {
		p->treeletNodes = reinterpret_cast<const float4*>(firstScene->nodes);
		p->treeletCount = firstScene->topCount;
	}
	size_t dummy = 0;
	st = uploadArray(ctx, p->descs, descs.data(), numBatches, dummy);
	if (st == PHX_OK) st = uploadArray(ctx, p->offsets, offsets.data(), numBatches + 1, dummy);
	if (st == PHX_OK && (flags & PHX_PLAN_TRACK_COMPLETION)) {
		BatchClock init;
		std::memset(&init, 0, sizeof(init));
		init.start = ~0ull;
		st = uploadArray(ctx, p->clock, &init, 1, dummy);
	}
	if (st == PHX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
		st = fail(PHX_ERR_HIP, "phx_plan_create: upload failed");
	if (st != PHX_OK) {
		hipFree(p->descs);
		hipFree(p->offsets);
		hipFree(p->clock);
		delete p;
		return st;
	}
	for (uint32_t b = 0; b < numBatches; b++)
		if (std::find(p->scenes.begin(), p->scenes.end(), batches[b].scene) == p->scenes.end()) {
			p->scenes.push_back(batches[b].scene);
			holdScene(batches[b].scene);
		}
	holdContext(ctx);
	registry().add(p);
	*outPlan = p;
	return ok();
}

phx_status phx_plan_create(PhxContext* ctx, const PhxBatch* batches, uint32_t numBatches,
                           uint32_t flags, PhxPlan** outPlan)
{
	return guarded("phx_plan_create", [&] {
		return phx_plan_create_impl(ctx, batches, numBatches, flags, outPlan);
	});
}

phx_status phx_plan_launch(PhxContext* ctx, PhxPlan* plan, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!plan || !registry().has(plan) || plan->ctx != ctx)
		return fail(PHX_ERR_INVALID_HANDLE, "invalid plan handle");
	const TraceLaunch L = plan->launch;
	auto launch = [&]() {
		return launchBatches(plan->descs, plan->offsets, plan->numBatches, plan->total, L,
		                     plan->treeletNodes, plan->treeletCount, plan->clock, ctx->stream);
	};
	if (flags & PHX_PLAN_USE_GRAPH) {
		if (!plan->graphExec) {
			PHX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
			        "phx_plan_launch: begin capture");
			hipError_t e = launch();
			hipGraph_t g = nullptr;
			hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
			if (e != hipSuccess) return hipFail(e, "phx_plan_launch: captured launch");
			if (e2 != hipSuccess) return hipFail(e2, "phx_plan_launch: end capture");
			plan->graph = g;
			PHX_HIP(hipGraphInstantiate(&plan->graphExec, g, nullptr, nullptr, 0),
			        "phx_plan_launch: instantiate");
		}
		PHX_HIP(hipGraphLaunch(plan->graphExec, ctx->stream), "phx_plan_launch: graph launch");
	} else {
		PHX_HIP(launch(), "phx_plan_launch");
	}
	if (!(flags & PHX_ASYNC)) PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_plan_launch");
	return ok();
}

phx_status phx_plan_batch_times(PhxContext* ctx, PhxPlan* plan, float* msSinceStart)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!plan || !registry().has(plan) || plan->ctx != ctx)
		return fail(PHX_ERR_INVALID_HANDLE, "invalid plan handle");
	if (!plan->clock)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_batch_times: plan not tracked");
	if (!msSinceStart) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_batch_times: NULL output");
	BatchClock host;
	PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_plan_batch_times");
	PHX_HIP(hipMemcpy(&host, plan->clock, sizeof(host), hipMemcpyDeviceToHost),
	        "phx_plan_batch_times");
	int rateKHz = 0;
	PHX_HIP(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, ctx->device),
	        "phx_plan_batch_times");
	const double ticksPerMs = double(rateKHz);
	for (uint32_t b = 0; b < plan->numBatches; b++) {
		if (plan->hostDescs[b].count == 0) {
			msSinceStart[b] = 0.0f;
			continue;
		}
		msSinceStart[b] = float(double(host.finish[b] - host.lastStart) / ticksPerMs);
	}
	return ok();
}

}  // extern "C"

namespace {

// Checks a launch shape against the device (block size, LDS, spill area); fills `L`.
// Blocks per CU of the kernel a shape launches: the persistent kernel for single casts, the fused
// kernel -- with the plan's batch count and completion tracking -- for a plan.
uint32_t blocksPerCU(const PhxPlan* plan, uint32_t block, uint32_t top, uint32_t helpers,
                     uint32_t depth)
{
	uint32_t bpc = 0;
	if (plan)
		batchBlocksPerCU(block, std::min(top, plan->treeletCount), helpers, depth, plan->numBatches,
		                 plan->clock != nullptr, &bpc);
	else
		persistentBlocksPerCU(block, top, helpers, depth, &bpc);
	return bpc;
}

phx_status shapeFor(PhxContext* ctx, const PhxLaunchParams& p, TraceLaunch& L,
                    const PhxPlan* plan)
{
	const uint32_t block = p.blockThreads;
	const uint32_t helpers = p.helperWaves == 0xFFFFFFFFu ? 0u : p.helperWaves;
	const uint32_t depth = p.stackLdsEntries ? p.stackLdsEntries : 16u;
	const uint32_t window = p.raysPerFetch ? p.raysPerFetch : 64u;
	if (block == 0 || block % 64 != 0 || block > 1024)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: blockThreads %u", block);
	if (p.topNodes > kTopCap)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: topNodes above %u", kTopCap);
	if (helpers > 16u || block + 64u * helpers > 1024u)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: %u threads and %u helper waves",
		            block, helpers);
	if (!validLdsDepth(depth))
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: stackLdsEntries %u outside 4..32", depth);
	if (window < 8u || window > 4096u)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: raysPerFetch %u outside 8..4096", window);
	if (blocksPerCU(plan, block, p.topNodes, helpers, depth) == 0)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: shape does not fit a CU");
	const uint32_t maxGrid = ctx->spillLanes / block;
	const uint32_t grid = p.gridBlocksClosest;
	if (grid == 0 || grid > maxGrid)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: grid %u outside 1..%u", grid, maxGrid);
	if (p.refillMin == 0 || p.refillMin > 64)
		return fail(PHX_ERR_INVALID_ARGUMENT, "launch params: refillMin %u", p.refillMin);
	L.blockThreads = block;
	L.topNodes = p.topNodes;
	L.helperWaves = helpers;
	L.ldsDepth = depth;
	L.window = window;
	L.gridBlocks = grid;
	L.refillMin = p.refillMin;
	L.spillLanes = ctx->spillLanes;
	L.work = ctx->work;
	L.spill = ctx->spill;
	return PHX_OK;
}

struct Candidate {
	TraceLaunch launch;
	PhxLaunchParams params;
	float ms = 0.0f;
};

// Candidate shapes: block 256/512, treelet 0/512/1024 nodes, a quarter, half or all of the
// resident grid, refill threshold 8 or 32, and for single batches 0 or 4 helper waves -- plus the
// one-ray-per-lane kernel for single batches.
std::vector<Candidate> candidates(PhxContext* ctx, const PhxPlan* plan)
{
	std::vector<Candidate> out;
	for (uint32_t helpers : {0u, 4u}) {
		if (plan && helpers) continue;
		for (uint32_t block : {256u, 512u})
			for (uint32_t top : {0u, 512u, 1024u}) {
				const uint32_t bpc = blocksPerCU(plan, block, top, helpers, 16u);
				if (bpc == 0) continue;
				const uint32_t resident = std::min(ctx->numCUs * bpc, ctx->spillLanes / block);
				for (uint32_t div : {4u, 2u, 1u})
					for (uint32_t refill : {8u, 32u}) {
						Candidate c;
						c.params = PhxLaunchParams{std::max(1u, resident / div),
						                           std::max(1u, resident / div), refill, block, top,
						                           resident, helpers, 64u, 16u};
						if (shapeFor(ctx, c.params, c.launch, plan) == PHX_OK) out.push_back(c);
					}
			}
	}
	if (!plan) {
		Candidate c;
		c.launch.kernel = TraceKernel::Simple;
		c.params = PhxLaunchParams{0, 0, 0, 256, 0, 0, 0, 0, 0};
		out.push_back(c);
		c.launch.kernel = TraceKernel::Packet;
		c.params = PhxLaunchParams{0, 0, 0, 256, 0, 0, 0, 0, 0};
		out.push_back(c);
	}
	return out;
}

// Times `launch(L)` for every candidate with HIP events on the context stream and returns the
// fastest. After one warm-up round, `reps` rounds each time every candidate once; a candidate's
// time is its median round (single launches vary by a few percent, and a minimum would reward
// the candidate with the luckiest draw). Interleaving the rounds spreads clock and cache drift
// over all candidates instead of favouring those timed last.
template <class Launch>
phx_status tune(PhxContext* ctx, std::vector<Candidate>& cands, int reps, Launch launch,
                Candidate& best)
{
	hipEvent_t a = nullptr, b = nullptr;
	PHX_HIP(hipEventCreate(&a), "autotune");
	hipError_t e = hipEventCreate(&b);
	for (Candidate& c : cands)
		if (e == hipSuccess) e = launch(c.launch);
	std::vector<std::vector<float>> times(cands.size());
	for (int r = 0; r < reps && e == hipSuccess; r++)
		for (size_t i = 0; i < cands.size() && e == hipSuccess; i++) {
			e = hipEventRecord(a, ctx->stream);
			if (e == hipSuccess) e = launch(cands[i].launch);
			if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
			if (e == hipSuccess) e = hipEventSynchronize(b);
			float ms = 0.0f;
			if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
			if (e == hipSuccess) times[i].push_back(ms);
		}
	hipEventDestroy(a);
	if (b) hipEventDestroy(b);
	if (e != hipSuccess) return hipFail(e, "autotune");
	float bestMs = 3.0e38f;
	for (size_t i = 0; i < cands.size(); i++) {
		std::vector<float>& t = times[i];
		std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
		cands[i].ms = t.empty() ? 3.0e38f : t[t.size() / 2];
	}
	for (const Candidate& c : cands)
		if (c.ms < bestMs) {
			bestMs = c.ms;
			best = c;
		}
	return ok();
}

// Second stage for a persistent winner: rays per fetch, then the stack split, one at a time around
// `best` (coordinate descent; the current value stays a candidate so `best` only improves). A
// split that lowers the blocks a CU holds caps the grid at the new resident grid.
template <class Launch>
phx_status refine(PhxContext* ctx, const PhxPlan* plan, int reps, Launch launch, Candidate& best)
{
	if (best.launch.kernel != TraceKernel::Persistent) return PHX_OK;
	for (int stage = 0; stage < 2; stage++) {
		std::vector<Candidate> cands{best};
		for (uint32_t v : stage == 0 ? std::vector<uint32_t>{16u, 32u, 64u, 128u, 256u}
		                             : std::vector<uint32_t>{8u, 16u, 24u, 32u}) {
			Candidate c = best;
			if (stage == 0) {
				if (v == best.params.raysPerFetch) continue;
				c.params.raysPerFetch = v;
			} else {
				if (v == best.params.stackLdsEntries) continue;
				c.params.stackLdsEntries = v;
				const uint32_t bpc = blocksPerCU(plan, c.params.blockThreads, c.params.topNodes,
				                                 c.params.helperWaves, v);
				const uint32_t resident =
					std::min(ctx->numCUs * bpc, ctx->spillLanes / c.params.blockThreads);
				if (resident == 0) continue;
				c.params.gridBlocksClosest = std::min(c.params.gridBlocksClosest, resident);
				c.params.gridBlocksAnyhit = std::min(c.params.gridBlocksAnyhit, resident);
				c.params.residentGridBlocks = resident;
			}
			if (shapeFor(ctx, c.params, c.launch, plan) == PHX_OK) cands.push_back(c);
		}
		phx_status st = tune(ctx, cands, reps, launch, best);
		if (st != PHX_OK) return st;
	}
	return PHX_OK;
}

}  // namespace

extern "C" {

static phx_status phx_autotune_impl(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                                    uint32_t numRays, uint32_t flags, PhxLaunchParams* chosen)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!rays || numRays == 0 || !(flags & PHX_MEM_DEVICE))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_autotune: needs device rays");
	const bool anyhit = (flags & PHX_AUTOTUNE_ANYHIT) != 0;
	void* out = nullptr;
	PHX_HIP(hipMalloc(&out, size_t(numRays) * (anyhit ? 1 : sizeof(PhxRayHit))), "phx_autotune");
	std::vector<Candidate> cands = candidates(ctx, nullptr);
	SceneView sv = view(scene);
	Candidate best;
	auto cast = [&](const TraceLaunch& L) {
		return anyhit ? launchAnyHit(sv, rays, numRays, static_cast<uint8_t*>(out), L, ctx->stream)
		              : launchClosest(sv, rays, numRays, static_cast<PhxRayHit*>(out), L, nullptr,
		                              ctx->stream);
	};
	st = tune(ctx, cands, 13, cast, best);
	if (st == PHX_OK) st = refine(ctx, nullptr, 15, cast, best);
	hipStreamSynchronize(ctx->stream);
	hipFree(out);
	if (st != PHX_OK) return st;
	// Apply: the one-ray-per-lane and packet kernels are selected per call by flag; record the
	// persistent shape.
	if (best.launch.kernel == TraceKernel::Persistent) {
		ctx->blockThreads = best.launch.blockThreads;
		ctx->topNodes = best.launch.topNodes;
		if (anyhit)
			ctx->gridAnyhit = best.launch.gridBlocks;
		else
			ctx->gridClosest = best.launch.gridBlocks;
		ctx->refillMin = best.launch.refillMin;
		ctx->helperWaves = best.launch.helperWaves;
		ctx->ldsDepth = best.launch.ldsDepth;
		ctx->window = best.launch.window;
	}
	if (chosen) {
		*chosen = best.params;
		chosen->residentGridBlocks = best.launch.kernel == TraceKernel::Simple   ? 0xFFFFFFFFu
		                             : best.launch.kernel == TraceKernel::Packet ? 0xFFFFFFFEu
		                                                                         : 0u;
	}
	return ok();
}

phx_status phx_autotune(PhxContext* ctx, const PhxScene* scene, const PhxRayIn* rays,
                        uint32_t numRays, uint32_t flags, PhxLaunchParams* chosen)
{
	return guarded("phx_autotune", [&] {
		return phx_autotune_impl(ctx, scene, rays, numRays, flags, chosen);
	});
}

phx_status phx_plan_set_launch_params(PhxContext* ctx, PhxPlan* plan, const PhxLaunchParams* p)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!plan || !registry().has(plan) || plan->ctx != ctx)
		return fail(PHX_ERR_INVALID_HANDLE, "invalid plan handle");
	if (!p) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_plan_set_launch_params: NULL params");
	TraceLaunch L;
	st = shapeFor(ctx, *p, L, plan);
	if (st != PHX_OK) return st;
	plan->launch = L;
	if (plan->graphExec) {
		hipGraphExecDestroy(plan->graphExec);
		hipGraphDestroy(plan->graph);
		plan->graphExec = nullptr;
		plan->graph = nullptr;
	}
	return ok();
}

static phx_status phx_plan_autotune_impl(PhxContext* ctx, PhxPlan* plan, PhxLaunchParams* chosen)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!plan || !registry().has(plan) || plan->ctx != ctx)
		return fail(PHX_ERR_INVALID_HANDLE, "invalid plan handle");
	std::vector<Candidate> cands = candidates(ctx, plan);
	Candidate best;
	// Timed as the plan launches: with its completion tracking, whose LDS can change residency.
	auto run = [&](const TraceLaunch& L) {
		return launchBatches(plan->descs, plan->offsets, plan->numBatches, plan->total, L,
		                     plan->treeletNodes, plan->treeletCount, plan->clock, ctx->stream);
	};
	st = tune(ctx, cands, 3, run, best);
	if (st == PHX_OK) st = refine(ctx, plan, 5, run, best);
	if (st != PHX_OK) return st;
	st = phx_plan_set_launch_params(ctx, plan, &best.params);
	if (st != PHX_OK) return st;
	if (chosen) *chosen = best.params;
	return ok();
}

phx_status phx_plan_autotune(PhxContext* ctx, PhxPlan* plan, PhxLaunchParams* chosen)
{
	return guarded("phx_plan_autotune", [&] {
		return phx_plan_autotune_impl(ctx, plan, chosen);
	});
}

void phx_plan_release(PhxPlan* plan)
{
	if (!plan || !registry().remove(plan)) return;
	hipSetDevice(plan->ctx->device);
	hipStreamSynchronize(plan->ctx->stream);
	if (plan->graphExec) hipGraphExecDestroy(plan->graphExec);
	if (plan->graph) hipGraphDestroy(plan->graph);
	hipFree(plan->descs);
	hipFree(plan->offsets);
	hipFree(plan->clock);
	for (const PhxScene* sc : plan->scenes) dropScene(sc);
	PhxContext* ctx = plan->ctx;
	delete plan;
	dropContext(ctx);
}

phx_status phx_gen_primary_rays(PhxContext* ctx, const PhxCameraDef* cam, uint32_t width,
                                uint32_t height, float tMin, float tMax, uint32_t order,
                                PhxRayIn* rays, uint32_t flags)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!cam || !rays || width == 0 || height == 0)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_primary_rays: bad arguments");
	if (order != PHX_RAYS_ROW_MAJOR && order != PHX_RAYS_TILED_8X8)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_primary_rays: unknown order %u", order);
	if (order == PHX_RAYS_TILED_8X8 && ((width % 8) || (height % 8)))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_primary_rays: tiled order needs w,h %% 8 == 0");
	uint64_t n = uint64_t(width) * height;
	if (n > 0xFFFFFFFFull) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_primary_rays: too many rays");
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	const bool async = (flags & PHX_ASYNC) != 0;
	if (async && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_gen_primary_rays: PHX_ASYNC needs PHX_MEM_DEVICE");
	PhxRayIn* dst = rays;
	if (!device) {
		st = growStage(ctx->stageOut, ctx->stageOutBytes, n * sizeof(PhxRayIn));
		if (st != PHX_OK) return st;
		dst = static_cast<PhxRayIn*>(ctx->stageOut);
	}
	PHX_HIP(launchGenPrimaryRays(*cam, width, height, 0.0f, 0.0f, tMin, tMax, order, dst,
	                             ctx->stream),
	        "phx_gen_primary_rays");
	if (!device)
		PHX_HIP(hipMemcpyAsync(rays, dst, n * sizeof(PhxRayIn), hipMemcpyDeviceToHost, ctx->stream),
		        "phx_gen_primary_rays");
	if (!async) PHX_HIP(hipStreamSynchronize(ctx->stream), "phx_gen_primary_rays");
	return ok();
}

}  // extern "C"
