// Host runtime objects shared by the translation units of the phx library: handle registry,
// contexts, scenes, plans and the launch-shape helpers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <mutex>
#include <unordered_set>
#include <vector>

#include "kernels.hpp"
#include "phx_internal.hpp"

namespace phx {

// Live handles, so a released or foreign pointer is reported instead of dereferenced.
struct Registry {
	std::mutex mu;
	std::unordered_set<const void*> live;
	void add(const void* p)
	{
		std::lock_guard<std::mutex> g(mu);
		live.insert(p);
	}
	bool remove(const void* p)
	{
		std::lock_guard<std::mutex> g(mu);
		return live.erase(p) != 0;
	}
	bool has(const void* p)
	{
		std::lock_guard<std::mutex> g(mu);
		return live.count(p) != 0;
	}
};

inline Registry& registry()
{
	static Registry r;
	return r;
}

inline phx_status hipFail(hipError_t e, const char* what)
{
	return fail(PHX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define PHX_HIP(call, what)                        \
	do {                                           \
		hipError_t e_ = (call);                    \
		if (e_ != hipSuccess) return hipFail(e_, what); \
	} while (0)

}  // namespace phx

struct PhxBVH {
	phx::BuiltBVH bvh;
};

struct PhxContext {
	int device = 0;
	hipStream_t stream = nullptr;
	uint32_t numCUs = 0;
	// Persistent-kernel state of this context's stream: the ray counter pair (left at zero by
	// every launch), the spill area of the LDS stacks, and the launch shape.
	uint32_t* work = nullptr;
	uint32_t* spill = nullptr;
	uint32_t gridClosest = 0;
	uint32_t gridAnyhit = 0;
	uint32_t blockThreads = 512;
	uint32_t topNodes = 0;
	uint32_t refillMin = 8;
	uint32_t helperWaves = 0;
	uint32_t ldsDepth = 16;   // stack entries per lane kept in LDS
	uint32_t window = 64;     // rays per fetch
	uint32_t spillLanes = 0;
	// Staging buffers for calls that pass host pointers, grown on demand.
	void* stageIn = nullptr;
	size_t stageInBytes = 0;
	void* stageOut = nullptr;
	size_t stageOutBytes = 0;
	// Scenes, plans, renderers, instance sets and servers made from this context (guarded by
	// sceneHoldMutex()). A context destroyed while any of them exists leaves the registry at once
	// but keeps its stream and buffers until the last of them is released.
	uint32_t holders = 0;
	bool released = false;
};

struct PhxPlan {
	PhxContext* ctx = nullptr;
	phx::TraceLaunch launch;  // launch shape of this plan (set at creation, tunable)
	uint32_t numBatches = 0;
	uint32_t total = 0;
	phx::BatchDesc* descs = nullptr;
	uint32_t* offsets = nullptr;
	phx::BatchClock* clock = nullptr;
	const float4* treeletNodes = nullptr;
	uint32_t treeletCount = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t graphExec = nullptr;
	std::vector<phx::BatchDesc> hostDescs;
	std::vector<const PhxScene*> scenes;  // distinct scenes of the batches, held by the plan
};

struct PhxScene {
	PhxContext* ctx = nullptr;
	PhxBVHNode* nodes = nullptr;
	PhxTriangleVertices* tris = nullptr;
	PhxTriangleData* datas = nullptr;
	PhxMaterial* materials = nullptr;
	PhxSphereLight* lights = nullptr;
	uint32_t numNodes = 0, numTris = 0, numMaterials = 0, numLights = 0;
	uint32_t depth = 0;
	uint32_t finiteBounds = 0;
	float rootMin[3] = {0, 0, 0}, rootMax[3] = {0, 0, 0};  // union of the root's child boxes
	uint32_t topCount = 0;
	size_t bytes = 0;
	void* texArena = nullptr;  // material textures (phx_scene_set_textures)
	size_t texBytes = 0;
	phx::TexSet tex;
	// Process-wide unique stamp of the scene's contents, renewed when the textures are replaced:
	// a captured frame graph is replayed only for the stamp it was captured with, so a scene
	// uploaded later at a released scene's address is never mistaken for it.
	uint64_t version = 0;
	// Plans and servers that launch on this scene (guarded by sceneHoldMutex()). A scene released
	// while held leaves the registry at once but keeps its device arrays until the last holder
	// lets go.
	uint32_t holders = 0;
	bool released = false;
};

namespace phx {

inline phx_status checkCtx(PhxContext* ctx)
{
	if (!ctx || !registry().has(ctx)) return fail(PHX_ERR_INVALID_HANDLE, "invalid context handle");
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return fail(PHX_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
	return PHX_OK;
}

inline phx_status checkScene(PhxContext* ctx, const PhxScene* scene)
{
	if (!scene || !registry().has(scene)) return fail(PHX_ERR_INVALID_HANDLE, "invalid scene handle");
	if (scene->ctx != ctx) return fail(PHX_ERR_INVALID_HANDLE, "scene belongs to another context");
	return PHX_OK;
}

// Views of an instance set (instances.cpp) for the renderer.
const PhxBVHNode* instanceTop(const PhxInstanceSet* s);
const DevInstance* instanceArray(const PhxInstanceSet* s);
const DevInstance* instanceByIndex(const PhxInstanceSet* s);
uint32_t instanceCount(const PhxInstanceSet* s);
uint64_t instanceStamp(const PhxInstanceSet* s);
bool instanceSetValid(PhxContext* ctx, const PhxInstanceSet* s);

// A new value of the process-wide stamp counter (PhxScene::version).
uint64_t nextSceneStamp();
// Takes / drops a plan's or server's hold on a scene; dropping the last hold of a released scene
// frees it.
void holdScene(const PhxScene* s);
void dropScene(const PhxScene* s);
// A context's holders: every object made from it holds it from creation to release.
void holdContext(PhxContext* ctx);
void dropContext(PhxContext* ctx);

inline SceneView view(const PhxScene* s)
{
	return SceneView{s->nodes, s->tris, s->numNodes, s->numTris, s->finiteBounds, s->topCount};
}

inline TraceLaunch traceLaunch(const PhxContext* ctx, uint32_t flags, bool anyhit)
{
	TraceLaunch L;
	L.kernel = (flags & PHX_KERNEL_ONE_RAY_PER_LANE) ? TraceKernel::Simple
	           : (flags & PHX_KERNEL_PACKET)        ? TraceKernel::Packet
	                                                : TraceKernel::Persistent;
	L.gridBlocks = anyhit ? ctx->gridAnyhit : ctx->gridClosest;
	L.blockThreads = ctx->blockThreads;
	L.topNodes = ctx->topNodes;
	L.refillMin = ctx->refillMin;
	L.spillLanes = ctx->spillLanes;
	L.helperWaves = ctx->helperWaves;
	L.ldsDepth = ctx->ldsDepth;
	L.window = ctx->window;
	L.work = ctx->work;
	L.spill = ctx->spill;
	return L;
}

}  // namespace phx
