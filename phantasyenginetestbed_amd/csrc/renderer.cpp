// The render-frame C ABI: frame buffers of one resolution and the wavefront frame sequence
// (CudaTracerRenderer::render's normal path, CudaTracerRenderer.cpp:637-720, with the G-buffer
// traced instead of rasterised) followed by the TAA resolve of GameScreen::render.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>

#include "runtime.hpp"

using namespace phx;

struct PhxRenderer {
	PhxContext* ctx = nullptr;
	uint32_t w = 0, h = 0, n = 0;    // full resolution
	uint32_t hw = 0, hh = 0, m = 0;  // half resolution (bounce records)
	uint32_t lightCap = 0;           // lights the per-light buffers hold
	PhxRayIn* prim = nullptr;
	PhxRayHit* primHits = nullptr;
	float4* gbuf = nullptr;
	PhxRayIn* bounce = nullptr;
	float4* throughput = nullptr;
	PhxRayHit* bounceHits = nullptr;
	float4* bounceInfo = nullptr;
	float4* contrib = nullptr;  // n * lightCap
	PhxRayIn* shadow = nullptr;  // n * lightCap
	uint8_t* inLight = nullptr;  // n * lightCap
	float4* color = nullptr;
	float4* hist[2] = {nullptr, nullptr};
	float2* vel[2] = {nullptr, nullptr};  // TAA: velocity of the frame that wrote hist[i]
	float4* out = nullptr;
	uint32_t histIndex = 0;
	bool haveHistory = false;
	PhxCameraDef lastCam{};  // the previous frame's camera (TAA velocity)
	bool haveLastCam = false;
	// Instances drawn besides the scene (phx_renderer_set_instances) and the instance index of
	// each primary / bounce hit.
	const PhxInstanceSet* instances = nullptr;
	uint32_t* primInst = nullptr;
	uint32_t* bounceInst = nullptr;
	float4* lastOut = nullptr;
	FrameParams* params = nullptr;  // per-frame values the frame kernels read (device)
	// The bounce rays' closest hits run on a side stream beside the primary shadow rays (fork
	// after the bounce rays are generated, join before the bounce records).
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
	// The captured frame (PHX_FRAME_GRAPH) and what it was captured for. The cameras, jitter and
	// seeds are read from the per-frame block, so a moving camera replays the same graph.
	struct GraphKey {
		const PhxScene* scene;
		uint64_t sceneVersion;  // unique per upload, renewed when the scene's textures are replaced
		uint64_t instanceStamp;  // the instance set's (0: none)
		uint32_t flags;  // PHX_FRAME_TAA | PHX_FRAME_NO_CLAMP
		uint32_t numLights;
		const void* lightBuffers;
		uint32_t launch[8];  // the context's launch shape
	};
	GraphKey key{};
	hipGraph_t graph = nullptr;
	hipGraphExec_t graphExec = nullptr;
};

namespace {

void dropGraph(PhxRenderer* r)
{
	if (r->graphExec) hipGraphExecDestroy(r->graphExec);
	if (r->graph) hipGraphDestroy(r->graph);
	r->graphExec = nullptr;
	r->graph = nullptr;
}

void freeRenderer(PhxRenderer* r)
{
	dropGraph(r);
	if (r->fork) hipEventDestroy(r->fork);
	if (r->join) hipEventDestroy(r->join);
	if (r->side) hipStreamDestroy(r->side);
	void* bufs[] = {r->prim,    r->primHits, r->gbuf,   r->bounce,  r->throughput,
	                r->bounceHits, r->bounceInfo, r->contrib, r->shadow, r->inLight,
	                r->color,   r->hist[0],  r->hist[1], r->out,    r->params,
	                r->vel[0],  r->vel[1], r->primInst, r->bounceInst};
	for (void* b : bufs) hipFree(b);
	delete r;
}

template <class T>
phx_status devAlloc(T*& p, size_t count, const char* what)
{
	p = nullptr;
	if (count == 0) return PHX_OK;
	if (hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)) != hipSuccess) {
		p = nullptr;
		return fail(PHX_ERR_OUT_OF_MEMORY, "%s: device allocation of %zu bytes failed", what,
		            count * sizeof(T));
	}
	return PHX_OK;
}

phx_status growLights(PhxRenderer* r, uint32_t lights)
{
	if (lights <= r->lightCap) return PHX_OK;
	hipStreamSynchronize(r->ctx->stream);  // earlier frames may still use the buffers
	dropGraph(r);                           // a captured frame uses them too
	hipFree(r->contrib);
	hipFree(r->shadow);
	hipFree(r->inLight);
	r->contrib = nullptr;
	r->shadow = nullptr;
	r->inLight = nullptr;
	r->lightCap = 0;
	const size_t k = size_t(r->n) * lights;
	phx_status st = devAlloc(r->contrib, k, "phx_render_frame");
	if (st == PHX_OK) st = devAlloc(r->shadow, k, "phx_render_frame");
	if (st == PHX_OK) st = devAlloc(r->inLight, k, "phx_render_frame");
	if (st == PHX_OK) r->lightCap = lights;
	return st;
}

// The frame's launches, all on stream s; per-frame values come from r->params, so this sequence is
// the same for every frame of a scene, camera and flags (and can be captured as a graph).
phx_status enqueueFrame(PhxContext* ctx, PhxRenderer* r, const PhxScene* scene,
                        const PhxCameraDef* cam, uint32_t flags, hipStream_t s)
{
	const uint32_t L = scene->numLights;
	const bool taa = (flags & PHX_FRAME_TAA) != 0;
	const uint32_t order = (r->w % 8u == 0u && r->h % 8u == 0u) ? PHX_RAYS_TILED_8X8 : PHX_RAYS_ROW_MAJOR;
	const SceneView sv = view(scene);
	const TraceLaunch closest = traceLaunch(ctx, 0, false);
	// Shadow rays are coherent -- a pixel tile's rays towards one light, tiles in order -- so they
	// are traced as packets, each wave expanding the union of 64 rays' nodes. So are camera rays in
	// 8x8-tile order (a wave's 64 rays are one tile): 0.50 ms instead of 0.77 ms for a 1080p
	// frame's closest hits; in row-major order a wave's rays span a 64-pixel row segment and
	// packets gain nothing, so those go to the persistent kernel.
	TraceLaunch shadows = traceLaunch(ctx, 0, true);
	shadows.kernel = TraceKernel::Packet;
	// The bounce hits' shadow rays start at scattered points (neighbouring half-resolution pixels'
	// bounce rays go anywhere), so a packet's union of nodes is large; one ray per lane traces
	// them fastest (798,504 such rays: 0.18 ms, against 0.25 persistent and 0.35 as packets).
	TraceLaunch bounceShadows = shadows;
	bounceShadows.kernel = TraceKernel::Simple;
	TraceLaunch primary = closest;
	if (order == PHX_RAYS_TILED_8X8) primary.kernel = TraceKernel::Packet;
	const char* what = "phx_render_frame";
	// Instances besides the scene: their closest hits are merged into the scene's (the scene
	// wins ties), their any hits occlude too, and hit records read the instance's mesh.
	const PhxInstanceSet* inst = r->instances && instanceCount(r->instances) ? r->instances : nullptr;
	const uint32_t numInst = inst ? instanceCount(inst) : 0u;
	const DevInstance* instByIndex = inst ? instanceByIndex(inst) : nullptr;

	// Primary hits -> hit records -> contributions and shadow rays of every light.
	PHX_HIP(launchGenPrimaryRays(*cam, r->w, r->h, 0.0f, 0.0f, 1e-4f, FLT_MAX, order, r->prim, s,
	                             r->params),
	        what);
	PHX_HIP(launchClosest(sv, r->prim, r->n, r->primHits, primary, nullptr, s), what);
	if (inst)
		PHX_HIP(launchInstances(instanceTop(inst), instanceArray(inst), numInst, r->prim, r->n,
		                        r->primHits, r->primInst, nullptr, true, s),
		        what);
	PHX_HIP(launchInterpret(r->prim, r->primHits, r->n, scene->datas, scene->numTris,
	                        scene->materials, scene->numMaterials, scene->tex, r->gbuf, s,
	                        inst ? r->primInst : nullptr, instByIndex, numInst),
	        what);
	PHX_HIP(launchShade(r->gbuf, nullptr, nullptr, scene->lights, L, r->n, r->params, 1u,
	                    r->contrib, r->shadow, s),
	        what);
	// Bounce rays (they need only the hit records): incoherent, traced by the persistent kernel.
	PHX_HIP(launchGenSecondary(r->gbuf, r->w, r->h, r->params, r->bounce, r->throughput, s),
	        what);
	// The bounce rays' closest hits (the persistent kernel, whose last rays take long) overlap the
	// primary shadow rays and their shadow logic; the bounce records need the shading buffers back,
	// so the streams join before them. No other persistent launch runs meanwhile, so the two
	// streams never share the context's ray counter or spill area.
	if (r->m) {
		PHX_HIP(hipEventRecord(r->fork, s), what);
		PHX_HIP(hipStreamWaitEvent(r->side, r->fork, 0), what);
		PHX_HIP(launchClosest(sv, r->bounce, r->m, r->bounceHits, closest, nullptr, r->side), what);
		if (inst)
			PHX_HIP(launchInstances(instanceTop(inst), instanceArray(inst), numInst, r->bounce,
			                        r->m, r->bounceHits, r->bounceInst, nullptr, true, r->side),
			        what);
		PHX_HIP(hipEventRecord(r->join, r->side), what);
	}
	PHX_HIP(launchAnyHit(sv, r->shadow, r->n * L, r->inLight, shadows, s), what);
	if (inst)
		PHX_HIP(launchInstances(instanceTop(inst), instanceArray(inst), numInst, r->shadow,
		                        r->n * L, nullptr, nullptr, r->inLight, true, s),
		        what);
	PHX_HIP(launchShadowLogic(r->color, r->w, r->h, 1u, false, r->inLight, r->contrib, L, s), what);

	// Bounce hits: records, shading x throughput with shadow rays, any hit, added at half res.
	if (r->m) {
		PHX_HIP(hipStreamWaitEvent(s, r->join, 0), what);
		PHX_HIP(launchInterpret(r->bounce, r->bounceHits, r->m, scene->datas, scene->numTris,
		                        scene->materials, scene->numMaterials, scene->tex, r->bounceInfo,
		                        s, inst ? r->bounceInst : nullptr, instByIndex, numInst),
		        what);
		PHX_HIP(launchShade(r->bounceInfo, r->bounce, r->throughput, scene->lights, L, r->m,
		                    r->params, 3u, r->contrib, r->shadow, s),
		        what);
		PHX_HIP(launchAnyHit(sv, r->shadow, r->m * L, r->inLight, bounceShadows, s), what);
		if (inst)
			PHX_HIP(launchInstances(instanceTop(inst), instanceArray(inst), numInst, r->shadow,
			                        r->m * L, nullptr, nullptr, r->inLight, true, s),
			        what);
		PHX_HIP(launchShadowLogic(r->color, r->w, r->h, 2u, true, r->inLight, r->contrib, L, s),
		        what);
	}

	if (taa)
		PHX_HIP(launchTaa(r->color, r->hist[0], r->hist[1], r->vel[0], r->vel[1], r->gbuf,
		                  r->prim, r->w, r->h, !(flags & PHX_FRAME_NO_CLAMP), r->params, r->out,
		                  s),
		        what);
	return PHX_OK;
}

}  // namespace

extern "C" {

void phx_halton23(uint32_t frameIndex, float* jx, float* jy)
{
	const uint32_t i = frameIndex % 16u + 1u;
	double r = 0.0, inv = 0.5;
	for (uint32_t k = i; k; k /= 2u, inv *= 0.5) r += double(k % 2u) * inv;
	if (jx) *jx = float(r);
	r = 0.0;
	inv = 1.0 / 3.0;
	for (uint32_t k = i; k; k /= 3u, inv /= 3.0) r += double(k % 3u) * inv;
	if (jy) *jy = float(r);
}

phx_status phx_renderer_create(PhxContext* ctx, uint32_t width, uint32_t height,
                               PhxRenderer** outRenderer)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outRenderer) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_renderer_create: NULL output");
	*outRenderer = nullptr;
	if (width == 0 || height == 0 || width > 16384 || height > 16384)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_renderer_create: resolution %ux%u", width,
		            height);
	PhxRenderer* r = new (std::nothrow) PhxRenderer();
	if (!r) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_renderer_create: out of host memory");
	r->ctx = ctx;
	r->w = width;
	r->h = height;
	r->n = width * height;
	r->hw = width / 2u;
	r->hh = height / 2u;
	r->m = r->hw * r->hh;
	const char* what = "phx_renderer_create";
	st = devAlloc(r->prim, r->n, what);
	if (st == PHX_OK) st = devAlloc(r->primHits, r->n, what);
	if (st == PHX_OK) st = devAlloc(r->gbuf, size_t(r->n) * 3u, what);
	if (st == PHX_OK) st = devAlloc(r->bounce, r->m, what);
	if (st == PHX_OK) st = devAlloc(r->throughput, r->m, what);
	if (st == PHX_OK) st = devAlloc(r->bounceHits, r->m, what);
	if (st == PHX_OK) st = devAlloc(r->bounceInfo, size_t(r->m) * 3u, what);
	if (st == PHX_OK) st = devAlloc(r->color, r->n, what);
	if (st == PHX_OK) st = devAlloc(r->hist[0], r->n, what);
	if (st == PHX_OK) st = devAlloc(r->hist[1], r->n, what);
	if (st == PHX_OK) st = devAlloc(r->primInst, r->n, what);
	if (st == PHX_OK) st = devAlloc(r->bounceInst, std::max(r->m, 1u), what);
	if (st == PHX_OK) st = devAlloc(r->vel[0], r->n, what);
	if (st == PHX_OK) st = devAlloc(r->vel[1], r->n, what);
	if (st == PHX_OK) st = devAlloc(r->out, r->n, what);
	if (st == PHX_OK) st = devAlloc(r->params, 1, what);
	if (st == PHX_OK && (hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking) != hipSuccess ||
	                     hipEventCreateWithFlags(&r->fork, hipEventDisableTiming) != hipSuccess ||
	                     hipEventCreateWithFlags(&r->join, hipEventDisableTiming) != hipSuccess))
		st = fail(PHX_ERR_HIP, "%s: stream or event creation failed", what);
	if (st != PHX_OK) {
		freeRenderer(r);
		return st;
	}
	r->lastOut = r->color;
	holdContext(ctx);
	registry().add(r);
	*outRenderer = r;
	return ok();
}

void phx_renderer_release(PhxRenderer* r)
{
	if (!r || !registry().remove(r)) return;
	PhxContext* ctx = r->ctx;
	hipSetDevice(ctx->device);
	hipStreamSynchronize(ctx->stream);
	freeRenderer(r);
	dropContext(ctx);
}

phx_status phx_renderer_buffer(PhxRenderer* r, uint32_t which, const void** devPtr, size_t* bytes)
{
	if (!r || !registry().has(r)) return fail(PHX_ERR_INVALID_HANDLE, "invalid renderer handle");
	if (!devPtr) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_renderer_buffer: NULL output");
	const void* p = nullptr;
	size_t b = 0;
	switch (which) {
	case PHX_BUFFER_COLOR: p = r->color; b = size_t(r->n) * 16u; break;
	case PHX_BUFFER_OUTPUT: p = r->lastOut; b = size_t(r->n) * 16u; break;
	case PHX_BUFFER_GBUFFER: p = r->gbuf; b = size_t(r->n) * 48u; break;
	case PHX_BUFFER_BOUNCE_RAYS: p = r->bounce; b = size_t(r->m) * sizeof(PhxRayIn); break;
	case PHX_BUFFER_HISTORY: p = r->hist[r->histIndex]; b = size_t(r->n) * 16u; break;
	case PHX_BUFFER_VELOCITY: p = r->vel[r->histIndex]; b = size_t(r->n) * 8u; break;
	default: return fail(PHX_ERR_INVALID_ARGUMENT, "phx_renderer_buffer: unknown buffer %u", which);
	}
	*devPtr = p;
	if (bytes) *bytes = b;
	return ok();
}

phx_status phx_renderer_set_instances(PhxRenderer* r, const PhxInstanceSet* set)
{
	if (!r || !registry().has(r)) return fail(PHX_ERR_INVALID_HANDLE, "invalid renderer handle");
	if (set && !instanceSetValid(r->ctx, set))
		return fail(PHX_ERR_INVALID_HANDLE, "phx_renderer_set_instances: invalid instance set");
	r->instances = set;
	return ok();
}

phx_status phx_render_frame(PhxContext* ctx, PhxRenderer* r, const PhxScene* scene,
                            const PhxCameraDef* cam, uint32_t frameIndex, uint64_t seed,
                            uint32_t flags, float* rgba)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	st = checkScene(ctx, scene);
	if (st != PHX_OK) return st;
	if (!r || !registry().has(r) || r->ctx != ctx)
		return fail(PHX_ERR_INVALID_HANDLE, "invalid renderer handle");
	if (!cam) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_render_frame: NULL camera");
	if (r->instances && !instanceSetValid(ctx, r->instances))
		return fail(PHX_ERR_INVALID_HANDLE, "phx_render_frame: the renderer's instance set was "
		            "released");
	if (!scene->datas) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_render_frame: scene has no TriangleData");
	if (uint64_t(r->n) * scene->numLights + r->m > 0x7FFFFFFFull)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_render_frame: more than 2^31 shadow rays");
	const bool device = (flags & PHX_MEM_DEVICE) != 0;
	const bool async = (flags & PHX_ASYNC) != 0;
	if (async && rgba && !device)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_render_frame: PHX_ASYNC needs PHX_MEM_DEVICE");
	st = growLights(r, scene->numLights);
	if (st != PHX_OK) return st;

	hipStream_t s = ctx->stream;
	const bool taa = (flags & PHX_FRAME_TAA) != 0;
	const char* what = "phx_render_frame";
	FrameParams fp{};
	if (taa) phx_halton23(frameIndex, &fp.jx, &fp.jy);
	fp.frame = frameIndex;
	fp.taaReset = (flags & PHX_FRAME_TAA_RESET) || !r->haveHistory ? 1u : 0u;
	fp.seed = seed;
	fp.histIn = r->histIndex;
	fp.cam = *cam;
	fp.prevCam = r->haveLastCam ? r->lastCam : *cam;
	PHX_HIP(launchSetFrameParams(r->params, fp, s), what);
	if (flags & PHX_FRAME_GRAPH) {
		PhxRenderer::GraphKey k{};
		k.scene = scene;
		k.sceneVersion = scene->version;
		k.instanceStamp = r->instances ? instanceStamp(r->instances) : 0u;
		k.flags = flags & (PHX_FRAME_TAA | PHX_FRAME_NO_CLAMP);
		k.numLights = scene->numLights;
		k.lightBuffers = r->contrib;
		const uint32_t launch[8] = {ctx->gridClosest, ctx->gridAnyhit, ctx->blockThreads,
		                            ctx->topNodes,      ctx->refillMin, ctx->helperWaves,
		                            ctx->ldsDepth,      ctx->window};
		std::memcpy(k.launch, launch, sizeof(launch));
		if (!r->graphExec || std::memcmp(&k, &r->key, sizeof(k)) != 0) {
			dropGraph(r);
			PHX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), what);
			st = enqueueFrame(ctx, r, scene, cam, flags, s);
			hipGraph_t g = nullptr;
			const hipError_t e = hipStreamEndCapture(s, &g);
			if (st != PHX_OK) {
				if (g) hipGraphDestroy(g);
				return st;
			}
			PHX_HIP(e, "phx_render_frame: capture");
			r->graph = g;
			PHX_HIP(hipGraphInstantiate(&r->graphExec, g, nullptr, nullptr, 0),
			        "phx_render_frame: graph instantiate");
			r->key = k;
		}
		PHX_HIP(hipGraphLaunch(r->graphExec, s), "phx_render_frame: graph launch");
	} else {
		st = enqueueFrame(ctx, r, scene, cam, flags, s);
		if (st != PHX_OK) return st;
	}

	r->lastOut = r->color;
	r->lastCam = *cam;
	r->haveLastCam = true;
	if (taa) {
		r->histIndex ^= 1u;
		r->haveHistory = true;
		r->lastOut = r->out;
	}
	if (rgba)
		PHX_HIP(hipMemcpyAsync(rgba, r->lastOut, size_t(r->n) * 16u,
		                       device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s),
		        what);
	if (!async) PHX_HIP(hipStreamSynchronize(s), what);
	return ok();
}

}  // extern "C"
