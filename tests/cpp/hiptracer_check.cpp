// Engine-side check of the C++ renderer interface (HipTracerRenderer) and the kernel-level launch
// functions, written the way engine code uses the reference's CudaTracerRenderer: set materials
// and the static scene, set the target resolution, update the camera, render frames. Every frame
// must equal the C ABI's phx_render_frame on the same inputs (textured materials included), byte
// for byte, and the launch functions must equal phx_cast_closest / phx_cast_anyhit. Built as a
// shared library and called in process (hiptracer_check_run returns the number of failed
// checks).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../phantasyenginetestbed_amd/cpp/HipTracerRenderer.hpp"

using namespace phe_hip;

static int failures = 0;

#define CHECK(cond, ...)                              \
	do {                                              \
		if (!(cond)) {                                \
			std::printf("FAILED %s:%d: ", __FILE__, __LINE__); \
			std::printf(__VA_ARGS__);                 \
			std::printf("\n");                        \
			failures++;                               \
		}                                             \
	} while (0)

static vec3 normalized(double x, double y, double z)
{
	const double l = std::sqrt(x * x + y * y + z * z);
	return vec3{float(x / l), float(y / l), float(z / l)};
}

extern "C" int hiptracer_check_run()
{
	failures = 0;
	std::setvbuf(stdout, nullptr, _IONBF, 0);
	std::printf("hiptracer_check: scene\n");
	const uint32_t grid = 64, boxes = 40, numMats = 16, numLights = 4;
	const uint32_t n = phx_scene_heightfield_boxes_count(grid, boxes);
	std::vector<PhxTriangleVertices> tris(n);
	std::vector<PhxTriangleData> datas(n);
	CHECK(phx_scene_heightfield_boxes(grid, boxes, 0xC0FFEE02ull, 0.1f, -3.0f, 0.5f, numMats,
	                                  tris.data(), datas.data()) == PHX_OK, "scene");
	PhxBVH* bvh = nullptr;
	CHECK(phx_bvh_build(tris.data(), datas.data(), n, 0, &bvh) == PHX_OK, "bvh build");
	const PhxBVHNode* nodes = nullptr;
	const PhxTriangleVertices* verts = nullptr;
	const PhxTriangleData* bdatas = nullptr;
	uint32_t numNodes = 0, numRefs = 0;
	CHECK(phx_bvh_arrays(bvh, &nodes, &numNodes, &verts, &bdatas, &numRefs) == PHX_OK, "arrays");

	StaticScene scene;
	scene.bvh.nodes.assign(nodes, nodes + numNodes);
	scene.bvh.triangleVerts.assign(verts, verts + numRefs);
	scene.bvh.triangleDatas.assign(bdatas, bdatas + numRefs);
	scene.sphereLights.resize(numLights);
	CHECK(phx_scene_lights_random(numLights, 3, scene.sphereLights.data()) == PHX_OK, "lights");
	std::vector<PhxMaterial> mats(numMats);
	CHECK(phx_scene_materials_random(numMats, 3, mats.data()) == PHX_OK, "materials");
	phx_bvh_release(bvh);

	HipTracerRenderer renderer(0);
	CHECK(renderer.lastStatus() == PHX_OK, "create: %s", renderer.lastError().c_str());
	// render before a scene is set: recorded error, no crash
	RenderResult none = renderer.render({}, {});
	CHECK(none.colorDevice == nullptr && renderer.lastStatus() != PHX_OK, "render without scene");

	renderer.setMaterialsAndTextures(mats, {});
	renderer.setStaticScene(scene);
	CHECK(renderer.lastStatus() == PHX_OK, "setStaticScene: %s", renderer.lastError().c_str());
	const int w = 64, h = 48;
	renderer.setTargetResolution(vec2i{w, h});
	CHECK(renderer.lastStatus() == PHX_OK, "resolution: %s", renderer.lastError().c_str());

	// ViewFrustum::set: dir = normalize(target - pos), up orthogonalised against dir.
	ViewFrustum cam;
	cam.pos = vec3{0.5f, 0.6f, -0.3f};
	cam.dir = normalized(0.0, 0.05 - 0.6, 0.6 + 0.3);
	const double d = double(cam.dir.y);
	cam.up = normalized(-d * cam.dir.x, 1.0 - d * cam.dir.y, -d * cam.dir.z);
	cam.verticalFovDeg = 60.0f;
	renderer.updateCamera(cam);
	renderer.setSeed(11);

	// The same frames through the C ABI directly.
	PhxContext* ctx = renderer.context();
	PhxRenderer* direct = nullptr;
	CHECK(phx_renderer_create(ctx, w, h, &direct) == PHX_OK, "direct renderer");
	PhxCameraDef def;
	const float pos[3] = {cam.pos.x, cam.pos.y, cam.pos.z};
	const float dir[3] = {cam.dir.x, cam.dir.y, cam.dir.z};
	const float up[3] = {cam.up.x, cam.up.y, cam.up.z};
	CHECK(phx_camera_def(pos, dir, up, 60.0f * (3.14159265358979323846f / 180.0f), float(w),
	                     float(h), &def) == PHX_OK, "camera def");
	std::vector<float> a(size_t(w) * h * 4), b(a.size());
	for (uint32_t f = 0; f < 4; f++) {
		std::printf("hiptracer_check: frame %u\n", f);
		RenderResult res = renderer.render({}, {});
		CHECK(res.colorDevice && res.renderedRes.x == w && res.renderedRes.y == h, "render %u: %s",
		      f, renderer.lastError().c_str());
		CHECK(phx_copy_to_host(ctx, a.data(), res.colorDevice, a.size() * 4) == PHX_OK, "copy");
		const uint32_t flags = PHX_FRAME_TAA | (f == 0 ? PHX_FRAME_TAA_RESET : 0u);
		CHECK(phx_render_frame(ctx, direct, renderer.scene(), &def, f, 11, flags, b.data()) ==
		          PHX_OK, "direct frame %u: %s", f, phx_last_error());
		CHECK(std::memcmp(a.data(), b.data(), a.size() * 4) == 0, "frame %u differs", f);
		double sum = 0.0;
		for (size_t i = 0; i < a.size(); i += 4) sum += a[i] + a[i + 1] + a[i + 2];
		CHECK(sum > 0.0 && std::isfinite(sum), "frame %u is black or not finite", f);
	}

	// Dynamic meshes: two meshes (a unit quad facing up and a small box), three objects placed
	// by sfz-layout transforms; each frame equals phx_render_frame with the same instances set on
	// a renderer of the C ABI, and differs from the frame without them.
	std::printf("hiptracer_check: dynamic meshes\n");
	{
		RawMesh quad;
		const float qp[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}};
		for (int k = 0; k < 4; k++) {
			Vertex v;
			v.pos = vec3{qp[k][0], qp[k][1], qp[k][2]};
			v.normal = vec3{0, 1, 0};
			quad.vertices.push_back(v);
			quad.materialIndices.push_back(uint16_t(k % 3));
		}
		quad.indices = {0, 1, 2, 0, 2, 3};
		RawMesh box;
		for (int k = 0; k < 8; k++) {
			Vertex v;
			v.pos = vec3{float(k & 1), float((k >> 1) & 1), float((k >> 2) & 1)};
			v.normal = normalized(v.pos.x - 0.5, v.pos.y - 0.5, v.pos.z - 0.5);
			box.vertices.push_back(v);
			box.materialIndices.push_back(uint16_t(5));
		}
		box.indices = {0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1,
		               2, 3, 7, 2, 7, 6, 0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3};
		RawMesh bad;
		bad.vertices.resize(2);
		bad.indices = {0, 1, 7};
		renderer.setDynamicMeshes({quad, bad});
		CHECK(renderer.lastStatus() == PHX_ERR_INVALID_SCENE && renderer.dynamicMeshes().empty(),
		      "a mesh with an out-of-range index is refused");
		renderer.setDynamicMeshes({quad});
		renderer.addDynamicMesh(box);
		CHECK(renderer.lastStatus() == PHX_OK && renderer.dynamicMeshes().size() == 2,
		      "dynamic meshes: %s", renderer.lastError().c_str());
		std::vector<DynObject> objs(4);
		const float s0 = 0.3f, s1 = 0.08f;
		// column-major: transform[4 * column + row]
		objs[0].meshIndex = 0;
		objs[0].transform[0] = s0, objs[0].transform[5] = 1.0f, objs[0].transform[10] = s0;
		objs[0].transform[12] = 0.35f, objs[0].transform[13] = 0.12f, objs[0].transform[14] = 0.3f;
		objs[1].meshIndex = 1;
		objs[1].transform[0] = s1, objs[1].transform[5] = 2 * s1, objs[1].transform[10] = s1;
		objs[1].transform[12] = 0.5f, objs[1].transform[13] = 0.05f, objs[1].transform[14] = 0.5f;
		objs[2] = objs[1];
		objs[2].transform[12] = 0.62f;
		objs[3].meshIndex = 7;  // no such mesh: skipped
		PhxRenderer *dyn = nullptr, *still = nullptr;
		CHECK(phx_renderer_create(ctx, w, h, &dyn) == PHX_OK, "dyn renderer");
		CHECK(phx_renderer_create(ctx, w, h, &still) == PHX_OK, "renderer without instances");
		renderer.resetTaa();
		for (uint32_t f = 0; f < 3; f++) {
			objs[1].transform[13] += 0.01f;  // one object moves
			RenderResult res = renderer.render(objs, {});
			CHECK(res.colorDevice, "dynamic render %u: %s", f, renderer.lastError().c_str());
			CHECK(phx_copy_to_host(ctx, a.data(), res.colorDevice, a.size() * 4) == PHX_OK, "copy");
			std::vector<PhxInstance> inst(3);
			for (int i = 0; i < 3; i++) {
				inst[i].mesh = renderer.dynamicMeshes()[objs[i].meshIndex];
				for (int r = 0; r < 3; r++)
					for (int c = 0; c < 4; c++)
						inst[i].objectToWorld[4 * r + c] = objs[i].transform[4 * c + r];
			}
			PhxInstanceSet* set = nullptr;
			CHECK(phx_instances_build(ctx, inst.data(), 3, &set) == PHX_OK, "instances");
			CHECK(phx_renderer_set_instances(dyn, set) == PHX_OK, "set instances");
			const uint32_t flags = PHX_FRAME_TAA | (f == 0 ? PHX_FRAME_TAA_RESET : 0u);
			CHECK(phx_render_frame(ctx, dyn, renderer.scene(), &def, f, 11, flags, b.data()) ==
			          PHX_OK, "direct dynamic frame %u: %s", f, phx_last_error());
			CHECK(std::memcmp(a.data(), b.data(), a.size() * 4) == 0, "dynamic frame %u differs", f);
			phx_instances_release(set);  // released while set: the next frame must refuse it
			CHECK(phx_render_frame(ctx, dyn, renderer.scene(), &def, f, 11, flags, nullptr) ==
			          PHX_ERR_INVALID_HANDLE, "a released instance set was used");
			CHECK(phx_renderer_set_instances(dyn, nullptr) == PHX_OK, "clear instances");
			std::vector<float> c(a.size());
			CHECK(phx_render_frame(ctx, still, renderer.scene(), &def, f, 11, flags, c.data()) ==
			          PHX_OK, "frame without instances");
			CHECK(std::memcmp(a.data(), c.data(), a.size() * 4) != 0, "instances had no effect");
		}
		phx_renderer_release(dyn);
		phx_renderer_release(still);
		renderer.render({}, {});  // no objects: back to the static scene alone
		CHECK(renderer.lastStatus() == PHX_OK && renderer.instances() == nullptr, "no objects");
	}

	// Material textures: setMaterialsAndTextures + addTexture give the frames of a scene uploaded
	// and textured through the C ABI, which differ from the untextured scene's frames.
	std::printf("hiptracer_check: textures\n");
	RawImage rgba;
	rgba.dim = vec2i{8, 4};
	rgba.bytesPerPixel = 4;
	rgba.pitch = 32;
	rgba.imgData.resize(128);
	for (size_t i = 0; i < rgba.imgData.size(); i++) rgba.imgData[i] = uint8_t(i * 37 + 11);
	RawImage gray;
	gray.dim = vec2i{3, 3};
	gray.bytesPerPixel = 1;
	gray.pitch = 4;  // one padding byte per row
	gray.imgData.resize(12);
	for (size_t i = 0; i < gray.imgData.size(); i++) gray.imgData[i] = uint8_t(200 - i * 13);
	std::vector<PhxMaterial> tmats = mats;
	for (PhxMaterial& m : tmats) {
		m.iData[0] = 0;
		m.iData[1] = 1;
	}
	renderer.setMaterialsAndTextures(tmats, {rgba});
	CHECK(renderer.lastStatus() == PHX_OK, "textures: %s", renderer.lastError().c_str());
	renderer.addTexture(gray);
	CHECK(renderer.lastStatus() == PHX_OK, "addTexture: %s", renderer.lastError().c_str());
	RawImage bad = gray;
	bad.bytesPerPixel = 3;
	renderer.addTexture(bad);
	CHECK(renderer.lastStatus() == PHX_ERR_INVALID_ARGUMENT, "3-byte texture accepted");
	const BVH& sb = scene.bvh;
	PhxScene* textured = nullptr;
	PhxScene* plain = nullptr;
	CHECK(phx_scene_upload(ctx, sb.nodes.data(), uint32_t(sb.nodes.size()),
	                       sb.triangleVerts.data(), sb.triangleDatas.data(),
	                       uint32_t(sb.triangleVerts.size()), tmats.data(), numMats,
	                       scene.sphereLights.data(), numLights, &textured) == PHX_OK, "upload");
	CHECK(phx_scene_upload(ctx, sb.nodes.data(), uint32_t(sb.nodes.size()),
	                       sb.triangleVerts.data(), sb.triangleDatas.data(),
	                       uint32_t(sb.triangleVerts.size()), tmats.data(), numMats,
	                       scene.sphereLights.data(), numLights, &plain) == PHX_OK, "upload");
	const PhxTexture tx[2] = {{rgba.imgData.data(), 8, 4, 4, 32}, {gray.imgData.data(), 3, 3, 1, 4}};
	CHECK(phx_scene_set_textures(ctx, textured, tx, 2) == PHX_OK, "set textures");
	std::vector<float> c(a.size());
	PhxRenderer* other = nullptr;  // its own TAA history
	CHECK(phx_renderer_create(ctx, w, h, &other) == PHX_OK, "second renderer");
	for (uint32_t f = 0; f < 2; f++) {
		RenderResult res = renderer.render({}, {});
		CHECK(res.colorDevice, "textured render %u: %s", f, renderer.lastError().c_str());
		CHECK(phx_copy_to_host(ctx, a.data(), res.colorDevice, a.size() * 4) == PHX_OK, "copy");
		const uint32_t flags = PHX_FRAME_TAA | (f == 0 ? PHX_FRAME_TAA_RESET : 0u);
		CHECK(phx_render_frame(ctx, direct, textured, &def, f, 11, flags, b.data()) == PHX_OK,
		      "direct textured frame %u: %s", f, phx_last_error());
		CHECK(std::memcmp(a.data(), b.data(), a.size() * 4) == 0, "textured frame %u differs", f);
		CHECK(phx_render_frame(ctx, other, plain, &def, f, 11, flags, c.data()) == PHX_OK,
		      "direct untextured frame %u: %s", f, phx_last_error());
		CHECK(std::memcmp(a.data(), c.data(), a.size() * 4) != 0, "textures had no effect");
	}
	phx_scene_release(textured);
	phx_scene_release(plain);
	phx_renderer_release(other);
	phx_renderer_release(direct);

	// Kernel-level API: primary rays on the device, the three launch functions.
	std::printf("hiptracer_check: launch functions\n");
	const uint32_t nr = uint32_t(w) * h;
	void *rays = nullptr, *hitsA = nullptr, *hitsB = nullptr, *light = nullptr;
	CHECK(phx_device_alloc(ctx, nr * sizeof(PhxRayIn), &rays) == PHX_OK, "alloc");
	CHECK(phx_device_alloc(ctx, nr * sizeof(PhxRayHit), &hitsA) == PHX_OK, "alloc");
	CHECK(phx_device_alloc(ctx, nr * sizeof(PhxRayHit), &hitsB) == PHX_OK, "alloc");
	CHECK(phx_device_alloc(ctx, nr, &light) == PHX_OK, "alloc");
	CHECK(phx_gen_primary_rays(ctx, &def, w, h, 1e-4f, 3.0e38f, PHX_RAYS_TILED_8X8,
	                           static_cast<PhxRayIn*>(rays), PHX_MEM_DEVICE) == PHX_OK, "rays");
	RayCastKernelInput in;
	in.scene = renderer.scene();
	in.numRays = nr;
	in.rays = static_cast<const PhxRayIn*>(rays);
	CHECK(launchRayCastKernel(ctx, in, static_cast<PhxRayHit*>(hitsA)) == PHX_OK, "persistent");
	CHECK(launchRayCastNoPersistenceKernel(ctx, in, static_cast<PhxRayHit*>(hitsB)) == PHX_OK,
	      "one ray per lane");
	std::vector<PhxRayIn> hostRays(nr);
	std::vector<PhxRayHit> ha(nr), hb(nr), hc(nr);
	CHECK(phx_copy_to_host(ctx, hostRays.data(), rays, nr * sizeof(PhxRayIn)) == PHX_OK, "copy");
	CHECK(phx_copy_to_host(ctx, ha.data(), hitsA, nr * sizeof(PhxRayHit)) == PHX_OK, "copy");
	CHECK(phx_copy_to_host(ctx, hb.data(), hitsB, nr * sizeof(PhxRayHit)) == PHX_OK, "copy");
	CHECK(phx_cast_closest(ctx, renderer.scene(), hostRays.data(), nr, hc.data(), 0) == PHX_OK,
	      "host cast");
	CHECK(std::memcmp(ha.data(), hc.data(), nr * sizeof(PhxRayHit)) == 0, "launchRayCastKernel");
	CHECK(std::memcmp(hb.data(), hc.data(), nr * sizeof(PhxRayHit)) == 0,
	      "launchRayCastNoPersistenceKernel");
	uint32_t hitsFound = 0;
	for (const PhxRayHit& hh : hc) hitsFound += hh.triangleIndex != 0xFFFFFFFFu;
	CHECK(hitsFound > nr / 2, "only %u of %u primary rays hit", hitsFound, nr);
	// shadow rays: towards a point above the scene from every hit
	std::vector<PhxRayIn> shadow(nr);
	for (uint32_t i = 0; i < nr; i++) {
		PhxRayIn& s = shadow[i];
		const PhxRayIn& r = hostRays[i];
		const float t = hc[i].triangleIndex == 0xFFFFFFFFu ? 1.0f : hc[i].t;
		float to[3];
		for (int k = 0; k < 3; k++) s.origin[k] = r.origin[k] + r.dir[k] * t;
		to[0] = 0.3f - s.origin[0];
		to[1] = 2.0f - s.origin[1];
		to[2] = 0.4f - s.origin[2];
		const float l = std::sqrt(to[0] * to[0] + to[1] * to[1] + to[2] * to[2]);
		for (int k = 0; k < 3; k++) s.dir[k] = to[k] / l;
		s.minDist = 1e-3f;
		s.maxDist = l - 1e-3f;
	}
	CHECK(phx_copy_to_device(ctx, rays, shadow.data(), nr * sizeof(PhxRayIn)) == PHX_OK, "copy");
	CHECK(launchShadowRayCastKernel(ctx, in, static_cast<bool*>(light)) == PHX_OK, "shadow");
	std::vector<uint8_t> la(nr), lb(nr);
	CHECK(phx_copy_to_host(ctx, la.data(), light, nr) == PHX_OK, "copy");
	CHECK(phx_cast_anyhit(ctx, renderer.scene(), shadow.data(), nr, lb.data(), 0) == PHX_OK,
	      "host anyhit");
	CHECK(la == lb, "launchShadowRayCastKernel");
	for (void* p : {rays, hitsA, hitsB, light}) phx_device_free(ctx, p);

	if (failures)
		std::printf("hiptracer_check: %d check(s) failed\n", failures);
	else
		std::printf("hiptracer_check: OK\n");
	return failures;
}
