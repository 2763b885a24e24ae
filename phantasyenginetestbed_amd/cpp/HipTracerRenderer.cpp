// HipTracerRenderer over the phx C ABI.
#include "HipTracerRenderer.hpp"

#include <cmath>

namespace phe_hip {

HipTracerRenderer::HipTracerRenderer(int device) noexcept
{
	record(phx_create(device, &mCtx));
}

HipTracerRenderer::~HipTracerRenderer() noexcept
{
	if (mRenderer) phx_renderer_release(mRenderer);
	if (mInstances) phx_instances_release(mInstances);
	for (PhxScene* m : mMeshes) phx_scene_release(m);
	if (mScene) phx_scene_release(mScene);
	if (mCtx) phx_destroy(mCtx);
}

bool HipTracerRenderer::record(phx_status st) noexcept
{
	mStatus = st;
	mError = st == PHX_OK ? std::string() : std::string(phx_last_error());
	return st == PHX_OK;
}

void HipTracerRenderer::uploadScene() noexcept
{
	if (!mCtx || !mHaveScene) return;
	PhxScene* s = nullptr;
	const BVH& b = mStatic.bvh;
	if (!record(phx_scene_upload(mCtx, b.nodes.data(), uint32_t(b.nodes.size()),
	                             b.triangleVerts.data(),
	                             b.triangleDatas.empty() ? nullptr : b.triangleDatas.data(),
	                             uint32_t(b.triangleVerts.size()),
	                             mMaterials.empty() ? nullptr : mMaterials.data(),
	                             uint32_t(mMaterials.size()),
	                             mStatic.sphereLights.empty() ? nullptr : mStatic.sphereLights.data(),
	                             uint32_t(mStatic.sphereLights.size()), &s)))
		return;
	uploadTextures(s);
	if (mStatus != PHX_OK) {
		phx_scene_release(s);
		return;
	}
	if (mScene) phx_scene_release(mScene);
	mScene = s;
	mFrame = 0;
}

void HipTracerRenderer::uploadTextures(PhxScene* scene) noexcept
{
	std::vector<PhxTexture> t(mTextures.size());
	for (size_t i = 0; i < t.size(); i++) {
		const RawImage& img = mTextures[i];
		t[i] = PhxTexture{img.imgData.data(), uint32_t(img.dim.x), uint32_t(img.dim.y),
		                  img.bytesPerPixel, img.pitch};
		// An image whose rows do not fit its buffer is rejected like an empty one.
		if (size_t(img.pitch) * uint32_t(img.dim.y) > img.imgData.size()) t[i].data = nullptr;
	}
	record(phx_scene_set_textures(mCtx, scene, t.empty() ? nullptr : t.data(),
	                              uint32_t(t.size())));
}

// The reference's CUDA backend uploads the images as bindless textures here
// (CudaTracerRenderer.cpp:192-239); the scene is re-uploaded with them.
void HipTracerRenderer::setMaterialsAndTextures(const std::vector<PhxMaterial>& materials,
                                                const std::vector<RawImage>& textures) noexcept
{
	mMaterials = materials;
	mTextures = textures;
	uploadScene();
}

// Unimplemented in the reference (CudaTracerRenderer.cpp:241-244); here the texture list grows
// by one and the current scene takes the new list.
void HipTracerRenderer::addTexture(const RawImage& texture) noexcept
{
	mTextures.push_back(texture);
	if (mScene) {
		uploadTextures(mScene);
		if (mStatus != PHX_OK) mTextures.pop_back();
	}
}

void HipTracerRenderer::addMaterial(const PhxMaterial& material) noexcept
{
	mMaterials.push_back(material);
	uploadScene();
}

void HipTracerRenderer::setStaticScene(const StaticScene& staticScene) noexcept
{
	mStatic = staticScene;
	mHaveScene = true;
	uploadScene();
}

// One mesh as an object-space scene: triangle i is vertices[indices[3i..3i+2]] with their normals
// and uvs, material materialIndices[indices[3i]] (as createDynamicBvh, DynamicBVH.cpp:127-160,
// reads them), under its own SBVH. Null (and the error recorded) for an unusable mesh.
PhxScene* HipTracerRenderer::uploadMesh(const RawMesh& mesh) noexcept
{
	const size_t nt = mesh.indices.size() / 3;
	if (nt == 0 || mesh.indices.size() % 3 != 0) {
		mStatus = PHX_ERR_INVALID_SCENE;
		mError = "HipTracerRenderer: a dynamic mesh needs a whole number of triangles";
		return nullptr;
	}
	std::vector<PhxTriangleVertices> tris(nt);
	std::vector<PhxTriangleData> datas(nt);
	for (size_t t = 0; t < nt; t++) {
		const Vertex* v[3];
		for (int k = 0; k < 3; k++) {
			const uint32_t idx = mesh.indices[3 * t + k];
			if (idx >= mesh.vertices.size()) {
				mStatus = PHX_ERR_INVALID_SCENE;
				mError = "HipTracerRenderer: a dynamic mesh index is out of range";
				return nullptr;
			}
			v[k] = &mesh.vertices[idx];
		}
		float* p[3] = {tris[t].v0, tris[t].v1, tris[t].v2};
		float* n[3] = {datas[t].n0, datas[t].n1, datas[t].n2};
		float* uv[3] = {datas[t].uv0, datas[t].uv1, datas[t].uv2};
		for (int k = 0; k < 3; k++) {
			p[k][0] = v[k]->pos.x, p[k][1] = v[k]->pos.y, p[k][2] = v[k]->pos.z, p[k][3] = 0.0f;
			n[k][0] = v[k]->normal.x, n[k][1] = v[k]->normal.y, n[k][2] = v[k]->normal.z;
			uv[k][0] = v[k]->uv[0], uv[k][1] = v[k]->uv[1];
		}
		const uint32_t first = mesh.indices[3 * t];
		datas[t].materialIndex = first < mesh.materialIndices.size() ? mesh.materialIndices[first] : 0u;
	}
	PhxBVH* bvh = nullptr;
	if (!record(phx_bvh_build(tris.data(), datas.data(), uint32_t(nt), 0, &bvh))) return nullptr;
	const PhxBVHNode* nodes = nullptr;
	const PhxTriangleVertices* v = nullptr;
	const PhxTriangleData* d = nullptr;
	uint32_t nn = 0, nr = 0;
	PhxScene* s = nullptr;
	if (record(phx_bvh_arrays(bvh, &nodes, &nn, &v, &d, &nr)))
		record(phx_scene_upload(mCtx, nodes, nn, v, d, nr, nullptr, 0, nullptr, 0, &s));
	phx_bvh_release(bvh);
	return mStatus == PHX_OK ? s : nullptr;
}

// The reference's CUDA backend leaves this a TODO (CudaTracerRenderer.cpp:317-327): here every
// mesh gets its SBVH once; render() places them.
void HipTracerRenderer::setDynamicMeshes(const std::vector<RawMesh>& meshes) noexcept
{
	if (!mCtx) return;
	std::vector<PhxScene*> up;
	for (const RawMesh& m : meshes) {
		PhxScene* s = uploadMesh(m);
		if (!s) {
			for (PhxScene* u : up) phx_scene_release(u);
			return;  // error recorded, previous meshes kept
		}
		up.push_back(s);
	}
	for (PhxScene* m : mMeshes) phx_scene_release(m);  // instance sets keep what they still use
	mMeshes = up;
	record(PHX_OK);
}

// Unimplemented in the reference (CudaTracerRenderer.cpp:412-415).
void HipTracerRenderer::addDynamicMesh(const RawMesh& mesh) noexcept
{
	if (!mCtx) return;
	PhxScene* s = uploadMesh(mesh);
	if (s) mMeshes.push_back(s);
}

void HipTracerRenderer::targetResolutionUpdated() noexcept
{
	if (!mCtx) return;
	PhxRenderer* r = nullptr;
	if (!record(phx_renderer_create(mCtx, uint32_t(mTargetResolution.x),
	                                uint32_t(mTargetResolution.y), &r)))
		return;
	if (mRenderer) phx_renderer_release(mRenderer);
	mRenderer = r;
	mFrame = 0;
}

RenderResult HipTracerRenderer::render(const std::vector<DynObject>& objects,
                                       const std::vector<PhxSphereLight>&) noexcept
{
	RenderResult res;
	res.renderedRes = mTargetResolution;
	if (!mCtx || !mScene || !mRenderer) {
		mStatus = PHX_ERR_INVALID_ARGUMENT;
		mError = "HipTracerRenderer::render: needs a static scene and a target resolution";
		return res;
	}
	// The objects' meshes as this frame's instances (objects without a mesh are skipped).
	std::vector<PhxInstance> inst;
	for (const DynObject& o : objects) {
		if (o.meshIndex >= mMeshes.size()) continue;
		PhxInstance in;
		in.mesh = mMeshes[o.meshIndex];
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 4; c++) in.objectToWorld[4 * r + c] = o.transform[4 * c + r];
		inst.push_back(in);
	}
	PhxInstanceSet* set = nullptr;
	if (!inst.empty() &&
	    !record(phx_instances_build(mCtx, inst.data(), uint32_t(inst.size()), &set)))
		return res;
	if (!record(phx_renderer_set_instances(mRenderer, set))) {
		if (set) phx_instances_release(set);
		return res;
	}
	if (mInstances) phx_instances_release(mInstances);
	mInstances = set;
	// generateCameraDef (GenerateRays.cpp:12-35) with the frustum's vertical FOV in radians.
	PhxCameraDef cam;
	const float pos[3] = {mCamera.pos.x, mCamera.pos.y, mCamera.pos.z};
	const float dir[3] = {mCamera.dir.x, mCamera.dir.y, mCamera.dir.z};
	const float up[3] = {mCamera.up.x, mCamera.up.y, mCamera.up.z};
	const float fov = mCamera.verticalFovDeg * (3.14159265358979323846f / 180.0f);
	if (!record(phx_camera_def(pos, dir, up, fov, float(mTargetResolution.x),
	                           float(mTargetResolution.y), &cam)))
		return res;
	uint32_t flags = 0;
	if (mTaa) flags |= PHX_FRAME_TAA | (mFrame == 0 ? PHX_FRAME_TAA_RESET : 0u);
	if (!record(phx_render_frame(mCtx, mRenderer, mScene, &cam, mFrame, mSeed, flags, nullptr)))
		return res;
	mFrame++;
	const void* p = nullptr;
	if (!record(phx_renderer_buffer(mRenderer, PHX_BUFFER_OUTPUT, &p, nullptr))) return res;
	res.colorDevice = static_cast<const float*>(p);
	return res;
}

phx_status launchRayCastKernel(PhxContext* ctx, const RayCastKernelInput& in,
                               PhxRayHit* rayResults) noexcept
{
	return phx_cast_closest(ctx, in.scene, in.rays, in.numRays, rayResults, PHX_MEM_DEVICE);
}

phx_status launchRayCastNoPersistenceKernel(PhxContext* ctx, const RayCastKernelInput& in,
                                            PhxRayHit* rayResults) noexcept
{
	return phx_cast_closest(ctx, in.scene, in.rays, in.numRays, rayResults,
	                        PHX_MEM_DEVICE | PHX_KERNEL_ONE_RAY_PER_LANE);
}

phx_status launchShadowRayCastKernel(PhxContext* ctx, const RayCastKernelInput& in,
                                     bool* inLight) noexcept
{
	static_assert(sizeof(bool) == 1, "inLight is one byte per ray");
	return phx_cast_anyhit(ctx, in.scene, in.rays, in.numRays, reinterpret_cast<uint8_t*>(inLight),
	                       PHX_MEM_DEVICE);
}

}  // namespace phe_hip
