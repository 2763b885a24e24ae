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

void HipTracerRenderer::setDynamicMeshes(const std::vector<RawMesh>&) noexcept
{
	mStatus = PHX_ERR_INVALID_ARGUMENT;
	mError = "HipTracerRenderer::setDynamicMeshes: dynamic meshes are not supported";
}

void HipTracerRenderer::addDynamicMesh(const RawMesh&) noexcept
{
	mStatus = PHX_ERR_INVALID_ARGUMENT;
	mError = "HipTracerRenderer::addDynamicMesh: dynamic meshes are not supported";
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

RenderResult HipTracerRenderer::render(const std::vector<DynObject>&,
                                       const std::vector<PhxSphereLight>&) noexcept
{
	RenderResult res;
	res.renderedRes = mTargetResolution;
	if (!mCtx || !mScene || !mRenderer) {
		mStatus = PHX_ERR_INVALID_ARGUMENT;
		mError = "HipTracerRenderer::render: needs a static scene and a target resolution";
		return res;
	}
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
