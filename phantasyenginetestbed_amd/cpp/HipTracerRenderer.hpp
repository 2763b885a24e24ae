// C++ host side above the phx C ABI, shaped like the reference's renderer interface so engine code
// written against it switches backends without changes:
//   BaseRenderer            rendering/BaseRenderer.hpp:43-96
//   CudaTracerRenderer      cuda_ray_tracer/CudaTracerRenderer.hpp:14-58
//   RayCastKernelInput and launchRayCastKernel / launchRayCastNoPersistenceKernel /
//   launchShadowRayCastKernel   cuda_ray_tracer/kernels/RayCastKernel.hpp:47-64
// The engine's own types (sfz vectors, DynArray, RawImage, RawMesh, DynObject, ViewFrustum) are
// replaced by the minimal structs below; the byte layouts of the scene arrays are the reference's.
//
// Error behaviour differs on purpose: where the reference prints CUDA errors and continues
// (CudaHelpers.hpp:20-25) or terminates on unimplemented calls (CudaTracerRenderer.cpp:241-249,
// :412-415), this renderer records the failure (lastStatus / lastError) and leaves its state
// unchanged.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phx.h"

namespace phe_hip {

struct vec2i {
	int32_t x = 0, y = 0;
};

struct vec3 {
	float x = 0.0f, y = 0.0f, z = 0.0f;
};

// ViewFrustum's camera state used by the tracer: position, direction, up, vertical FOV (degrees).
struct ViewFrustum {
	vec3 pos;
	vec3 dir{0.0f, 0.0f, 1.0f};
	vec3 up{0.0f, 1.0f, 0.0f};
	float verticalFovDeg = 60.0f;
};

// ray_tracer_common/BVH.hpp:19-26 and level/StaticScene.hpp:22-26.
struct BVH {
	std::vector<PhxBVHNode> nodes;
	std::vector<PhxTriangleVertices> triangleVerts;
	std::vector<PhxTriangleData> triangleDatas;
};

struct StaticScene {
	BVH bvh;
	std::vector<PhxSphereLight> sphereLights;
};

// rendering/RawImage.hpp:21-29: dim.y rows of `pitch` bytes, bytesPerPixel bytes a texel.
struct RawImage {
	std::vector<uint8_t> imgData;
	vec2i dim;
	uint32_t bytesPerPixel = 0;
	uint32_t pitch = 0;
};
// rendering/RawMesh.hpp:21-37: vertices, a material index per vertex, three indices a triangle.
struct Vertex {
	vec3 pos;
	vec3 normal;
	float uv[2] = {0.0f, 0.0f};
};
struct RawMesh {
	std::vector<Vertex> vertices;
	std::vector<uint16_t> materialIndices;
	std::vector<uint32_t> indices;
};

// level/DynObject.hpp:16-20: a mesh of setDynamicMeshes placed by `transform`, stored like
// sfz::mat4 (column-major, elements[column][row]; world = transform * (p, 1)).
struct DynObject {
	uint32_t meshIndex = ~0u;
	float transform[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
	vec3 velocity;
};

// RenderResult (BaseRenderer.hpp:29-41): the reference returns GL texture names; this backend
// returns the device address of the RGBA32F output (row-major, renderedRes.x * renderedRes.y).
struct RenderResult {
	vec2i renderedRes;
	const float* colorDevice = nullptr;
};

class BaseRenderer {
public:
	virtual ~BaseRenderer() noexcept {}

	virtual void setMaterialsAndTextures(const std::vector<PhxMaterial>& materials,
	                                     const std::vector<RawImage>& textures) noexcept = 0;
	virtual void addTexture(const RawImage& texture) noexcept = 0;
	virtual void addMaterial(const PhxMaterial& material) noexcept = 0;
	virtual void setStaticScene(const StaticScene& staticScene) noexcept = 0;
	virtual void setDynamicMeshes(const std::vector<RawMesh>& meshes) noexcept = 0;
	virtual void addDynamicMesh(const RawMesh& mesh) noexcept = 0;
	virtual RenderResult render(const std::vector<DynObject>& objects,
	                            const std::vector<PhxSphereLight>& lights) noexcept = 0;

	void updateCamera(const ViewFrustum& camera) noexcept { mCamera = camera; }
	vec2i targetResolution() const noexcept { return mTargetResolution; }
	void setTargetResolution(vec2i targetResolution) noexcept
	{
		mTargetResolution = targetResolution;
		targetResolutionUpdated();
	}

protected:
	virtual void targetResolutionUpdated() noexcept = 0;

	ViewFrustum mCamera;
	vec2i mTargetResolution;
};

// CudaTracerRenderer's replacement. The static scene's sphere lights are the lights rendered (as
// in the reference, whose render() ignores its `lights` argument for the static path). Each
// render() is one wavefront frame; with TAA on (the default, as GameScreen enables it) frames
// are jittered by Halton(2,3) and resolved into an FP32 history. Dynamic meshes (left untraced by
// the reference's CUDA backend, CudaTracerRenderer.cpp:317-327, :412-415) each get an
// object-space SBVH once; every render() places the DynObjects' meshes as instances (a top tree
// rebuilt per frame) that the frame traces besides the static scene. Their normals take the
// inverse transpose of the object transform (the reference's createDynamicBvh applies the plain
// transform, DynamicBVH.cpp:152-154).
class HipTracerRenderer final : public BaseRenderer {
public:
	HipTracerRenderer(const HipTracerRenderer&) = delete;
	HipTracerRenderer& operator=(const HipTracerRenderer&) = delete;

	explicit HipTracerRenderer(int device = 0) noexcept;
	~HipTracerRenderer() noexcept override;

	void setMaterialsAndTextures(const std::vector<PhxMaterial>& materials,
	                             const std::vector<RawImage>& textures) noexcept override;
	void addTexture(const RawImage& texture) noexcept override;
	void addMaterial(const PhxMaterial& material) noexcept override;
	void setStaticScene(const StaticScene& staticScene) noexcept override;
	void setDynamicMeshes(const std::vector<RawMesh>& meshes) noexcept override;
	void addDynamicMesh(const RawMesh& mesh) noexcept override;
	RenderResult render(const std::vector<DynObject>& objects,
	                    const std::vector<PhxSphereLight>& lights) noexcept override;

	// Frame options: TAA (jitter + resolve) and the random seed; resetTaa restarts the history.
	void setTaa(bool enabled) noexcept { mTaa = enabled; }
	void resetTaa() noexcept { mFrame = 0; }
	void setSeed(uint64_t seed) noexcept { mSeed = seed; }
	uint32_t frameIndex() const noexcept { return mFrame; }

	phx_status lastStatus() const noexcept { return mStatus; }
	const std::string& lastError() const noexcept { return mError; }
	PhxContext* context() const noexcept { return mCtx; }
	const PhxScene* scene() const noexcept { return mScene; }
	// The uploaded dynamic meshes (object space) and the instance set of the last render().
	const std::vector<PhxScene*>& dynamicMeshes() const noexcept { return mMeshes; }
	const PhxInstanceSet* instances() const noexcept { return mInstances; }

protected:
	void targetResolutionUpdated() noexcept override;

private:
	bool record(phx_status st) noexcept;
	void uploadScene() noexcept;
	void uploadTextures(PhxScene* scene) noexcept;
	PhxScene* uploadMesh(const RawMesh& mesh) noexcept;

	PhxContext* mCtx = nullptr;
	PhxScene* mScene = nullptr;
	PhxRenderer* mRenderer = nullptr;
	std::vector<PhxScene*> mMeshes;
	PhxInstanceSet* mInstances = nullptr;
	std::vector<PhxMaterial> mMaterials;
	std::vector<RawImage> mTextures;
	StaticScene mStatic;
	bool mHaveScene = false;
	bool mTaa = true;
	uint64_t mSeed = 1;
	uint32_t mFrame = 0;
	phx_status mStatus = PHX_OK;
	std::string mError;
};

// RayCastKernelInput (RayCastKernel.hpp:47-52): the node and triangle textures become the
// uploaded scene; rays and results are device pointers, as in the reference.
struct RayCastKernelInput {
	const PhxScene* scene = nullptr;
	uint32_t numRays = 0;
	const PhxRayIn* rays = nullptr;
};

// launchRayCastKernel / launchRayCastNoPersistenceKernel / launchShadowRayCastKernel
// (RayCastKernel.hpp:54-64). Like the reference they return after the launch completed; the
// device properties argument becomes the context.
phx_status launchRayCastKernel(PhxContext* ctx, const RayCastKernelInput& input,
                               PhxRayHit* rayResults) noexcept;
phx_status launchRayCastNoPersistenceKernel(PhxContext* ctx, const RayCastKernelInput& input,
                                            PhxRayHit* rayResults) noexcept;
phx_status launchShadowRayCastKernel(PhxContext* ctx, const RayCastKernelInput& input,
                                     bool* inLight) noexcept;

}  // namespace phe_hip
