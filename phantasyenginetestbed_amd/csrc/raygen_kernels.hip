// Ray generation kernels.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "rng.hpp"
#include "shading.hpp"

namespace phx {
namespace {

// Pinhole primary ray of pixel (x, y): calculatePrimaryRayDir (RayCastKernel.cu:891-897) ==
// the CPU renderer's formula (CPUTracerRenderer.cpp:93-97): c = (x, y) / res * 2 - 1,
// dir = normalize((camDir + cx * dX) + cy * dY), normalize(v) = v / sqrt((x*x + y*y) + z*z).
// Compiled without contraction and with correctly rounded sqrt/div, so it matches the CPU bits.
// A jitter (jx, jy) in pixels moves the sample to (x + jx, y + jy); (0, 0) is the formula above.
__global__ __launch_bounds__(256) void k_gen_primary(PhxCameraDef cam, uint32_t w, uint32_t h,
                                                     float jx, float jy, float tMin, float tMax,
                                                     uint32_t order, float4* __restrict__ rays,
                                                     const FrameParams* __restrict__ fp)
{
	const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
	const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
	if (x >= w || y >= h) return;
	if (fp) {
		cam = fp->cam;
		jx = fp->jx;
		jy = fp->jy;
	}
	const float cx = ((float(x) + jx) / float(w)) * 2.0f - 1.0f;
	const float cy = ((float(y) + jy) / float(h)) * 2.0f - 1.0f;
	float d[3];
	for (int k = 0; k < 3; k++) {
		float a = cam.dX[k] * cx;
		float s = cam.dir[k] + a;
		float b = cam.dY[k] * cy;
		d[k] = s + b;
	}
	float dd = d[0] * d[0] + d[1] * d[1];
	dd = dd + d[2] * d[2];
	const float len = sqrtf(dd);
	d[0] = d[0] / len;
	d[1] = d[1] / len;
	d[2] = d[2] / len;

	size_t id;
	if (order == PHX_RAYS_TILED_8X8) {
		// getRayIdx (CudaDeviceHelpers.cuh:49-62) with 8x8 tiles.
		const uint32_t tile = (y / 8u) * (w / 8u) + (x / 8u);
		id = size_t(tile) * 64u + (y % 8u) * 8u + (x % 8u);
	} else {
		id = size_t(y) * w + x;
	}
	rays[2 * id] = make_float4(cam.origin[0], cam.origin[1], cam.origin[2], tMin);
	rays[2 * id + 1] = make_float4(d[0], d[1], d[2], tMax);
}

__device__ __forceinline__ void normalize3(float v[3])
{
	float d = v[0] * v[0] + v[1] * v[1];
	d = d + v[2] * v[2];
	const float l = sqrtf(d);
	v[0] = v[0] / l;
	v[1] = v[1] / l;
	v[2] = v[2] / l;
}

// Bounce ray of ray i: hit position o + t*d (InterpretRayHitKernel.cu:64-70), shading normal
// normalize(n0 + (n1-n0)u + (n2-n0)v) (BVHTraversal.hpp:259-262) turned to face the incoming ray,
// cosine-weighted direction (sampleLambertian + hemisphereToWorld, MaterialKernel.cu:206-246,
// with cos/sin(2 pi r1) from shading::turn_sincos), PCG32 stream pixel_rng(seed, frame, i),
// tMin 0.001, tMax FLT_MAX. A miss gets the dummy ray of MaterialKernel.cu:41-47.
__global__ __launch_bounds__(256) void k_gen_bounce(const float4* __restrict__ rays,
                                                    const uint4* __restrict__ hits,
                                                    const PhxTriangleData* __restrict__ datas,
                                                    uint32_t numTris, uint32_t n, uint64_t seed,
                                                    uint32_t frame, float4* __restrict__ out)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint4 h = hits[i];
	if (h.x >= numTris) {  // a miss (0xFFFFFFFF), or an index of another scene
		out[2 * size_t(i)] = make_float4(1000000.0f, 1000000.0f, 1000000.0f, 0.001f);
		out[2 * size_t(i) + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.001f);
		return;
	}
	const float4 ro = rays[2 * size_t(i)];
	const float4 rd = rays[2 * size_t(i) + 1];
	const float t = __uint_as_float(h.y), bu = __uint_as_float(h.z), bv = __uint_as_float(h.w);
	const float pos[3] = {ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t};
	const PhxTriangleData& d = datas[h.x];
	float nrm[3];
	for (int k = 0; k < 3; k++) {
		float a = (d.n1[k] - d.n0[k]) * bu;
		float b = (d.n2[k] - d.n0[k]) * bv;
		nrm[k] = (d.n0[k] + a) + b;
	}
	normalize3(nrm);
	float dn = nrm[0] * rd.x + nrm[1] * rd.y;
	dn = dn + nrm[2] * rd.z;
	if (dn > 0.0f) {
		nrm[0] = -nrm[0];
		nrm[1] = -nrm[1];
		nrm[2] = -nrm[2];
	}

	Pcg32 rng = pixel_rng(seed, frame, i);
	const float r1 = rng.uniform();
	const float r2 = rng.uniform();
	const shading::V3 dv = shading::sample_lambert(r1, r2, shading::v3(nrm[0], nrm[1], nrm[2]));
	const float dir[3] = {dv.x, dv.y, dv.z};
	out[2 * size_t(i)] = make_float4(pos[0], pos[1], pos[2], 0.001f);
	out[2 * size_t(i) + 1] = make_float4(dir[0], dir[1], dir[2], FLT_MAX);
}

}  // namespace

hipError_t launchGenBounceRays(const PhxRayIn* rays, const PhxRayHit* hits,
                               const PhxTriangleData* datas, uint32_t numTris, uint32_t n,
                               uint64_t seed, uint32_t frame, PhxRayIn* out, hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	const uint32_t grid = n / 256u + (n % 256u != 0u);
	hipLaunchKernelGGL(k_gen_bounce, dim3(grid), dim3(256), 0, stream,
	                   reinterpret_cast<const float4*>(rays), reinterpret_cast<const uint4*>(hits),
	                   datas, numTris, n, seed, frame, reinterpret_cast<float4*>(out));
	return hipGetLastError();
}

hipError_t launchGenPrimaryRays(const PhxCameraDef& cam, uint32_t w, uint32_t h, float jx,
                                float jy, float tMin, float tMax, uint32_t order, PhxRayIn* rays,
                                hipStream_t stream, const FrameParams* fp)
{
	dim3 grid((w + 15u) / 16u, (h + 15u) / 16u);
	hipLaunchKernelGGL(k_gen_primary, grid, dim3(256), 0, stream, cam, w, h, jx, jy, tMin, tMax,
	                   order, reinterpret_cast<float4*>(rays), fp);
	return hipGetLastError();
}

}  // namespace phx
