// Wavefront frame kernels: interpret hits, shade + shadow rays per light, bounce-ray generation,
// shadow logic and the TAA resolve. Traversal itself is trace_kernels.hip.
//
// Records of full-resolution passes are indexed by pixel_id (8x8 tiles when both sizes are
// multiples of 8, as getRayIdx in CudaDeviceHelpers.cuh:49-62; row-major otherwise); images are
// row-major RGBA32F. Per-light arrays are light-major: record id + light * count.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "rng.hpp"
#include "shading.hpp"

namespace phx {
namespace {

using namespace shading;

constexpr uint32_t kBlock1D = 256;

__device__ __forceinline__ uint32_t pixel_id(uint32_t x, uint32_t y, uint32_t w, uint32_t h)
{
	if ((w % 8u) == 0u && (h % 8u) == 0u) {
		const uint32_t tile = (y / 8u) * (w / 8u) + (x / 8u);
		return tile * 64u + (y % 8u) * 8u + (x % 8u);
	}
	return y * w + x;
}

__device__ __forceinline__ Pcg32 frame_rng(uint64_t seed, uint32_t frame, uint32_t purpose,
                                           uint32_t id)
{
	return Pcg32::seeded(seed ^ (uint64_t(frame) << 32), (uint64_t(purpose) << 32) | id);
}

__device__ __forceinline__ void store_ray(float4* rays, size_t i, V3 o, float tMin, V3 d,
                                          float tMax)
{
	rays[2 * i] = make_float4(o.x, o.y, o.z, tMin);
	rays[2 * i + 1] = make_float4(d.x, d.y, d.z, tMax);
}

// setToDummyRay, MaterialKernel.cu:41-47.
__device__ __forceinline__ void store_dummy(float4* rays, size_t i)
{
	rays[2 * i] = make_float4(1000000.0f, 1000000.0f, 1000000.0f, 0.001f);
	rays[2 * i + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.001f);
}

__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// Texel coordinate along one axis of a wrapping, nearest-texel lookup
// (CPURayTracerRenderer::sampleImage, CPUTracerRenderer.cpp:160-180):
// floor(fmod(fmod(c * dim, dim) + dim, dim)), in [0, dim) for finite c; NaN or infinite c reads 0.
__device__ __forceinline__ uint32_t texel_coord(float c, uint32_t dim)
{
	const float d = float(dim);
	float s = fmodf(c * d, d);
	s = fmodf(s + d, d);
	const float f = floorf(s);
	return f >= 0.0f ? uint32_t(f) : 0u;
}

// First byte of the texel of texture `idx` at (u, v); its size in *bpp.
__device__ __forceinline__ const uint8_t* texel_at(const TexSet& tex, uint32_t idx, float u,
                                                   float v, uint32_t* bpp)
{
	const TexDesc d = tex.descs[idx];
	const uint32_t x = texel_coord(u, d.width), y = texel_coord(v, d.height);
	*bpp = d.bpp;
	return tex.texels + d.offset + size_t(y) * d.pitch + size_t(x) * d.bpp;
}

// interpretRayHitKernel, InterpretRayHitKernel.cu:58-121: position o + d*t, normal
// normalize((n0 + (n1-n0)u) + (n2-n0)v), uv interpolated the same way, material from the table
// (an index outside it reads albedo 1, roughness 1, metallic 0). A texture index in [0, count)
// replaces the constant: albedo from a 4-byte texel, rgb through the linearisation table, alpha
// c / 255 (:86-93; other texel sizes keep the constant, CPUTracerRenderer.cpp:200-206); roughness
// and metallic from byte 0 as c / 255 (:96-111). Record layout (48 B): (pos, roughness)
// (normal, metallic) (albedo, alpha), alpha negative for a miss.
// With instances (instIds not null): a hit of instance k reads that instance's mesh data and its
// interpolated normal is mapped to world space by the transpose of the world -> object matrix
// (the inverse transpose of the object -> world one) before normalisation.
__global__ __launch_bounds__(kBlock1D) void k_interpret(const float4* __restrict__ rays,
                                                        const uint4* __restrict__ hits, uint32_t n,
                                                        const PhxTriangleData* __restrict__ datas,
                                                        uint32_t numTris,
                                                        const uint32_t* __restrict__ instIds,
                                                        const DevInstance* __restrict__ inst,
                                                        uint32_t numInst,
                                                        const PhxMaterial* __restrict__ mats,
                                                        uint32_t numMats, TexSet tex,
                                                        float4* __restrict__ info)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint4 h = hits[i];
	const uint32_t k = instIds ? instIds[i] : 0xFFFFFFFFu;
	const DevInstance* in = k < numInst ? &inst[k] : nullptr;
	if (in) {
		datas = in->datas;
		numTris = datas ? in->numTris : 0u;
	}
	if (h.x >= numTris) {  // a miss (0xFFFFFFFF), or an index of another scene
		info[3 * size_t(i)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		info[3 * size_t(i) + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		info[3 * size_t(i) + 2] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
		return;
	}
	const float4 ro = rays[2 * size_t(i)];
	const float4 rd = rays[2 * size_t(i) + 1];
	const float t = __uint_as_float(h.y), bu = __uint_as_float(h.z), bv = __uint_as_float(h.w);
	const PhxTriangleData& d = datas[h.x];
	float nn[3];
	for (int k = 0; k < 3; k++) {
		const float a = (d.n1[k] - d.n0[k]) * bu;
		const float b = (d.n2[k] - d.n0[k]) * bv;
		nn[k] = (d.n0[k] + a) + b;
	}
	if (in) {
		const float* A = in->worldToObject;
		float w[3];
		for (int c = 0; c < 3; c++) {
			const float a = A[c] * nn[0] + A[4 + c] * nn[1];
			w[c] = a + A[8 + c] * nn[2];
		}
		nn[0] = w[0];
		nn[1] = w[1];
		nn[2] = w[2];
	}
	const V3 nrm = normalize(v3(nn[0], nn[1], nn[2]));
	float4 alb = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
	float rough = 1.0f, metal = 0.0f;
	if (d.materialIndex < numMats) {
		const PhxMaterial& m = mats[d.materialIndex];
		alb = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], m.albedo[3]);
		rough = m.fData2[0];
		metal = m.fData2[1];
		if (tex.count != 0u) {
			float uv[2];
			for (int k = 0; k < 2; k++) {
				const float a = (d.uv1[k] - d.uv0[k]) * bu;
				const float b = (d.uv2[k] - d.uv0[k]) * bv;
				uv[k] = (d.uv0[k] + a) + b;
			}
			uint32_t bpp;
			const int32_t ai = m.iData[0], ri = m.iData[1], mi = m.iData[2];
			if (ai >= 0 && uint32_t(ai) < tex.count) {
				const uint8_t* p = texel_at(tex, uint32_t(ai), uv[0], uv[1], &bpp);
				if (bpp == 4u) {
					const uchar4 c = *reinterpret_cast<const uchar4*>(p);
					alb = make_float4(tex.linear[c.x], tex.linear[c.y], tex.linear[c.z],
					                  float(c.w) / 255.0f);
				}
			}
			if (ri >= 0 && uint32_t(ri) < tex.count)
				rough = float(*texel_at(tex, uint32_t(ri), uv[0], uv[1], &bpp)) / 255.0f;
			if (mi >= 0 && uint32_t(mi) < tex.count)
				metal = float(*texel_at(tex, uint32_t(mi), uv[0], uv[1], &bpp)) / 255.0f;
		}
	}
	info[3 * size_t(i)] = make_float4(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t, rough);
	info[3 * size_t(i) + 1] = make_float4(nrm.x, nrm.y, nrm.z, metal);
	info[3 * size_t(i) + 2] = make_float4(alb.x, alb.y, alb.z, fabsf(alb.w));
}

// shadeHit (MaterialKernel.cu:49-114) of every record: per light the contribution
// mask * falloff * shade and a shadow ray towards a point on the light's disc (soft shadows,
// radius 0 = point light), or the dummy ray when the light has no static shadows or contributes
// nothing (every channel <= 1e-4). Primary records (rays == nullptr) view towards camPos with
// mask 1 (gBufferMaterialKernel, :119-140); bounce records view along -ray.dir with their path
// throughput (materialKernel, :145-184). Records without a hit get dummy rays and zero.
__global__ __launch_bounds__(kBlock1D) void k_shade(const float4* __restrict__ info,
                                                    const float4* __restrict__ rays,
                                                    const float4* __restrict__ throughput,
                                                    const PhxSphereLight* __restrict__ lights,
                                                    uint32_t numLights, uint32_t count,
                                                    const FrameParams* __restrict__ fp,
                                                    uint32_t purpose,
                                                    float4* __restrict__ contrib,
                                                    float4* __restrict__ shadow)
{
	const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
	if (id >= count) return;
	const uint64_t seed = fp->seed;
	const uint32_t frame = fp->frame;
	const float4 d1 = info[3 * size_t(id)];
	const float4 d2 = info[3 * size_t(id) + 1];
	const float4 d3 = info[3 * size_t(id) + 2];
	if (!(d3.w >= 0.0f)) {
		for (uint32_t i = 0; i < numLights; i++) {
			const size_t s = size_t(id) + size_t(i) * count;
			contrib[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			store_dummy(shadow, s);
		}
		return;
	}
	const V3 p = xyz(d1), n = xyz(d2), albedo = xyz(d3);
	const float rough = d1.w, metal = d2.w;
	V3 view, mask;
	if (rays) {
		const float4 rd = rays[2 * size_t(id) + 1];
		view = v3(-rd.x, -rd.y, -rd.z);
		mask = xyz(throughput[id]);
	} else {
		view = normalize(sub(v3(fp->cam.origin[0], fp->cam.origin[1], fp->cam.origin[2]), p));
		mask = v3(1.0f, 1.0f, 1.0f);
	}
	Pcg32 rng = frame_rng(seed, frame, purpose, id);
	for (uint32_t i = 0; i < numLights; i++) {
		const PhxSphereLight& L = lights[i];
		const V3 lp = v3(L.pos[0], L.pos[1], L.pos[2]);
		const V3 toL = sub(lp, p);
		const float dist = length(toL);
		const V3 l = divs(toL, dist);
		const float fo = falloff(dist, L.range);
		const V3 sh = shade(n, view, albedo, rough, metal, l, dist,
		                    v3(L.strength[0], L.strength[1], L.strength[2]), L.range);
		const V3 c = mul(scale(mask, fo), sh);
		const size_t s = size_t(id) + size_t(i) * count;
		contrib[s] = make_float4(c.x, c.y, c.z, 0.0f);
		const float eps = 0.0001f;
		const bool lit = c.x > eps || c.y > eps || c.z > eps;
		if (L.staticShadows && lit) {
			const V3 cu = fabsf(n.z) > 0.01f ? normalize(v3(0.0f, -n.z, n.y))
			                                 : normalize(v3(-n.y, n.x, 0.0f));
			const V3 cv = cross(cu, l);
			const float r1 = rng.uniform();
			const float r2 = rng.uniform();
			const float cd = L.radius * (2.0f * r2 - 1.0f);
			float sn, cs;
			turn_sincos(r1, sn, cs);
			const V3 off = add(scale(scale(cu, cs), cd), scale(scale(cv, sn), cd));
			const V3 diff = sub(add(lp, off), p);
			store_ray(shadow, s, p, 0.001f, normalize(diff), length(diff));
		} else {
			store_dummy(shadow, s);
		}
	}
}

// createSecondaryRaysKernel (MaterialKernel.cu:251-312) on the half-resolution grid: a random
// pixel of the 2x2 block, then diffuse (Lambert) or GGX-around-the-reflection by Russian roulette
// on roughness, throughput albedo * 0.6, tMin 1e-3; alpha <= 0.5 takes the reference's glass
// path. A block pixel without a hit gets the dummy ray and zero throughput.
__global__ __launch_bounds__(kBlock1D) void k_gen_secondary(const float4* __restrict__ gbuf,
                                                            uint32_t w, uint32_t h,
                                                            const FrameParams* __restrict__ fp,
                                                            float4* __restrict__ rays,
                                                            float4* __restrict__ throughput)
{
	const uint32_t hw = w / 2u, hh = h / 2u;
	const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
	const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
	if (x >= hw || y >= hh) return;
	const uint64_t seed = fp->seed;
	const uint32_t frame = fp->frame;
	const uint32_t hid = pixel_id(x, y, hw, hh);
	Pcg32 rng = frame_rng(seed, frame, 2u, hid);
	const uint32_t pix = rng.next() % 4u;
	const uint32_t fid = pixel_id(2u * x + pix % 2u, 2u * y + pix / 2u, w, h);
	const float4 d1 = gbuf[3 * size_t(fid)];
	const float4 d2 = gbuf[3 * size_t(fid) + 1];
	const float4 d3 = gbuf[3 * size_t(fid) + 2];
	if (!(d3.w >= 0.0f)) {
		store_dummy(rays, hid);
		throughput[hid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		return;
	}
	const V3 gp = xyz(d1), gn = xyz(d2), alb = xyz(d3);
	const float rough = d1.w;
	const V3 fromCam =
		normalize(sub(gp, v3(fp->cam.origin[0], fp->cam.origin[1], fp->cam.origin[2])));
	const float r1 = rng.uniform();
	const float r2 = rng.uniform();
	float tMin = 0.001f;
	V3 dir, tp;
	if (fabsf(d3.w) > 0.5f) {
		const float pick = rng.uniform();
		tp = scale(alb, 0.6f);
		if (pick < rough)
			dir = sample_lambert(r1, r2, gn);
		else
			dir = sample_ggx(r1, r2, rough * rough, reflect(fromCam, gn), gn);
	} else {
		dir = normalize(add(scale(fromCam, 0.8f), scale(v3(-gn.x, -gn.y, -gn.z), 0.2f)));
		tp = v3(1.0f, 1.0f, 1.0f);
		tMin = 0.2f;
	}
	store_ray(rays, hid, gp, tMin, dir, FLT_MAX);
	throughput[hid] = make_float4(tp.x, tp.y, tp.z, 0.0f);
}

// shadowLogicKernel, MaterialKernel.cu:361-388: colour (+)= the contributions of the lights whose
// shadow ray reached them, in light order; a 1/scale pass is read at loc / scale.
__global__ __launch_bounds__(kBlock1D) void k_shadow_logic(float4* __restrict__ img, uint32_t w,
                                                           uint32_t h, uint32_t sc, uint32_t addTo,
                                                           const uint8_t* __restrict__ inLight,
                                                           const float4* __restrict__ contrib,
                                                           uint32_t numLights)
{
	const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
	const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
	if (x >= w || y >= h) return;
	const uint32_t sw = w / sc, sh = h / sc;
	const uint32_t sx = min(x / sc, sw - 1u), sy = min(y / sc, sh - 1u);
	const uint32_t sid = pixel_id(sx, sy, sw, sh);
	const size_t sN = size_t(sw) * sh;
	float4* px = img + size_t(y) * w + x;
	float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
	if (addTo) {
		const float4 cur = *px;
		c0 = cur.x;
		c1 = cur.y;
		c2 = cur.z;
	}
	for (uint32_t i = 0; i < numLights; i++) {
		const size_t s = size_t(sid) + size_t(i) * sN;
		if (inLight[s]) {
			const float4 c = contrib[s];
			c0 = c0 + c.x;
			c1 = c1 + c.y;
			c2 = c2 + c.z;
		}
	}
	*px = make_float4(c0, c1, c2, 1.0f);
}

// Continuous pixel coordinates (px, py) -- the (x + jx, y + jy) of the primary rays -- at which
// camera `cam` sees the point cam.origin + v (v may be a direction: a point at infinity), the
// inverse of the pinhole formula normalize(dir + cx dX + cy dY), cx = px / w * 2 - 1, for
// orthogonal dir, dX, dY: cx = (v . dX) / ((v . dir) |dX|^2) (getUVCoord, velocity_gen.frag:40-47,
// for the tracer's camera). False for a point not in front of the camera.
__device__ __forceinline__ bool project(const PhxCameraDef& cam, V3 v, uint32_t w, uint32_t h,
                                        float& px, float& py)
{
	const V3 dir = v3(cam.dir[0], cam.dir[1], cam.dir[2]);
	const V3 dX = v3(cam.dX[0], cam.dX[1], cam.dX[2]), dY = v3(cam.dY[0], cam.dY[1], cam.dY[2]);
	const float a = dot(v, dir);
	if (!(a > 0.0f)) return false;
	const float cx = dot(v, dX) / (a * dot(dX, dX));
	const float cy = dot(v, dY) / (a * dot(dY, dY));
	px = ((cx + 1.0f) * 0.5f) * float(w);
	py = ((cy + 1.0f) * 0.5f) * float(h);
	return true;
}

// Bilinear tap of a w x h image at continuous pixel coordinates (hx, hy), texel centres at
// i + 0.5, edge texels repeated (GL_LINEAR, clamp to edge); rows blended first, each with
// lerp (1 - f) a + f b (MathHelpers.inl:66-70).
struct Taps {
	size_t i00, i01, i10, i11;
	float fx, fy;
};

__device__ __forceinline__ Taps taps(uint32_t w, uint32_t h, float hx, float hy)
{
	const float tx = hx - 0.5f, ty = hy - 0.5f;
	const float x0f = floorf(tx), y0f = floorf(ty);
	const int x0 = int(x0f), y0 = int(y0f);
	const int xa = min(max(x0, 0), int(w) - 1), xb = min(max(x0 + 1, 0), int(w) - 1);
	const int ya = min(max(y0, 0), int(h) - 1), yb = min(max(y0 + 1, 0), int(h) - 1);
	return Taps{size_t(ya) * w + xa, size_t(ya) * w + xb, size_t(yb) * w + xa,
	            size_t(yb) * w + xb, tx - x0f, ty - y0f};
}

__device__ __forceinline__ float lerp2(float a, float b, float d, float e, const Taps& t)
{
	const float top = a * (1.0f - t.fx) + b * t.fx;
	const float bot = d * (1.0f - t.fx) + e * t.fx;
	return top * (1.0f - t.fy) + bot * t.fy;
}

// TAA resolve (taa.frag:25-71) with the screen-space velocity of velocity_gen.frag:49-71.
// Velocity of a pixel, in pixels: its primary hit (record position; for a miss the primary ray's
// direction) projected by this frame's camera minus its projection by the previous frame's
// (FLT_MAX when the point is behind the previous camera). The history is read bilinearly at
// (x + 0.5, y + 0.5) - velocity; a position off the image starts the pixel afresh. Trust =
// clamp(1 - mix(3, 1, roughness) * min(1, 1.2 sqrt(| |v| - |v_hist| |)), 0, 1) with v_hist the
// previous frame's velocity read at the same position. FP32 history with the effective sample
// count in alpha; the history colour is clamped to the current frame's 3x3 neighbourhood
// min/max (clampHistory; not in the reference).
__global__ __launch_bounds__(kBlock1D) void k_taa(const float4* __restrict__ cur,
                                                  float4* hist0, float4* hist1, float2* vel0,
                                                  float2* vel1, const float4* __restrict__ gbuf,
                                                  const float4* __restrict__ prim, uint32_t w,
                                                  uint32_t h, uint32_t clampHistory,
                                                  const FrameParams* __restrict__ fp,
                                                  float4* __restrict__ out)
{
	const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
	const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
	if (x >= w || y >= h) return;
	const uint32_t reset = fp->taaReset;
	const float4* histIn = fp->histIn ? hist1 : hist0;
	float4* histOut = fp->histIn ? hist0 : hist1;
	const float2* velIn = fp->histIn ? vel1 : vel0;
	float2* velOut = fp->histIn ? vel0 : vel1;
	const size_t p = size_t(y) * w + x;
	const uint32_t rec = pixel_id(x, y, w, h);

	// velocity of this pixel
	const float4 g1 = gbuf[3 * size_t(rec)];
	const float4 g3 = gbuf[3 * size_t(rec) + 2];
	V3 vc, vp;
	if (g3.w >= 0.0f) {
		const V3 pos = xyz(g1);
		vc = sub(pos, v3(fp->cam.origin[0], fp->cam.origin[1], fp->cam.origin[2]));
		vp = sub(pos, v3(fp->prevCam.origin[0], fp->prevCam.origin[1], fp->prevCam.origin[2]));
	} else {
		vc = vp = xyz(prim[2 * size_t(rec) + 1]);
	}
	float cx = 0.0f, cy = 0.0f, px = 0.0f, py = 0.0f;
	float vx = FLT_MAX, vy = FLT_MAX;
	if (project(fp->cam, vc, w, h, cx, cy) && project(fp->prevCam, vp, w, h, px, py)) {
		vx = cx - px;
		vy = cy - py;
	}
	velOut[p] = make_float2(vx, vy);

	const float4 c = cur[p];
	const float hx = (float(x) + 0.5f) - vx, hy = (float(y) + 0.5f) - vy;
	const bool onScreen = hx >= 0.0f && hx <= float(w) && hy >= 0.0f && hy <= float(h);
	float o0, o1, o2, n2;
	if (reset || !onScreen) {
		o0 = c.x;
		o1 = c.y;
		o2 = c.z;
		n2 = 1.0f;
	} else {
		const Taps t = taps(w, h, hx, hy);
		const float2 v00 = velIn[t.i00], v01 = velIn[t.i01], v10 = velIn[t.i10],
		             v11 = velIn[t.i11];
		const float pvx = lerp2(v00.x, v01.x, v10.x, v11.x, t);
		const float pvy = lerp2(v00.y, v01.y, v10.y, v11.y, t);
		const float lv = sqrtf(vx * vx + vy * vy);
		const float lp = sqrtf(pvx * pvx + pvy * pvy);
		const float rough = g1.w;
		const float vDistrust = min_num(1.0f, 1.2f * sqrtf(fabsf(lv - lp)));
		const float rDistrust = 3.0f * (1.0f - rough) + 1.0f * rough;
		const float trust = clamp_num(1.0f - rDistrust * vDistrust, 0.0f, 1.0f);
		const float4 a = histIn[t.i00], b = histIn[t.i01], d = histIn[t.i10], e = histIn[t.i11];
		float h0 = lerp2(a.x, b.x, d.x, e.x, t);
		float h1 = lerp2(a.y, b.y, d.y, e.y, t);
		float h2 = lerp2(a.z, b.z, d.z, e.z, t);
		const float hn = lerp2(a.w, b.w, d.w, e.w, t);
		if (clampHistory) {
			float lo0 = c.x, lo1 = c.y, lo2 = c.z, hi0 = c.x, hi1 = c.y, hi2 = c.z;
			const uint32_t y0 = y ? y - 1u : 0u, y1 = y + 1u < h ? y + 1u : h - 1u;
			const uint32_t x0 = x ? x - 1u : 0u, x1 = x + 1u < w ? x + 1u : w - 1u;
			for (uint32_t yy = y0; yy <= y1; yy++)
				for (uint32_t xx = x0; xx <= x1; xx++) {
					const float4 q = cur[size_t(yy) * w + xx];
					lo0 = min_num(lo0, q.x);
					lo1 = min_num(lo1, q.y);
					lo2 = min_num(lo2, q.z);
					hi0 = max_num(hi0, q.x);
					hi1 = max_num(hi1, q.y);
					hi2 = max_num(hi2, q.z);
				}
			h0 = clamp_num(h0, lo0, hi0);
			h1 = clamp_num(h1, lo1, hi1);
			h2 = clamp_num(h2, lo2, hi2);
		}
		n2 = hn * trust + 1.0f;
		const float blend = 1.0f - 1.0f / n2;
		o0 = c.x * (1.0f - blend) + h0 * blend;
		o1 = c.y * (1.0f - blend) + h1 * blend;
		o2 = c.z * (1.0f - blend) + h2 * blend;
	}
	histOut[p] = make_float4(o0, o1, o2, n2);
	out[p] = make_float4(o0, o1, o2, 1.0f);
}

uint32_t blocks1d(uint32_t n) { return n / kBlock1D + (n % kBlock1D != 0u); }
dim3 grid2d(uint32_t w, uint32_t h) { return dim3((w + 15u) / 16u, (h + 15u) / 16u); }

}  // namespace

hipError_t launchInterpret(const PhxRayIn* rays, const PhxRayHit* hits, uint32_t n,
                           const PhxTriangleData* datas, uint32_t numTris, const PhxMaterial* mats,
                           uint32_t numMats, const TexSet& tex, float4* info, hipStream_t stream,
                           const uint32_t* instIds, const DevInstance* inst, uint32_t numInst)
{
	if (n == 0) return hipSuccess;
	hipLaunchKernelGGL(k_interpret, dim3(blocks1d(n)), dim3(kBlock1D), 0, stream,
	                   reinterpret_cast<const float4*>(rays), reinterpret_cast<const uint4*>(hits),
	                   n, datas, numTris, instIds, inst, numInst, mats, numMats, tex, info);
	return hipGetLastError();
}

hipError_t launchShade(const float4* info, const PhxRayIn* rays, const float4* throughput,
                       const PhxSphereLight* lights, uint32_t numLights, uint32_t count,
                       const FrameParams* fp, uint32_t purpose, float4* contrib, PhxRayIn* shadow,
                       hipStream_t stream)
{
	if (count == 0 || numLights == 0) return hipSuccess;
	hipLaunchKernelGGL(k_shade, dim3(blocks1d(count)), dim3(kBlock1D), 0, stream, info,
	                   reinterpret_cast<const float4*>(rays), throughput, lights, numLights, count,
	                   fp, purpose, contrib, reinterpret_cast<float4*>(shadow));
	return hipGetLastError();
}

hipError_t launchGenSecondary(const float4* gbuf, uint32_t w, uint32_t h, const FrameParams* fp,
                              PhxRayIn* rays, float4* throughput, hipStream_t stream)
{
	if (w < 2u || h < 2u) return hipSuccess;
	hipLaunchKernelGGL(k_gen_secondary, grid2d(w / 2u, h / 2u), dim3(kBlock1D), 0, stream, gbuf, w,
	                   h, fp, reinterpret_cast<float4*>(rays), throughput);
	return hipGetLastError();
}

hipError_t launchShadowLogic(float4* img, uint32_t w, uint32_t h, uint32_t scale, bool addTo,
                             const uint8_t* inLight, const float4* contrib, uint32_t numLights,
                             hipStream_t stream)
{
	if (w == 0 || h == 0) return hipSuccess;
	hipLaunchKernelGGL(k_shadow_logic, grid2d(w, h), dim3(kBlock1D), 0, stream, img, w, h, scale,
	                   uint32_t(addTo), inLight, contrib, numLights);
	return hipGetLastError();
}

hipError_t launchTaa(const float4* cur, float4* hist0, float4* hist1, float2* vel0, float2* vel1,
                     const float4* gbuf, const PhxRayIn* prim, uint32_t w, uint32_t h,
                     bool clampHistory, const FrameParams* fp, float4* out, hipStream_t stream)
{
	if (w == 0 || h == 0) return hipSuccess;
	hipLaunchKernelGGL(k_taa, grid2d(w, h), dim3(kBlock1D), 0, stream, cur, hist0, hist1, vel0,
	                   vel1, gbuf, reinterpret_cast<const float4*>(prim), w, h,
	                   uint32_t(clampHistory), fp, out);
	return hipGetLastError();
}

namespace {
__global__ void k_set_frame_params(FrameParams* dst, FrameParams value)
{
	if (threadIdx.x == 0) *dst = value;  // per-lane condition: an ordinary vector store
}
}  // namespace

hipError_t launchSetFrameParams(FrameParams* dst, const FrameParams& value, hipStream_t stream)
{
	hipLaunchKernelGGL(k_set_frame_params, dim3(1), dim3(1), 0, stream, dst, value);
	return hipGetLastError();
}

}  // namespace phx
