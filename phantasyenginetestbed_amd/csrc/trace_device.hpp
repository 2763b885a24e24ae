// Device-side arithmetic of the SBVH traversal, shared by the HIP kernels.
//
// Every function reproduces the reference CPU traversal's IEEE float operation sequence
// (ray_tracer_common/Intersection.hpp:24-90, Ray.hpp:35-37, sfzCore Vector.inl:345-367 and
// :425-470, :569-601). The translation units that include this header are compiled with
// -ffp-contract=off and correctly rounded division, so a*b+c is never fused and 1/x is IEEE
// division: hits, t, u and v are the CPU's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace phx {
namespace dev {

// std::min / std::max: return the FIRST operand when the comparison is false (NaN, +-0 ties).
// Deliberately not fminf/fmaxf (which drop NaNs) and not v_min3/v_max3.
__device__ __forceinline__ float min_first(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_first(float a, float b) { return (a < b) ? b : a; }

struct Ray {
	float ox, oy, oz;
	float dx, dy, dz;
	float ix, iy, iz;  // 1 / dir, elementwise IEEE division (Ray::setDir, Ray.hpp:35-37)
	float tMin, tMax;
	bool nanFree;      // no slab value of this ray can be NaN (see slab_fast)
};

__device__ __forceinline__ Ray make_ray(const float4 a, const float4 b)
{
	Ray r;
	r.ox = a.x, r.oy = a.y, r.oz = a.z, r.tMin = a.w;
	r.dx = b.x, r.dy = b.y, r.dz = b.z, r.tMax = b.w;
	r.ix = 1.0f / b.x;
	r.iy = 1.0f / b.y;
	r.iz = 1.0f / b.z;
	// |o| <= 2^126 and node bounds within the same range (finiteBounds) keep b - o finite, so
	// (b - o) * (1/d) is never NaN, also where 1/d is 0 (an infinite direction).
	r.nanFree = __builtin_isfinite(r.ix) && __builtin_isfinite(r.iy) && __builtin_isfinite(r.iz) &&
	            __builtin_fabsf(r.ox) <= 0x1p126f && __builtin_fabsf(r.oy) <= 0x1p126f &&
	            __builtin_fabsf(r.oz) <= 0x1p126f;
	return r;
}

// Slab test of one AABB (Intersection.hpp:77-90): t1 = (min - o) * invDir, t2 = (max - o) *
// invDir, tIn = maxElement(min(t1, t2)), tOut = minElement(max(t1, t2)); hit iff tOut >= tIn.
__device__ __forceinline__ void slab(const Ray& r, float mnx, float mny, float mnz, float mxx,
                                     float mxy, float mxz, float& tIn, float& tOut)
{
	float t1x = (mnx - r.ox) * r.ix;
	float t1y = (mny - r.oy) * r.iy;
	float t1z = (mnz - r.oz) * r.iz;
	float t2x = (mxx - r.ox) * r.ix;
	float t2y = (mxy - r.oy) * r.iy;
	float t2z = (mxz - r.oz) * r.iz;
	float lx = min_first(t1x, t2x), ly = min_first(t1y, t2y), lz = min_first(t1z, t2z);
	float hx = max_first(t1x, t2x), hy = max_first(t1y, t2y), hz = max_first(t1z, t2z);
	tIn = max_first(max_first(lx, ly), lz);
	tOut = min_first(min_first(hx, hy), hz);
}

// v_min_f32 / v_max_f32 / v_min3_f32 / v_max3_f32 on values known not to be NaN. Written out
// because the minnum/maxnum intrinsics first canonicalise their inputs (one extra v_max per
// operand) when the compiler cannot prove them canonical. ALU only: no memory access in asm.
__device__ __forceinline__ float vmin(float a, float b)
{
	float r;
	asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
	return r;
}
__device__ __forceinline__ float vmax(float a, float b)
{
	float r;
	asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
	return r;
}
__device__ __forceinline__ float vmin3(float a, float b, float c)
{
	float r;
	asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
	return r;
}
__device__ __forceinline__ float vmax3(float a, float b, float c)
{
	float r;
	asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
	return r;
}

// The same slab test for a ray whose 1/dir is finite and whose origin lies within +-2^126. Box
// bounds lie within +-2^126 too (checked at upload), so (b - o) is finite and (b - o) * (1/dir) is
// finite, +-inf or 0 (1/dir may be 0 for an infinite direction): no slab value is NaN.
// Without NaNs, std::min/std::max and the hardware min/max can differ only in the sign of a zero
// result, and tIn/tOut are only ever compared (-0 == +0), so every visit decision equals
// slab()'s.
__device__ __forceinline__ void slab_fast(const Ray& r, float mnx, float mny, float mnz, float mxx,
                                          float mxy, float mxz, float& tIn, float& tOut)
{
	float t1x = (mnx - r.ox) * r.ix;
	float t1y = (mny - r.oy) * r.iy;
	float t1z = (mnz - r.oz) * r.iz;
	float t2x = (mxx - r.ox) * r.ix;
	float t2y = (mxy - r.oy) * r.iy;
	float t2z = (mxz - r.oz) * r.iz;
	tIn = vmax3(vmin(t1x, t2x), vmin(t1y, t2y), vmin(t1z, t2z));
	tOut = vmin3(vmax(t1x, t2x), vmax(t1y, t2y), vmax(t1z, t2z));
}

// Moller-Trumbore, Intersection.hpp:24-67 (RTR3 p. 750). Returns false on a miss. dot() is
// (x*x' + y*y') + z*z' and cross() is (y*z'-z*y', z*x'-x*z', x*y'-y*x').
__device__ __forceinline__ bool triangle(const Ray& r, const float4 p0, const float4 p1,
                                         const float4 p2, float& t, float& u, float& v)
{
	const float EPS = 0.00001f;
	float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
	float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
	float qx = r.dy * e2z - r.dz * e2y;
	float qy = r.dz * e2x - r.dx * e2z;
	float qz = r.dx * e2y - r.dy * e2x;
	float a = e1x * qx + e1y * qy;
	a = a + e1z * qz;
	if (-EPS < a && a < EPS) return false;
	float f = 1.0f / a;
	float sx = r.ox - p0.x, sy = r.oy - p0.y, sz = r.oz - p0.z;
	float sq = sx * qx + sy * qy;
	sq = sq + sz * qz;
	u = f * sq;
	if (u < 0.0f) return false;
	float rx = sy * e1z - sz * e1y;
	float ry = sz * e1x - sx * e1z;
	float rz = sx * e1y - sy * e1x;
	float dr = r.dx * rx + r.dy * ry;
	dr = dr + r.dz * rz;
	v = f * dr;
	if (v < 0.0f || (u + v) > 1.0f) return false;
	float er = e2x * rx + e2y * ry;
	er = er + e2z * rz;
	t = f * er;
	return true;
}

}  // namespace dev
}  // namespace phx

namespace phx {
namespace dev {

// Moller-Trumbore on a pre-transformed triangle (p0, e1 = p1 - p0, e2 = p2 - p0). e1 and e2 are
// computed once at upload with the same IEEE subtraction the CPU performs per test, so every
// subsequent operation -- and the result -- is bit-identical to triangle().
__device__ __forceinline__ bool triangle_pre(const Ray& r, const float4 p0, const float4 e1,
                                             const float4 e2, float& t, float& u, float& v)
{
	const float EPS = 0.00001f;
	float qx = r.dy * e2.z - r.dz * e2.y;
	float qy = r.dz * e2.x - r.dx * e2.z;
	float qz = r.dx * e2.y - r.dy * e2.x;
	float a = e1.x * qx + e1.y * qy;
	a = a + e1.z * qz;
	if (-EPS < a && a < EPS) return false;
	float f = 1.0f / a;
	float sx = r.ox - p0.x, sy = r.oy - p0.y, sz = r.oz - p0.z;
	float sq = sx * qx + sy * qy;
	sq = sq + sz * qz;
	u = f * sq;
	if (u < 0.0f) return false;
	float rx = sy * e1.z - sz * e1.y;
	float ry = sz * e1.x - sx * e1.z;
	float rz = sx * e1.y - sy * e1.x;
	float dr = r.dx * rx + r.dy * ry;
	dr = dr + r.dz * rz;
	v = f * dr;
	if (v < 0.0f || (u + v) > 1.0f) return false;
	float er = e2.x * rx + e2.y * ry;
	er = er + e2.z * rz;
	t = f * er;
	return true;
}

}  // namespace dev
}  // namespace phx

namespace phx {
namespace dev {

// 1/a rounded to nearest even, as IEEE division gives it, at a fraction of the cost of the
// division sequence: v_rcp_f32 (error below one ulp) refined by one Newton-Raphson step in fused
// multiply-adds (Markstein's reciprocal iteration: e = 1 - a*y is exact, y' = y + e*y). On gfx950
// the result equals IEEE 1/a for every a with 2^-126 <= |a| < 2^126 -- established by comparing
// the two over all 2^32 inputs on the device (phx_check_reciprocal repeats that comparison). Outside
// that range (denormal results, infinities, NaN) the caller divides.
__device__ __forceinline__ float recip_newton(float a)
{
	const float y = __builtin_amdgcn_rcpf(a);
	const float e = __builtin_fmaf(-a, y, 1.0f);
	return __builtin_fmaf(e, y, y);
}
__device__ __forceinline__ bool recip_newton_ok(float a) { return __builtin_fabsf(a) < 0x1p126f; }

// Moller-Trumbore split in two halves so that the determinants of several triangles can share
// one range check of the fast reciprocal. Same operations, in the same order, as triangle_pre.
struct TriDet {
	float qx, qy, qz, a;
};

__device__ __forceinline__ TriDet tri_det(const Ray& r, const float4 e1, const float4 e2)
{
	TriDet d;
	d.qx = r.dy * e2.z - r.dz * e2.y;
	d.qy = r.dz * e2.x - r.dx * e2.z;
	d.qz = r.dx * e2.y - r.dy * e2.x;
	float a = e1.x * d.qx + e1.y * d.qy;
	d.a = a + e1.z * d.qz;
	return d;
}

// Second half, branch-free: true iff triangle_pre returns true; t, u, v are then its outputs.
// `f` is IEEE 1/d.a (only its value for accepted determinants matters).
__device__ __forceinline__ bool tri_finish(const Ray& r, const TriDet& d, float f, const float4 p0,
                                           const float4 e1, const float4 e2, float& t, float& u,
                                           float& v)
{
	const float EPS = 0.00001f;
	const bool det = !(-EPS < d.a && d.a < EPS);
	float sx = r.ox - p0.x, sy = r.oy - p0.y, sz = r.oz - p0.z;
	float sq = sx * d.qx + sy * d.qy;
	sq = sq + sz * d.qz;
	u = f * sq;
	float rx = sy * e1.z - sz * e1.y;
	float ry = sz * e1.x - sx * e1.z;
	float rz = sx * e1.y - sy * e1.x;
	float dr = r.dx * rx + r.dy * ry;
	dr = dr + r.dz * rz;
	v = f * dr;
	float er = e2.x * rx + e2.y * ry;
	er = er + e2.z * rz;
	t = f * er;
	return det && !(u < 0.0f) && !(v < 0.0f || (u + v) > 1.0f);
}

}  // namespace dev
}  // namespace phx
