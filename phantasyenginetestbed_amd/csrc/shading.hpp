// Shading arithmetic of the wavefront frame, on the device.
//
// Every function uses only IEEE-754 operations that are correctly rounded on both gfx950 and the
// host (+, -, *, /, sqrt, compares), in a fixed order, compiled without contraction. Library calls
// whose results differ between implementations are restated from their definitions:
//   powf(x, 2) -> x*x, powf(x, 4) -> (x*x)*(x*x), powf(x, 5) -> ((x*x)*(x*x))*x
//   cos/sin(2*pi*r) -> turn_sincos(r) (quadrant reduction + Taylor polynomials, below)
//   fmaxf/fminf/clamp -> max_num/min_num/clamp_num (NaN yields the other operand, as fmaxf)
// so a CPU restatement with the same order reproduces every bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace phx {
namespace shading {

struct V3 {
	float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }

// sfz dot: x*x' + y*y' + z*z' left to right (Vector.inl:339-344).
__device__ __forceinline__ float dot(V3 a, V3 b)
{
	float d = a.x * b.x + a.y * b.y;
	return d + a.z * b.z;
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, length(a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
	return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ float max_num(float a, float b)
{
	return (a != a) ? b : ((b != b) ? a : (a > b ? a : b));
}
__device__ __forceinline__ float min_num(float a, float b)
{
	return (a != a) ? b : ((b != b) ? a : (a < b ? a : b));
}
// clamp of CudaDeviceHelpers.cuh:38-41: fminf(fmaxf(v, lo), hi).
__device__ __forceinline__ float clamp_num(float v, float lo, float hi)
{
	return min_num(max_num(v, lo), hi);
}
__device__ __forceinline__ bool finite(float x) { return x - x == 0.0f; }

// cos(2*pi*r) and sin(2*pi*r). 4r is split into the nearest quadrant q and a remainder f in
// [-1/2, 1/2] (exact: r has at most 24 significant bits), x = f*pi/2 in [-pi/4, pi/4], and
// sin x, cos x come from their Taylor polynomials to x^9 / x^8 (truncation below 2e-9 there).
__device__ __forceinline__ void turn_sincos(float r, float& s, float& c)
{
	const float r4 = r * 4.0f;
	const float q = rintf(r4);
	const float f = r4 - q;
	const float x = f * 1.57079632679489662f;
	const float x2 = x * x;
	float ps = x2 * 2.75573192e-6f;
	ps = ps + -1.98412698e-4f;
	ps = ps * x2;
	ps = ps + 8.33333333e-3f;
	ps = ps * x2;
	ps = ps + -1.66666667e-1f;
	ps = ps * x2;
	ps = ps * x;
	const float sx = x + ps;
	float pc = x2 * 2.48015873e-5f;
	pc = pc + -1.38888889e-3f;
	pc = pc * x2;
	pc = pc + 4.16666667e-2f;
	pc = pc * x2;
	pc = pc + -0.5f;
	pc = pc * x2;
	const float cx = 1.0f + pc;
	const int quadrant = int(q) & 3;
	if (quadrant == 0) {
		c = cx;
		s = sx;
	} else if (quadrant == 1) {
		c = -sx;
		s = cx;
	} else if (quadrant == 2) {
		c = -cx;
		s = -sx;
	} else {
		c = sx;
		s = -cx;
	}
}

constexpr double kPiD = 3.14159265358979323846;  // CUDART_PI (a double)
constexpr float kPiF = 3.14159265358979323846f;  // float(CUDART_PI)

// ggx, CudaPbr.cuh:26-31; CUDART_PI * powf(.., 2) is evaluated in double, as there.
__device__ __forceinline__ float ggx(float nDotH, float a)
{
	const float a2 = a * a;
	float x = nDotH * nDotH;
	x = x * (a2 - 1.0f);
	x = x + 1.0f;
	const float div = float(kPiD * double(x * x));
	return a2 / div;
}

// geometricSchlick, CudaPbr.cuh:40-45.
__device__ __forceinline__ float geometric_schlick(float nDotL, float nDotV, float k)
{
	const float g1 = nDotL / (nDotL * (1.0f - k) + k);
	const float g2 = nDotV / (nDotV * (1.0f - k) + k);
	return g1 * g2;
}

// shade, CudaPbr.cuh:58-103: Lambert + Cook-Torrance (GGX, Smith-Schlick, Schlick Fresnel).
__device__ __forceinline__ V3 shade(V3 n, V3 v, V3 albedo, float roughness, float metallic,
                                    V3 l, float toLightDist, V3 strength, float range)
{
	const V3 zero = v3(0.0f, 0.0f, 0.0f);
	if (toLightDist > range) return zero;
	const V3 h = normalize(add(l, v));
	const float nDotL = dot(n, l);
	if (nDotL <= 0.0f) return zero;
	const float nDotV = max_num(0.001f, dot(n, v));
	const V3 diffuse = divs(albedo, kPiF);
	const float nDotH = max_num(0.0f, dot(n, h));
	const float ctD = ggx(nDotH, roughness * roughness);
	const float r1 = roughness + 1.0f;
	const float k = (r1 * r1) / 8.0f;
	const float ctG = geometric_schlick(nDotL, nDotV, k);
	// f0 = lerp(0.04, albedo, metallic) = (1 - m) * 0.04 + m * albedo (MathHelpers.inl:66-70)
	const float om = 1.0f - metallic;
	const V3 f0 = v3(om * 0.04f + metallic * albedo.x, om * 0.04f + metallic * albedo.y,
	                 om * 0.04f + metallic * albedo.z);
	const float p = 1.0f - nDotV;
	const float p2 = p * p;
	const float p5 = clamp_num((p2 * p2) * p, 0.0f, 1.0f);
	const V3 ctF = v3(f0.x + (1.0f - f0.x) * p5, f0.y + (1.0f - f0.y) * p5,
	                  f0.z + (1.0f - f0.z) * p5);
	const float den = (4.0f * nDotL) * nDotV;
	const V3 spec = divs(scale(scale(ctF, ctD), ctG), den);
	const V3 kd = scale(v3(1.0f - ctF.x, 1.0f - ctF.y, 1.0f - ctF.z), om);
	V3 out;
	out.x = ((kd.x * diffuse.x + spec.x) * strength.x) * nDotL;
	out.y = ((kd.y * diffuse.y + spec.y) * strength.y) * nDotL;
	out.z = ((kd.z * diffuse.z + spec.z) * strength.z) * nDotL;
	if (!finite(out.x) || !finite(out.y) || !finite(out.z)) return zero;
	return out;
}

// falloffFactor, CudaPbr.cuh:105-110 (the denominator's + 1.0 is a double addition there).
__device__ __forceinline__ float falloff(float d, float range)
{
	const float q = d / range;
	const float q2 = q * q;
	const float c = clamp_num(1.0f - q2 * q2, 0.0f, 1.0f);
	const float num = c * c;
	const float den = float(double(d * d) + 1.0);
	return num / den;
}

// hemisphereToWorld, MaterialKernel.cu:206-218, with (cos phi, sin phi) given.
__device__ __forceinline__ V3 hemisphere_to_world(float cosPhi, float sinPhi, float cosT,
                                                  float sinT, V3 n)
{
	const float hx = sinT * cosPhi, hy = sinT * sinPhi, hz = cosT;
	const bool zUp = fabsf(n.z) < 0.999f;
	const V3 up = zUp ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
	const V3 u = normalize(cross(up, n));
	const V3 v = cross(n, u);
	return add(add(scale(u, hx), scale(v, hy)), scale(n, hz));
}

// sampleLambertian, MaterialKernel.cu:239-246.
__device__ __forceinline__ V3 sample_lambert(float r1, float r2, V3 n)
{
	float s, c;
	turn_sincos(r1, s, c);
	return hemisphere_to_world(c, s, sqrtf(1.0f - r2), sqrtf(r2), n);
}

// sampleGgx + clampToHemisphere, MaterialKernel.cu:220-236.
__device__ __forceinline__ V3 sample_ggx(float r1, float r2, float a, V3 refl, V3 n)
{
	float s, c;
	turn_sincos(r1, s, c);
	const float cosT = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
	const float sinT = sqrtf(1.0f - cosT * cosT);
	const V3 w = hemisphere_to_world(c, s, cosT, sinT, refl);
	const float nd = dot(w, n);
	const float off = -((nd < 0.0f) ? nd : 0.0f);
	return normalize(add(w, scale(n, off)));
}

// reflect, CudaDeviceHelpers.cuh:44-47.
__device__ __forceinline__ V3 reflect(V3 in, V3 n)
{
	const float d2 = 2.0f * dot(n, in);
	return sub(in, scale(n, d2));
}

}  // namespace shading
}  // namespace phx
