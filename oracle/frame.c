/*
 * frame.c -- CPU restatement of the wavefront frame and the TAA resolve. TEST INFRASTRUCTURE ONLY
 * (see oracle.h): loaded by tests/ to check the device frame, never by the product.
 *
 * Built like oracle.c (-O2 -ffp-contract=off, no fast-math, SSE float): every expression is one
 * IEEE single-precision operation in the written order. The shading formulas avoid library
 * functions whose results vary between implementations: powf with integer exponents is written
 * as products, cos/sin(2*pi*r) as quadrant reduction plus Taylor polynomials (turn_sincos), and
 * min/max/clamp as explicit compares with fmaxf's NaN rule.
 *
 * Parity pinning: the closed forms of tests/test_frame_oracle.py -- shading of a light along
 * the normal against the textbook Cook-Torrance value computed in double, turn_sincos against
 * math.cos/sin, the cosine-weighted sample's mean and first moments, the TAA running mean of a
 * static image, the shadow logic's sum over lights -- and a second evaluation of ggx / Fresnel /
 * falloff in numpy. The bounce stage (secondary rays with the 1-of-4 pixel pick, the Lambert/GGX
 * roulette and the glass ray, throughput, shading of the bounce hit, the half-resolution add) is
 * pinned by tests/test_frame_bounce_oracle.py: plane scenes whose colour has a closed form per
 * pixel, with the random draws recomputed from PCG32's published definition (itself checked
 * against the published pcg32 demo output); that file also plants one-constant mistakes in a copy
 * of this file and shows each is caught. The reference's own random stream (curand XORWOW) is not
 * reproduced: the frame draws from PCG32, so sampled directions are "parity unpinned" against the
 * reference's stream and pinned to its formulas only.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct V3 {
	float x, y, z;
} V3;

static inline V3 v3(float x, float y, float z)
{
	V3 r = {x, y, z};
	return r;
}
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
/* sfz dot, left to right (Vector.inl:339-344) */
static inline float dot(V3 a, V3 b)
{
	float d = a.x * b.x + a.y * b.y;
	return d + a.z * b.z;
}
static inline float len(V3 a) { return sqrtf(dot(a, a)); }
static inline V3 nrm(V3 a) { return divs(a, len(a)); }
static inline V3 cross(V3 a, V3 b)
{
	return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float max_num(float a, float b)
{
	if (a != a) return b;
	if (b != b) return a;
	return a > b ? a : b;
}
static inline float min_num(float a, float b)
{
	if (a != a) return b;
	if (b != b) return a;
	return a < b ? a : b;
}
static inline float clamp_num(float v, float lo, float hi) { return min_num(max_num(v, lo), hi); }
static inline int finite_f(float x) { return x - x == 0.0f; }

/* PCG32 (O'Neill 2014): pcg32_random_r / pcg32_srandom_r. */
typedef struct Pcg {
	uint64_t state, inc;
} Pcg;

static inline uint32_t pcg_next(Pcg* g)
{
	uint64_t old = g->state;
	g->state = old * 6364136223846793005ull + g->inc;
	uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
	uint32_t rot = (uint32_t)(old >> 59u);
	return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static inline float pcg_uniform(Pcg* g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }

static inline Pcg pcg_seeded(uint64_t init, uint64_t seq)
{
	Pcg g;
	g.state = 0u;
	g.inc = (seq << 1u) | 1u;
	pcg_next(&g);
	g.state += init;
	pcg_next(&g);
	return g;
}

/* Stream of ray/pixel `id` for one use (`purpose`) in one frame. */
static inline Pcg frame_rng(uint64_t seed, uint32_t frame, uint32_t purpose, uint32_t id)
{
	return pcg_seeded(seed ^ ((uint64_t)frame << 32), ((uint64_t)purpose << 32) | id);
}

void oracle_turn_sincos(float r, float* s, float* c)
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
	switch ((int)q & 3) {
	case 0: *c = cx; *s = sx; break;
	case 1: *c = -sx; *s = cx; break;
	case 2: *c = -cx; *s = -sx; break;
	default: *c = sx; *s = -cx; break;
	}
}

static const double PI_D = 3.14159265358979323846;
static const float PI_F = 3.14159265358979323846f;

/* ggx, CudaPbr.cuh:26-31 (CUDART_PI * powf(x, 2) in double). */
float oracle_ggx(float nDotH, float a)
{
	float a2 = a * a;
	float x = nDotH * nDotH;
	x = x * (a2 - 1.0f);
	x = x + 1.0f;
	float div = (float)(PI_D * (double)(x * x));
	return a2 / div;
}

/* falloffFactor, CudaPbr.cuh:105-110. */
float oracle_falloff(float d, float range)
{
	float q = d / range;
	float q2 = q * q;
	float c = clamp_num(1.0f - q2 * q2, 0.0f, 1.0f);
	float num = c * c;
	float den = (float)((double)(d * d) + 1.0);
	return num / den;
}

/* shade, CudaPbr.cuh:58-103. */
static V3 shade(V3 n, V3 v, V3 albedo, float rough, float metal, V3 l, float dist, V3 strength,
                float range)
{
	const V3 zero = v3(0.0f, 0.0f, 0.0f);
	if (dist > range) return zero;
	V3 h = nrm(add(l, v));
	float nDotL = dot(n, l);
	if (nDotL <= 0.0f) return zero;
	float nDotV = max_num(0.001f, dot(n, v));
	V3 diffuse = divs(albedo, PI_F);
	float nDotH = max_num(0.0f, dot(n, h));
	float ctD = oracle_ggx(nDotH, rough * rough);
	float r1 = rough + 1.0f;
	float k = (r1 * r1) / 8.0f;
	float g1 = nDotL / (nDotL * (1.0f - k) + k);
	float g2 = nDotV / (nDotV * (1.0f - k) + k);
	float ctG = g1 * g2;
	float om = 1.0f - metal;
	V3 f0 = v3(om * 0.04f + metal * albedo.x, om * 0.04f + metal * albedo.y,
	           om * 0.04f + metal * albedo.z);
	float p = 1.0f - nDotV;
	float p2 = p * p;
	float p5 = clamp_num((p2 * p2) * p, 0.0f, 1.0f);
	V3 ctF = v3(f0.x + (1.0f - f0.x) * p5, f0.y + (1.0f - f0.y) * p5, f0.z + (1.0f - f0.z) * p5);
	float den = (4.0f * nDotL) * nDotV;
	V3 spec = divs(scale(scale(ctF, ctD), ctG), den);
	V3 kd = scale(v3(1.0f - ctF.x, 1.0f - ctF.y, 1.0f - ctF.z), om);
	V3 o;
	o.x = ((kd.x * diffuse.x + spec.x) * strength.x) * nDotL;
	o.y = ((kd.y * diffuse.y + spec.y) * strength.y) * nDotL;
	o.z = ((kd.z * diffuse.z + spec.z) * strength.z) * nDotL;
	if (!finite_f(o.x) || !finite_f(o.y) || !finite_f(o.z)) return zero;
	return o;
}

void oracle_shade(const float n[3], const float v[3], const float albedo[3], float rough,
                  float metal, const float l[3], float dist, const float strength[3], float range,
                  float out[3])
{
	V3 r = shade(v3(n[0], n[1], n[2]), v3(v[0], v[1], v[2]), v3(albedo[0], albedo[1], albedo[2]),
	             rough, metal, v3(l[0], l[1], l[2]), dist,
	             v3(strength[0], strength[1], strength[2]), range);
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
}

/* hemisphereToWorld, MaterialKernel.cu:206-218. */
static V3 hemi_to_world(float cphi, float sphi, float cosT, float sinT, V3 n)
{
	float hx = sinT * cphi, hy = sinT * sphi, hz = cosT;
	V3 up = fabsf(n.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
	V3 u = nrm(cross(up, n));
	V3 w = cross(n, u);
	return add(add(scale(u, hx), scale(w, hy)), scale(n, hz));
}

/* sampleLambertian, MaterialKernel.cu:239-246. */
static V3 sample_lambert(float r1, float r2, V3 n)
{
	float s, c;
	oracle_turn_sincos(r1, &s, &c);
	return hemi_to_world(c, s, sqrtf(1.0f - r2), sqrtf(r2), n);
}

/* sampleGgx + clampToHemisphere, MaterialKernel.cu:220-236. */
static V3 sample_ggx(float r1, float r2, float a, V3 refl, V3 n)
{
	float s, c;
	oracle_turn_sincos(r1, &s, &c);
	float cosT = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
	float sinT = sqrtf(1.0f - cosT * cosT);
	V3 w = hemi_to_world(c, s, cosT, sinT, refl);
	float nd = dot(w, n);
	float off = -((nd < 0.0f) ? nd : 0.0f);
	return nrm(add(w, scale(n, off)));
}

void oracle_sample_lambert(float r1, float r2, const float n[3], float out[3])
{
	V3 d = sample_lambert(r1, r2, v3(n[0], n[1], n[2]));
	out[0] = d.x;
	out[1] = d.y;
	out[2] = d.z;
}

void oracle_sample_ggx(float r1, float r2, float a, const float refl[3], const float n[3],
                       float out[3])
{
	V3 d = sample_ggx(r1, r2, a, v3(refl[0], refl[1], refl[2]), v3(n[0], n[1], n[2]));
	out[0] = d.x;
	out[1] = d.y;
	out[2] = d.z;
}

void oracle_pcg32(uint64_t init, uint64_t seq, uint32_t n, uint32_t* out)
{
	Pcg g = pcg_seeded(init, seq);
	for (uint32_t i = 0; i < n; i++) out[i] = pcg_next(&g);
}

/* Pixel -> ray/record index: 8x8 tiles (getRayIdx, CudaDeviceHelpers.cuh:49-62) when both sizes
 * are multiples of 8, row-major otherwise. */
static inline uint32_t pixel_id(uint32_t x, uint32_t y, uint32_t w, uint32_t h)
{
	if ((w % 8u) == 0u && (h % 8u) == 0u) {
		uint32_t tile = (y / 8u) * (w / 8u) + (x / 8u);
		return tile * 64u + (y % 8u) * 8u + (x % 8u);
	}
	return y * w + x;
}

static inline void set_dummy(OrRay* r)
{
	r->o[0] = r->o[1] = r->o[2] = 1000000.0f;
	r->tMin = 0.001f;
	r->d[0] = 0.0f;
	r->d[1] = 1.0f;
	r->d[2] = 0.0f;
	r->tMax = 0.001f;
}

static inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

/* CPURayTracerRenderer::sampleImage (CPUTracerRenderer.cpp:160-180) along one axis: the texel
 * index floor(fmod(fmod(c * dim, dim) + dim, dim)); NaN and infinite c (for which the reference
 * converts NaN to an integer) read texel 0. */
uint32_t oracle_texel_coord(float c, uint32_t dim)
{
	const float d = (float)dim;
	float s = fmodf(c * d, d);
	s = s + d;
	s = fmodf(s, d);
	const float f = floorf(s);
	if (isnan(f)) return 0u;
	return (uint32_t)f;
}

/* linearize (InterpretRayHitKernel.cu:39-45) of the 256 byte values c / 255.0f: powf(x, 2.2f),
 * here evaluated in long double and rounded to float. */
void oracle_linear_table(float out[256])
{
	for (int c = 0; c < 256; c++) {
		const float x = (float)c / 255.0f;
		out[c] = (float)powl((long double)x, (long double)2.2f);
	}
}

static const uint8_t* texel_of(const OrTexture* t, float u, float v)
{
	const uint32_t x = oracle_texel_coord(u, t->width), y = oracle_texel_coord(v, t->height);
	return t->data + (size_t)y * t->pitch + (size_t)x * t->bpp;
}

/* interpretRayHitKernel, InterpretRayHitKernel.cu:13-121; a material index outside the table
 * reads the default material (albedo 1, roughness 1, metallic 0), a texture index outside
 * [0, numTex) means no texture. Albedo textures are read only with 4 bytes per texel
 * (CPUTracerRenderer.cpp:200-206; the CUDA path reads uchar4). */
/* Hit records; with instances (instIds not NULL), a hit of instance k reads that instance's mesh
 * data and its interpolated normal is mapped to world space by the transpose of the instance's
 * world -> object matrix (A, rows of 4) before normalisation. */
static void interpret_records(const OrRay* rays, const OrHit* hits, const uint32_t* instIds,
                              uint32_t n, const OrTriData* datas, const OrInstance* inst,
                              const float* instA, uint32_t numInst, const OrMaterial* mats,
                              uint32_t numMats, const OrTexture* texs, uint32_t numTex,
                              OrHitInfo* out)
{
	float lin[256];
	oracle_linear_table(lin);
	for (uint32_t i = 0; i < n; i++) {
		OrHitInfo hi;
		memset(&hi, 0, sizeof(hi));
		const uint32_t ik = instIds ? instIds[i] : 0xFFFFFFFFu;
		const OrTriData* dtab = ik < numInst ? inst[ik].datas : datas;
		if (hits[i].tri != 0xFFFFFFFFu) {
			const OrRay* r = &rays[i];
			const OrTriData* d = &dtab[hits[i].tri];
			float t = hits[i].t, bu = hits[i].u, bv = hits[i].v;
			for (int k = 0; k < 3; k++) hi.d1[k] = r->o[k] + r->d[k] * t;
			V3 nn;
			float* np = &nn.x;
			for (int k = 0; k < 3; k++) {
				float a = (d->n1[k] - d->n0[k]) * bu;
				float b = (d->n2[k] - d->n0[k]) * bv;
				np[k] = (d->n0[k] + a) + b;
			}
			if (ik < numInst) {
				const float* A = &instA[12 * ik];
				float w[3];
				for (int c = 0; c < 3; c++) {
					const float a = A[c] * np[0] + A[4 + c] * np[1];
					w[c] = a + A[8 + c] * np[2];
				}
				nn = v3(w[0], w[1], w[2]);
			}
			nn = nrm(nn);
			hi.d2[0] = nn.x;
			hi.d2[1] = nn.y;
			hi.d2[2] = nn.z;
			float alb[4] = {1.0f, 1.0f, 1.0f, 1.0f}, rough = 1.0f, metal = 0.0f;
			if (d->materialIndex < numMats) {
				const OrMaterial* m = &mats[d->materialIndex];
				memcpy(alb, m->albedo, sizeof(alb));
				rough = m->f2[0];
				metal = m->f2[1];
				float uv[2];
				for (int k = 0; k < 2; k++) {
					float a = (d->uv1[k] - d->uv0[k]) * bu;
					float b = (d->uv2[k] - d->uv0[k]) * bv;
					uv[k] = (d->uv0[k] + a) + b;
				}
				const int32_t ai = m->i[0], ri = m->i[1], mi = m->i[2];
				if (ai >= 0 && (uint32_t)ai < numTex && texs[ai].bpp == 4u) {
					const uint8_t* p = texel_of(&texs[ai], uv[0], uv[1]);
					alb[0] = lin[p[0]];
					alb[1] = lin[p[1]];
					alb[2] = lin[p[2]];
					alb[3] = (float)p[3] / 255.0f;
				}
				if (ri >= 0 && (uint32_t)ri < numTex)
					rough = (float)texel_of(&texs[ri], uv[0], uv[1])[0] / 255.0f;
				if (mi >= 0 && (uint32_t)mi < numTex)
					metal = (float)texel_of(&texs[mi], uv[0], uv[1])[0] / 255.0f;
			}
			hi.d1[3] = rough;
			hi.d2[3] = metal;
			hi.d3[0] = alb[0];
			hi.d3[1] = alb[1];
			hi.d3[2] = alb[2];
			hi.d3[3] = fabsf(alb[3]);
		} else {
			hi.d3[3] = -1.0f;
		}
		out[i] = hi;
	}
}

void oracle_interpret(const OrRay* rays, const OrHit* hits, uint32_t n, const OrTriData* datas,
                      const OrMaterial* mats, uint32_t numMats, const OrTexture* texs,
                      uint32_t numTex, OrHitInfo* out)
{
	interpret_records(rays, hits, 0, n, datas, 0, 0, 0u, mats, numMats, texs, numTex, out);
}

/* shadeHit, MaterialKernel.cu:49-114, for one record: contribution and shadow ray per light,
 * light-major (record id + light * count). */
static void shade_hit(uint32_t id, uint32_t count, V3 mask, Pcg* rng, V3 n, V3 v, V3 p, V3 albedo,
                      float metal, float rough, const OrLight* lights, uint32_t numLights,
                      float* contrib, OrRay* shadow)
{
	for (uint32_t i = 0; i < numLights; i++) {
		const OrLight* L = &lights[i];
		V3 lp = ld3(L->pos);
		V3 toL = sub(lp, p);
		float dist = len(toL);
		V3 l = divs(toL, dist);
		float fo = oracle_falloff(dist, L->range);
		V3 sh = shade(n, v, albedo, rough, metal, l, dist, ld3(L->strength), L->range);
		V3 c = mulv(scale(mask, fo), sh);
		size_t s = (size_t)id + (size_t)i * count;
		contrib[4 * s + 0] = c.x;
		contrib[4 * s + 1] = c.y;
		contrib[4 * s + 2] = c.z;
		contrib[4 * s + 3] = 0.0f;
		const float eps = 0.0001f;
		int lit = c.x > eps || c.y > eps || c.z > eps;
		OrRay* r = &shadow[s];
		if (L->staticShadows && lit) {
			V3 cu = fabsf(n.z) > 0.01f ? nrm(v3(0.0f, -n.z, n.y)) : nrm(v3(-n.y, n.x, 0.0f));
			V3 cv = cross(cu, l);
			float r1 = pcg_uniform(rng);
			float r2 = pcg_uniform(rng);
			float cd = L->radius * (2.0f * r2 - 1.0f);
			float sn, cs;
			oracle_turn_sincos(r1, &sn, &cs);
			V3 off = add(scale(scale(cu, cs), cd), scale(scale(cv, sn), cd));
			V3 diff = sub(add(lp, off), p);
			V3 dir = nrm(diff);
			r->o[0] = p.x;
			r->o[1] = p.y;
			r->o[2] = p.z;
			r->tMin = 0.001f;
			r->d[0] = dir.x;
			r->d[1] = dir.y;
			r->d[2] = dir.z;
			r->tMax = len(diff);
		} else {
			set_dummy(r);
		}
	}
}

/* shadowLogicKernel, MaterialKernel.cu:361-388: colour (+)= sum of lit contributions, lights in
 * order; records of a 1/scale-resolution pass map by loc / scale (clamped to that pass). */
static void shadow_logic(float* img, uint32_t w, uint32_t h, uint32_t sc, int addTo,
                         const uint8_t* inLight, const float* contrib, uint32_t numLights)
{
	uint32_t sw = w / sc, sh = h / sc;
	size_t sN = (size_t)sw * sh;
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			uint32_t sx = x / sc, sy = y / sc;
			if (sx >= sw) sx = sw - 1u;
			if (sy >= sh) sy = sh - 1u;
			uint32_t sid = pixel_id(sx, sy, sw, sh);
			float* px = &img[4 * ((size_t)y * w + x)];
			float c[3] = {0.0f, 0.0f, 0.0f};
			if (addTo) {
				c[0] = px[0];
				c[1] = px[1];
				c[2] = px[2];
			}
			for (uint32_t i = 0; i < numLights; i++) {
				size_t s = (size_t)sid + (size_t)i * sN;
				if (inLight[s]) {
					c[0] = c[0] + contrib[4 * s + 0];
					c[1] = c[1] + contrib[4 * s + 1];
					c[2] = c[2] + contrib[4 * s + 2];
				}
			}
			px[0] = c[0];
			px[1] = c[1];
			px[2] = c[2];
			px[3] = 1.0f;
		}
}

void oracle_primary_rays_jittered(const float cam[12], uint32_t w, uint32_t h, float jx, float jy,
                                  float tMin, float tMax, OrRay* rays)
{
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			float cx = (((float)x + jx) / (float)w) * 2.0f - 1.0f;
			float cy = (((float)y + jy) / (float)h) * 2.0f - 1.0f;
			float d[3];
			for (int k = 0; k < 3; k++) {
				float a = cam[6 + k] * cx;
				float s = cam[3 + k] + a;
				float b = cam[9 + k] * cy;
				d[k] = s + b;
			}
			float dd = d[0] * d[0] + d[1] * d[1];
			dd = dd + d[2] * d[2];
			float l = sqrtf(dd);
			OrRay* r = &rays[pixel_id(x, y, w, h)];
			for (int k = 0; k < 3; k++) {
				r->o[k] = cam[k];
				r->d[k] = d[k] / l;
			}
			r->tMin = tMin;
			r->tMax = tMax;
		}
}

/* Bounce ray of each hit (phx_gen_bounce_rays): origin o + d*t, interpolated normal turned to
 * face the ray, cosine-weighted direction, tMin 1e-3, tMax FLT_MAX; misses get the dummy ray. */
void oracle_bounce_rays(const OrTriData* datas, const OrRay* rays, const OrHit* hits, uint32_t n,
                        uint64_t seed, uint32_t frame, OrRay* out)
{
	for (uint32_t i = 0; i < n; i++) {
		OrRay* o = &out[i];
		if (hits[i].tri == 0xFFFFFFFFu) {
			set_dummy(o);
			continue;
		}
		const OrRay* r = &rays[i];
		const OrTriData* d = &datas[hits[i].tri];
		float t = hits[i].t, bu = hits[i].u, bv = hits[i].v;
		float pos[3];
		for (int k = 0; k < 3; k++) pos[k] = r->o[k] + r->d[k] * t;
		float nn[3];
		for (int k = 0; k < 3; k++) {
			float a = (d->n1[k] - d->n0[k]) * bu;
			float b = (d->n2[k] - d->n0[k]) * bv;
			nn[k] = (d->n0[k] + a) + b;
		}
		V3 N = nrm(v3(nn[0], nn[1], nn[2]));
		float dn = N.x * r->d[0] + N.y * r->d[1];
		dn = dn + N.z * r->d[2];
		if (dn > 0.0f) N = v3(-N.x, -N.y, -N.z);
		Pcg g = frame_rng(seed, frame, 0u, i);
		float r1 = pcg_uniform(&g);
		float r2 = pcg_uniform(&g);
		V3 dir = sample_lambert(r1, r2, N);
		o->o[0] = pos[0];
		o->o[1] = pos[1];
		o->o[2] = pos[2];
		o->tMin = 0.001f;
		o->d[0] = dir.x;
		o->d[1] = dir.y;
		o->d[2] = dir.z;
		o->tMax = FLT_MAX;
	}
}

/* One frame (CudaTracerRenderer.cpp:637-720, the "normal path") with the G-buffer traced
 * instead of rasterised: primary rays (jitter (jx, jy) in pixels), closest hit, interpret, shade
 * every light + shadow rays, any-hit, shadow logic; one bounce per 2x2 pixel block (random pixel
 * of the block, diffuse or GGX by roughness, throughput albedo * 0.6), closest hit, interpret,
 * shade x throughput + shadow rays, any-hit, shadow logic added at half resolution.
 * With instances (numInst > 0), closest hits and any hits are oracle_cast_instances' over the scene
 * and the instances, and hit records of instance hits read the instance's mesh data.
 * Outputs: colour (w*h RGBA, row-major); optional G-buffer (record order) and bounce rays. */
void oracle_frame(const OrNode* nodes, const OrTri* tris, const OrTriData* datas,
                  const OrMaterial* mats, uint32_t numMats, const OrTexture* texs,
                  uint32_t numTex, const OrLight* lights,
                  uint32_t numLights, const float cam[12], uint32_t w, uint32_t h, float jx,
                  float jy, uint64_t seed, uint32_t frame, float* color, OrHitInfo* gbufOut,
                  OrRay* bounceOut, uint32_t threads, const OrInstance* inst, uint32_t numInst)
{
	const uint32_t N = w * h;
	const uint32_t hw = w / 2u, hh = h / 2u, M = hw * hh;
	const V3 camPos = v3(cam[0], cam[1], cam[2]);
	float* instA = (float*)malloc(sizeof(float) * 12 * (numInst ? numInst : 1));
	for (uint32_t k = 0; k < numInst; k++) {
		float mn[3], mx[3];
		oracle_instance_prepare(&inst[k], &instA[12 * k], mn, mx);
	}
	uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
	OrRay* prim = (OrRay*)malloc(sizeof(OrRay) * N);
	OrHit* ph = (OrHit*)malloc(sizeof(OrHit) * N);
	OrHitInfo* gb = (OrHitInfo*)malloc(sizeof(OrHitInfo) * N);
	size_t ns = (size_t)N * numLights;
	float* contrib = (float*)malloc(sizeof(float) * 4 * (ns ? ns : 1));
	OrRay* shadow = (OrRay*)malloc(sizeof(OrRay) * (ns ? ns : 1));
	uint8_t* inl = (uint8_t*)malloc(ns ? ns : 1);

	oracle_primary_rays_jittered(cam, w, h, jx, jy, 1e-4f, FLT_MAX, prim);
	if (numInst) {
		oracle_cast_instances(nodes, tris, inst, numInst, prim, N, ph, ids, 0);
		interpret_records(prim, ph, ids, N, datas, inst, instA, numInst, mats, numMats, texs,
		                  numTex, gb);
	} else {
		oracle_cast_closest(nodes, tris, prim, N, ph, threads);
		oracle_interpret(prim, ph, N, datas, mats, numMats, texs, numTex, gb);
	}

	/* gBufferMaterialKernel (MaterialKernel.cu:119-140); a pixel without a hit gets dummy rays
	 * and no contribution, as materialKernel does for a secondary miss (:151-164). */
	for (uint32_t id = 0; id < N; id++) {
		const OrHitInfo* g = &gb[id];
		if (!(g->d3[3] >= 0.0f)) {
			for (uint32_t i = 0; i < numLights; i++) {
				size_t s = (size_t)id + (size_t)i * N;
				set_dummy(&shadow[s]);
				contrib[4 * s] = contrib[4 * s + 1] = contrib[4 * s + 2] = contrib[4 * s + 3] = 0.0f;
			}
			continue;
		}
		Pcg rng = frame_rng(seed, frame, 1u, id);
		V3 p = ld3(g->d1), n = ld3(g->d2);
		V3 toCam = nrm(sub(camPos, p));
		shade_hit(id, N, v3(1.0f, 1.0f, 1.0f), &rng, n, toCam, p, ld3(g->d3), g->d2[3], g->d1[3],
		          lights, numLights, contrib, shadow);
	}
	if (numInst)
		oracle_cast_instances(nodes, tris, inst, numInst, shadow, (uint32_t)ns, 0, 0, inl);
	else
		oracle_cast_anyhit(nodes, tris, shadow, (uint32_t)ns, inl, threads);
	shadow_logic(color, w, h, 1u, 0, inl, contrib, numLights);

	/* createSecondaryRaysKernel (MaterialKernel.cu:251-312) on the half-resolution grid. */
	if (M > 0) {
		OrRay* sec = (OrRay*)malloc(sizeof(OrRay) * M);
		float* tp = (float*)malloc(sizeof(float) * 3 * M);
		OrHit* sh = (OrHit*)malloc(sizeof(OrHit) * M);
		OrHitInfo* sgb = (OrHitInfo*)malloc(sizeof(OrHitInfo) * M);
		for (uint32_t y = 0; y < hh; y++)
			for (uint32_t x = 0; x < hw; x++) {
				uint32_t hid = pixel_id(x, y, hw, hh);
				Pcg rng = frame_rng(seed, frame, 2u, hid);
				uint32_t pix = pcg_next(&rng) % 4u;
				uint32_t fx = 2u * x + pix % 2u, fy = 2u * y + pix / 2u;
				const OrHitInfo* g = &gb[pixel_id(fx, fy, w, h)];
				OrRay* r = &sec[hid];
				if (!(g->d3[3] >= 0.0f)) {
					set_dummy(r);
					tp[3 * hid] = tp[3 * hid + 1] = tp[3 * hid + 2] = 0.0f;
					continue;
				}
				V3 gp = ld3(g->d1), gn = ld3(g->d2), alb = ld3(g->d3);
				float rough = g->d1[3];
				V3 fromCam = nrm(sub(gp, camPos));
				float r1 = pcg_uniform(&rng);
				float r2 = pcg_uniform(&rng);
				float tMin = 0.001f;
				V3 dir, t3;
				if (fabsf(g->d3[3]) > 0.5f) {
					float pick = pcg_uniform(&rng);
					t3 = scale(alb, 0.6f);
					if (pick < rough) {
						dir = sample_lambert(r1, r2, gn);
					} else {
						V3 refl = sub(fromCam, scale(gn, 2.0f * dot(gn, fromCam)));
						dir = sample_ggx(r1, r2, rough * rough, refl, gn);
					}
				} else {
					dir = nrm(add(scale(fromCam, 0.8f), scale(v3(-gn.x, -gn.y, -gn.z), 0.2f)));
					t3 = v3(1.0f, 1.0f, 1.0f);
					tMin = 0.2f;
				}
				r->o[0] = gp.x;
				r->o[1] = gp.y;
				r->o[2] = gp.z;
				r->tMin = tMin;
				r->d[0] = dir.x;
				r->d[1] = dir.y;
				r->d[2] = dir.z;
				r->tMax = FLT_MAX;
				tp[3 * hid] = t3.x;
				tp[3 * hid + 1] = t3.y;
				tp[3 * hid + 2] = t3.z;
			}
		if (bounceOut) memcpy(bounceOut, sec, sizeof(OrRay) * M);
		if (numInst) {
			oracle_cast_instances(nodes, tris, inst, numInst, sec, M, sh, ids, 0);
			interpret_records(sec, sh, ids, M, datas, inst, instA, numInst, mats, numMats, texs,
			                  numTex, sgb);
		} else {
			oracle_cast_closest(nodes, tris, sec, M, sh, threads);
			oracle_interpret(sec, sh, M, datas, mats, numMats, texs, numTex, sgb);
		}
		/* materialKernel (MaterialKernel.cu:145-184) */
		size_t ms = (size_t)M * numLights;
		for (uint32_t id = 0; id < M; id++) {
			const OrHitInfo* g = &sgb[id];
			if (!(g->d3[3] >= 0.0f)) {
				for (uint32_t i = 0; i < numLights; i++) {
					size_t s = (size_t)id + (size_t)i * M;
					set_dummy(&shadow[s]);
					contrib[4 * s] = contrib[4 * s + 1] = contrib[4 * s + 2] = contrib[4 * s + 3] =
						0.0f;
				}
				continue;
			}
			Pcg rng = frame_rng(seed, frame, 3u, id);
			const OrRay* r = &sec[id];
			V3 v = v3(-r->d[0], -r->d[1], -r->d[2]);
			shade_hit(id, M, ld3(&tp[3 * id]), &rng, ld3(g->d2), v, ld3(g->d1), ld3(g->d3),
			          g->d2[3], g->d1[3], lights, numLights, contrib, shadow);
		}
		if (numInst)
			oracle_cast_instances(nodes, tris, inst, numInst, shadow, (uint32_t)ms, 0, 0, inl);
		else
			oracle_cast_anyhit(nodes, tris, shadow, (uint32_t)ms, inl, threads);
		shadow_logic(color, w, h, 2u, 1, inl, contrib, numLights);
		free(sec);
		free(tp);
		free(sh);
		free(sgb);
	}
	if (gbufOut) memcpy(gbufOut, gb, sizeof(OrHitInfo) * N);
	free(instA);
	free(ids);
	free(prim);
	free(ph);
	free(gb);
	free(contrib);
	free(shadow);
	free(inl);
}

/* Continuous pixel coordinates (px, py) -- the (x + jx, y + jy) of the primary rays -- at which
 * the camera cam (CameraDef: origin, dir, dX, dY) sees the point origin + v; v may be a direction
 * (a point at infinity). Inverts the pinhole formula normalize(dir + cx dX + cy dY), cx = px / w *
 * 2 - 1, using that dir, dX and dY are orthogonal: cx = (v . dX) / ((v . dir) |dX|^2). This is
 * getUVCoord of velocity_gen.frag:40-47 for the tracer's camera. Returns 0 for a point that is
 * not in front of the camera. */
static int project(const float cam[12], V3 v, uint32_t w, uint32_t h, float* px, float* py)
{
	const V3 dir = v3(cam[3], cam[4], cam[5]), dX = v3(cam[6], cam[7], cam[8]),
	         dY = v3(cam[9], cam[10], cam[11]);
	const float a = dot(v, dir);
	if (!(a > 0.0f)) return 0;
	const float cx = dot(v, dX) / (a * dot(dX, dX));
	const float cy = dot(v, dY) / (a * dot(dY, dY));
	*px = ((cx + 1.0f) * 0.5f) * (float)w;
	*py = ((cy + 1.0f) * 0.5f) * (float)h;
	return 1;
}

/* Screen-space velocity of every pixel (velocity_gen.frag:49-71), in pixels, row-major (w*h*2):
 * the pixel's primary hit (G-buffer record position; for a miss the primary ray's direction, a
 * point at infinity) projected by the current camera minus its projection by the previous camera.
 * A point behind the previous camera gets FLT_MAX (its history is off screen). */
void oracle_velocity(const float cam[12], const float prevCam[12], const OrHitInfo* gbuf,
                     const OrRay* prim, uint32_t w, uint32_t h, float* vel)
{
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			const uint32_t rec = pixel_id(x, y, w, h);
			V3 vc, vp;
			if (gbuf[rec].d3[3] >= 0.0f) {
				const V3 pos = ld3(gbuf[rec].d1);
				vc = sub(pos, v3(cam[0], cam[1], cam[2]));
				vp = sub(pos, v3(prevCam[0], prevCam[1], prevCam[2]));
			} else {
				vc = vp = ld3(prim[rec].d);
			}
			float cx = 0.0f, cy = 0.0f, px = 0.0f, py = 0.0f;
			float* o = &vel[2 * ((size_t)y * w + x)];
			if (project(cam, vc, w, h, &cx, &cy) && project(prevCam, vp, w, h, &px, &py)) {
				o[0] = cx - px;
				o[1] = cy - py;
			} else {
				o[0] = FLT_MAX;
				o[1] = FLT_MAX;
			}
		}
}

/* Bilinear sample of a w x h image of `ch` floats per pixel at continuous pixel coordinates
 * (hx, hy) (texel centres at i + 0.5), edge texels repeated (GL_LINEAR with clamp to edge):
 * rows blended with lerp (1 - f) * a + f * b (MathHelpers.inl:66-70), first along x. */
static void bilinear(const float* img, uint32_t ch, uint32_t w, uint32_t h, float hx, float hy,
                     float* o)
{
	const float tx = hx - 0.5f, ty = hy - 0.5f;
	const float x0f = floorf(tx), y0f = floorf(ty);
	const float fx = tx - x0f, fy = ty - y0f;
	const int x0 = (int)x0f, y0 = (int)y0f;
	const int xa = x0 < 0 ? 0 : (x0 > (int)w - 1 ? (int)w - 1 : x0);
	const int xb = x0 + 1 < 0 ? 0 : (x0 + 1 > (int)w - 1 ? (int)w - 1 : x0 + 1);
	const int ya = y0 < 0 ? 0 : (y0 > (int)h - 1 ? (int)h - 1 : y0);
	const int yb = y0 + 1 < 0 ? 0 : (y0 + 1 > (int)h - 1 ? (int)h - 1 : y0 + 1);
	for (uint32_t c = 0; c < ch; c++) {
		const float a = img[((size_t)ya * w + xa) * ch + c], b = img[((size_t)ya * w + xb) * ch + c];
		const float d = img[((size_t)yb * w + xa) * ch + c], e = img[((size_t)yb * w + xb) * ch + c];
		const float top = a * (1.0f - fx) + b * fx;
		const float bot = d * (1.0f - fx) + e * fx;
		o[c] = top * (1.0f - fy) + bot * fy;
	}
}

/* TAA resolve (taa.frag:25-71), FP32 history. With this frame's velocity vel (NULL: zero) the
 * history is read bilinearly at (x + 0.5, y + 0.5) - vel; a reprojected position outside the
 * image starts the pixel afresh. The previous frame's velocity velIn (NULL: zero) at that position
 * sets the velocity distrust min(1, 1.2 sqrt(| |v| - |v_hist| |)), the roughness the roughness
 * distrust mix(3, 1, roughness). With clampHistory != 0 the history colour is clamped to the
 * current frame's 3x3 neighbourhood min/max (not in the reference). cur/histIn/histOut/out: w*h
 * RGBA row-major; roughness: w*h (may be NULL = 0); vel, velIn: w*h*2. reset != 0 starts the
 * history from this frame. */
void oracle_taa(const float* cur, const float* histIn, const float* roughness, const float* velIn,
                const float* vel, uint32_t w, uint32_t h, int reset, int clampHistory,
                float* histOut, float* out)
{
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			size_t p = (size_t)y * w + x;
			const float* c = &cur[4 * p];
			const float vx = vel ? vel[2 * p] : 0.0f, vy = vel ? vel[2 * p + 1] : 0.0f;
			const float hx = ((float)x + 0.5f) - vx, hy = ((float)y + 0.5f) - vy;
			const int onScreen = hx >= 0.0f && hx <= (float)w && hy >= 0.0f && hy <= (float)h;
			float o[3], n2;
			if (reset || !onScreen) {
				o[0] = c[0];
				o[1] = c[1];
				o[2] = c[2];
				n2 = 1.0f;
			} else {
				float pv[2] = {0.0f, 0.0f};
				if (velIn) bilinear(velIn, 2u, w, h, hx, hy, pv);
				const float lv = sqrtf(vx * vx + vy * vy);
				const float lp = sqrtf(pv[0] * pv[0] + pv[1] * pv[1]);
				float rough = roughness ? roughness[p] : 0.0f;
				float vDistrust = min_num(1.0f, 1.2f * sqrtf(fabsf(lv - lp)));
				float rDistrust = 3.0f * (1.0f - rough) + 1.0f * rough;
				float trust = clamp_num(1.0f - rDistrust * vDistrust, 0.0f, 1.0f);
				float hs[4];
				bilinear(histIn, 4u, w, h, hx, hy, hs);
				float hc[3] = {hs[0], hs[1], hs[2]};
				if (clampHistory) {
					float lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
					uint32_t y0 = y ? y - 1u : 0u, y1 = y + 1u < h ? y + 1u : h - 1u;
					uint32_t x0 = x ? x - 1u : 0u, x1 = x + 1u < w ? x + 1u : w - 1u;
					for (uint32_t yy = y0; yy <= y1; yy++)
						for (uint32_t xx = x0; xx <= x1; xx++) {
							const float* q = &cur[4 * ((size_t)yy * w + xx)];
							for (int k = 0; k < 3; k++) {
								lo[k] = min_num(lo[k], q[k]);
								hi[k] = max_num(hi[k], q[k]);
							}
						}
					for (int k = 0; k < 3; k++) hc[k] = clamp_num(hc[k], lo[k], hi[k]);
				}
				n2 = hs[3] * trust + 1.0f;
				float blend = 1.0f - 1.0f / n2;
				for (int k = 0; k < 3; k++) o[k] = c[k] * (1.0f - blend) + hc[k] * blend;
			}
			float* ho = &histOut[4 * p];
			float* oo = &out[4 * p];
			for (int k = 0; k < 3; k++) {
				ho[k] = o[k];
				oo[k] = o[k];
			}
			ho[3] = n2;
			oo[3] = 1.0f;
		}
}

/* Halton(2,3) pixel jitter of frame f: radical inverses of f % 16 + 1 in bases 2 and 3
 * (HALTON_SEQ, GameScreen.cpp:21-38, lists the same 16 points to 12 digits). */
void oracle_halton23(uint32_t f, float* jx, float* jy)
{
	uint32_t i = f % 16u + 1u;
	double r = 0.0, inv = 0.5;
	for (uint32_t k = i; k; k /= 2u, inv *= 0.5) r += (double)(k % 2u) * inv;
	*jx = (float)r;
	r = 0.0;
	inv = 1.0 / 3.0;
	for (uint32_t k = i; k; k /= 3u, inv /= 3.0) r += (double)(k % 3u) * inv;
	*jy = (float)r;
}
