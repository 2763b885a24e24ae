/*
 * oracle.c -- CPU restatement of the reference's ray-query path. TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
 *
 * Built with -O2 -ffp-contract=off (no fast-math) on x86-64, where float arithmetic is IEEE single
 * precision (SSE), so every expression below is evaluated exactly in the written order.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <unistd.h>

/* std::min(a, b) = (b < a) ? b : a and std::max(a, b) = (a < b) ? b : a: the first operand is
 * returned when the comparison is false (NaN, +-0). sfz::min/max/minElement/maxElement build on
 * them (externals/sfzCore/include/sfz/math/Vector.inl:425-470, 569-601). */
static inline float min_first(float a, float b) { return (b < a) ? b : a; }
static inline float max_first(float a, float b) { return (a < b) ? b : a; }

typedef struct Ray {
	float o[3], d[3], inv[3];
	float tMin, tMax;
} Ray;

/* Ray(origin, dir): invDir = vec3(1) / dir, elementwise (Ray.hpp:22-37, Vector.inl:904-910). */
static inline void make_ray(const OrRay* in, Ray* r)
{
	for (int k = 0; k < 3; k++) {
		r->o[k] = in->o[k];
		r->d[k] = in->d[k];
		r->inv[k] = 1.0f / in->d[k];
	}
	r->tMin = in->tMin;
	r->tMax = in->tMax;
}

/* intersects(ray, min, max), Intersection.hpp:77-90. */
static inline int aabb(const Ray* r, const float* mn, const float* mx, float* tIn, float* tOut)
{
	float t1[3], t2[3], lo[3], hi[3];
	for (int k = 0; k < 3; k++) {
		t1[k] = (mn[k] - r->o[k]) * r->inv[k];
		t2[k] = (mx[k] - r->o[k]) * r->inv[k];
	}
	for (int k = 0; k < 3; k++) {
		lo[k] = min_first(t1[k], t2[k]);
		hi[k] = max_first(t1[k], t2[k]);
	}
	float tmin = max_first(max_first(lo[0], lo[1]), lo[2]);
	float tmax = min_first(min_first(hi[0], hi[1]), hi[2]);
	*tIn = tmin;
	*tOut = tmax;
	return tmax >= tmin;
}

/* dot = x*x' + y*y' + z*z' evaluated left to right (Vector.inl:345-351). */
static inline float dot3(const float* a, const float* b)
{
	float s = a[0] * b[0];
	s = s + a[1] * b[1];
	s = s + a[2] * b[2];
	return s;
}

/* cross (Vector.inl:362-367). */
static inline void cross3(const float* a, const float* b, float* out)
{
	out[0] = a[1] * b[2] - a[2] * b[1];
	out[1] = a[2] * b[0] - a[0] * b[2];
	out[2] = a[0] * b[1] - a[1] * b[0];
}

/* intersects(tri, origin, dir), Intersection.hpp:24-67 (Moller-Trumbore, RTR3 p.750). */
static inline int triangle(const float* p0, const float* p1, const float* p2, const float* o,
                           const float* d, float* tOut, float* uOut, float* vOut)
{
	const float EPS = 0.00001f;
	float e1[3], e2[3], q[3], s[3], r[3];
	for (int k = 0; k < 3; k++) {
		e1[k] = p1[k] - p0[k];
		e2[k] = p2[k] - p0[k];
	}
	cross3(d, e2, q);
	float a = dot3(e1, q);
	if (-EPS < a && a < EPS) return 0;
	float f = 1.0f / a;
	for (int k = 0; k < 3; k++) s[k] = o[k] - p0[k];
	float u = f * dot3(s, q);
	if (u < 0.0f) return 0;
	cross3(s, e1, r);
	float v = f * dot3(d, r);
	if (v < 0.0f || (u + v) > 1.0f) return 0;
	float t = f * dot3(e2, r);
	*tOut = t;
	*uOut = u;
	*vOut = v;
	return 1;
}

#define STACK_MAX 144

/* castRay / castDebugRay, BVHTraversal.hpp:37-235. `cnt` may be NULL. */
static void cast_one(const OrNode* nodes, const OrTri* tris, const OrRay* in, OrHit* out,
                     uint32_t* cnt)
{
	Ray ray;
	make_ray(in, &ray);
	const float tMin = ray.tMin, tMax = ray.tMax;
	uint32_t stack[STACK_MAX];
	stack[0] = 0u;
	uint32_t sp = 1u;
	OrHit c = {~0u, FLT_MAX, FLT_MAX, FLT_MAX};
	uint32_t nodesVisited = 0, trisHit = 0, triTests = 0, aabbTests = 0, aabbHits = 0;

	while (sp > 0u) {
		sp--;
		const OrNode* node = &nodes[stack[sp]];
		nodesVisited++;
		float lmn[3] = {node->f[0], node->f[1], node->f[2]};
		float lmx[3] = {node->f[3], node->f[4], node->f[5]};
		float rmn[3] = {node->f[6], node->f[7], node->f[8]};
		float rmx[3] = {node->f[9], node->f[10], node->f[11]};
		float lIn, lOut, rIn, rOut;
		int lHit = aabb(&ray, lmn, lmx, &lIn, &lOut);
		int rHit = aabb(&ray, rmn, rmx, &rIn, &rOut);
		aabbTests += 2;
		float tCurrMax = min_first(tMax, c.t);
		int visit[2];
		visit[0] = lHit && lOut > tMin && lIn < tCurrMax;
		visit[1] = rHit && rOut > tMin && rIn < tCurrMax;

		for (int side = 0; side < 2; side++) {
			if (!visit[side]) continue;
			aabbHits++;
			int32_t raw = node->i[side];
			uint32_t idx = (uint32_t)(raw < 0 ? ~raw : raw);
			int32_t count = node->i[2 + side];
			if (count == 0) {
				stack[sp] = idx;
				sp += 1u;
				continue;
			}
			for (int32_t k = 0; k < count; k++) {
				const OrTri* tri = &tris[idx + (uint32_t)k];
				float t, u, v;
				int hit = triangle(tri->v[0], tri->v[1], tri->v[2], ray.o, ray.d, &t, &u, &v);
				triTests++;
				if (hit && t < c.t && tMin <= t && t <= tMax) {
					trisHit++;
					c.tri = idx + (uint32_t)k;
					c.t = t;
					c.u = u;
					c.v = v;
				}
			}
		}
	}
	*out = c;
	if (cnt) {
		cnt[0] = nodesVisited;
		cnt[1] = trisHit;
		cnt[2] = triTests;
		cnt[3] = aabbTests;
		cnt[4] = aabbHits;
	}
}

/* Any hit: castRay's traversal with tCurrMax = tMax, ending at the first triangle with
 * tMin <= t < tMax. `cnt` (may be NULL) receives castDebugRay-style counters of that traversal
 * (trianglesIntersected = 1 if occluded). */
static int anyhit_counted(const OrNode* nodes, const OrTri* tris, const OrRay* in, uint32_t* cnt)
{
	Ray ray;
	make_ray(in, &ray);
	const float tMin = ray.tMin, tMax = ray.tMax;
	uint32_t stack[STACK_MAX];
	stack[0] = 0u;
	uint32_t sp = 1u;
	uint32_t c[5] = {0, 0, 0, 0, 0};
	int occluded = 0;
	while (sp > 0u && !occluded) {
		sp--;
		const OrNode* node = &nodes[stack[sp]];
		c[0]++;
		c[3] += 2;
		float lIn, lOut, rIn, rOut;
		int lHit = aabb(&ray, &node->f[0], &node->f[3], &lIn, &lOut);
		int rHit = aabb(&ray, &node->f[6], &node->f[9], &rIn, &rOut);
		int visit[2];
		visit[0] = lHit && lOut > tMin && lIn < tMax;
		visit[1] = rHit && rOut > tMin && rIn < tMax;
		for (int side = 0; side < 2 && !occluded; side++) {
			if (!visit[side]) continue;
			c[4]++;
			int32_t raw = node->i[side];
			uint32_t idx = (uint32_t)(raw < 0 ? ~raw : raw);
			int32_t count = node->i[2 + side];
			if (count == 0) {
				stack[sp] = idx;
				sp += 1u;
				continue;
			}
			for (int32_t k = 0; k < count; k++) {
				const OrTri* tri = &tris[idx + (uint32_t)k];
				float t, u, v;
				c[2]++;
				if (triangle(tri->v[0], tri->v[1], tri->v[2], ray.o, ray.d, &t, &u, &v) &&
				    tMin <= t && t < tMax) {
					c[1]++;
					occluded = 1;
					break;
				}
			}
		}
	}
	if (cnt)
		for (int k = 0; k < 5; k++) cnt[k] = c[k];
	return occluded;
}

static int anyhit_one(const OrNode* nodes, const OrTri* tris, const OrRay* in)
{
	return anyhit_counted(nodes, tris, in, 0);
}

/* ---- threading: chunks of rays pulled from an atomic counter ---- */

typedef struct Job {
	const OrNode* nodes;
	const OrTri* tris;
	const OrRay* rays;
	uint32_t n;
	OrHit* hits;
	uint8_t* inLight;
	uint32_t* perRay;
	_Atomic uint32_t next;
	_Atomic uint64_t sum[5];
	int mode; /* 0 closest, 1 debug, 2 anyhit */
} Job;

#define CHUNK 1024u

static void* worker(void* arg)
{
	Job* j = (Job*)arg;
	uint64_t local[5] = {0, 0, 0, 0, 0};
	for (;;) {
		uint32_t b = atomic_fetch_add(&j->next, CHUNK);
		if (b >= j->n) break;
		uint32_t e = b + CHUNK < j->n ? b + CHUNK : j->n;
		for (uint32_t i = b; i < e; i++) {
			if (j->mode == 3) {
				uint32_t c[5];
				j->inLight[i] = (uint8_t)!anyhit_counted(j->nodes, j->tris, &j->rays[i], c);
				for (int k = 0; k < 5; k++) local[k] += c[k];
			} else if (j->mode == 2) {
				j->inLight[i] = (uint8_t)!anyhit_one(j->nodes, j->tris, &j->rays[i]);
			} else if (j->mode == 1) {
				uint32_t c[5];
				cast_one(j->nodes, j->tris, &j->rays[i], &j->hits[i], c);
				for (int k = 0; k < 5; k++) local[k] += c[k];
				if (j->perRay)
					for (int k = 0; k < 5; k++) j->perRay[5 * (size_t)i + k] = c[k];
			} else {
				cast_one(j->nodes, j->tris, &j->rays[i], &j->hits[i], 0);
			}
		}
	}
	for (int k = 0; k < 5; k++) atomic_fetch_add(&j->sum[k], local[k]);
	return 0;
}

static void run(Job* j, uint32_t numThreads)
{
	if (numThreads == 0) {
		long c = sysconf(_SC_NPROCESSORS_ONLN);
		numThreads = c > 0 ? (uint32_t)c : 1u;
	}
	if (numThreads > 256) numThreads = 256;
	atomic_store(&j->next, 0u);
	for (int k = 0; k < 5; k++) atomic_store(&j->sum[k], 0u);
	pthread_t th[256];
	uint32_t started = 0;
	for (uint32_t t = 1; t < numThreads; t++)
		if (pthread_create(&th[started], 0, worker, j) == 0) started++;
	worker(j);
	for (uint32_t t = 0; t < started; t++) pthread_join(th[t], 0);
}

void oracle_cast_closest(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                         OrHit* hits, uint32_t numThreads)
{
	Job j = {nodes, tris, rays, n, hits, 0, 0};
	j.mode = 0;
	run(&j, numThreads);
}

void oracle_cast_debug(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                       OrHit* hits, OrCounters* total, uint32_t* perRay, uint32_t numThreads)
{
	Job j = {nodes, tris, rays, n, hits, 0, perRay};
	j.mode = 1;
	run(&j, numThreads);
	if (total) {
		total->nodesVisited = atomic_load(&j.sum[0]);
		total->trianglesIntersected = atomic_load(&j.sum[1]);
		total->triangleIntersectionTests = atomic_load(&j.sum[2]);
		total->aabbIntersectionTests = atomic_load(&j.sum[3]);
		total->aabbIntersections = atomic_load(&j.sum[4]);
	}
}

void oracle_cast_anyhit(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                        uint8_t* inLight, uint32_t numThreads)
{
	Job j = {nodes, tris, rays, n, 0, inLight, 0};
	j.mode = 2;
	run(&j, numThreads);
}

void oracle_anyhit_debug(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                         uint8_t* inLight, OrCounters* total, uint32_t numThreads)
{
	Job j = {nodes, tris, rays, n, 0, inLight, 0};
	j.mode = 3;
	run(&j, numThreads);
	total->nodesVisited = atomic_load(&j.sum[0]);
	total->trianglesIntersected = atomic_load(&j.sum[1]);
	total->triangleIntersectionTests = atomic_load(&j.sum[2]);
	total->aabbIntersectionTests = atomic_load(&j.sum[3]);
	total->aabbIntersections = atomic_load(&j.sum[4]);
}

void oracle_brute_closest(const OrTri* tris, uint32_t numTris, const OrRay* rays, uint32_t n,
                          OrHit* hits)
{
	for (uint32_t i = 0; i < n; i++) {
		Ray ray;
		make_ray(&rays[i], &ray);
		OrHit c = {~0u, FLT_MAX, FLT_MAX, FLT_MAX};
		for (uint32_t k = 0; k < numTris; k++) {
			float t, u, v;
			if (triangle(tris[k].v[0], tris[k].v[1], tris[k].v[2], ray.o, ray.d, &t, &u, &v) &&
			    t < c.t && ray.tMin <= t && t <= ray.tMax) {
				c.tri = k;
				c.t = t;
				c.u = u;
				c.v = v;
			}
		}
		hits[i] = c;
	}
}

/* CPUTracerRenderer.cpp:93-97 with sfz operation order: locNormalized = loc / res,
 * c = locNormalized * 2 - 1, dir = normalize(camDir + c.x * dX + c.y * dY). */
void oracle_primary_rays(const float cam[12], uint32_t w, uint32_t h, float tMin, float tMax,
                         uint32_t order, OrRay* rays)
{
	const float* org = cam;
	const float* dir = cam + 3;
	const float* dX = cam + 6;
	const float* dY = cam + 9;
	for (uint32_t y = 0; y < h; y++)
		for (uint32_t x = 0; x < w; x++) {
			float cx = ((float)x / (float)w) * 2.0f - 1.0f;
			float cy = ((float)y / (float)h) * 2.0f - 1.0f;
			float d[3];
			for (int k = 0; k < 3; k++) {
				float s = dir[k] + cx * dX[k];
				d[k] = s + cy * dY[k];
			}
			float len = sqrtf(dot3(d, d));
			size_t id;
			if (order == 1u) {
				uint32_t tile = (y / 8u) * (w / 8u) + (x / 8u);
				id = (size_t)tile * 64u + (y % 8u) * 8u + (x % 8u);
			} else {
				id = (size_t)y * w + x;
			}
			OrRay* r = &rays[id];
			for (int k = 0; k < 3; k++) {
				r->o[k] = org[k];
				r->d[k] = d[k] / len;
			}
			r->tMin = tMin;
			r->tMax = tMax;
		}
}

int oracle_triangle(const float p0[3], const float p1[3], const float p2[3], const float o[3],
                    const float d[3], float* t, float* u, float* v)
{
	return triangle(p0, p1, p2, o, d, t, u, v);
}

int oracle_aabb(const float o[3], const float d[3], const float mn[3], const float mx[3],
                float* tIn, float* tOut)
{
	OrRay in = {{o[0], o[1], o[2]}, 0.0f, {d[0], d[1], d[2]}, 0.0f};
	Ray r;
	make_ray(&in, &r);
	return aabb(&r, mn, mx, tIn, tOut);
}

/* ---- instances: the two-level path (phx_instances_build, phx_cast_*_instances) ----
 *
 * The library's definition, restated: an instance is traced iff the ray enters its world box
 * (castRay's box test with tIn <= min(tMax, best t)); it runs castRay on its mesh with the ray in
 * object space (o' = A o + b, d' = A d, the same tMin/tMax), from a miss; the closest result wins,
 * on equal t the static scene, then the lowest instance index. Instances are taken here in index
 * order, one after the other (the device walks a top tree instead; the box test never drops an
 * instance that could win or tie, so the order does not change the result). */

/* World -> object rows (A | b) of the object -> world rows (R | t): the adjugate of R over its
 * determinant, b = -A t, in double, each value rounded once to float. */
static int invert_affine(const float m[12], float out[12])
{
	double r[3][3], t[3];
	for (int i = 0; i < 3; i++) {
		for (int j = 0; j < 3; j++) r[i][j] = m[4 * i + j];
		t[i] = m[4 * i + 3];
	}
	const double c00 = r[1][1] * r[2][2] - r[1][2] * r[2][1];
	const double c01 = r[1][2] * r[2][0] - r[1][0] * r[2][2];
	const double c02 = r[1][0] * r[2][1] - r[1][1] * r[2][0];
	const double det = (r[0][0] * c00 + r[0][1] * c01) + r[0][2] * c02;
	if (!(fabs(det) > 0.0) || !isfinite(det)) return 0;
	double a[3][3];
	a[0][0] = c00 / det;
	a[1][0] = c01 / det;
	a[2][0] = c02 / det;
	a[0][1] = (r[0][2] * r[2][1] - r[0][1] * r[2][2]) / det;
	a[1][1] = (r[0][0] * r[2][2] - r[0][2] * r[2][0]) / det;
	a[2][1] = (r[0][1] * r[2][0] - r[0][0] * r[2][1]) / det;
	a[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) / det;
	a[1][2] = (r[0][2] * r[1][0] - r[0][0] * r[1][2]) / det;
	a[2][2] = (r[0][0] * r[1][1] - r[0][1] * r[1][0]) / det;
	for (int i = 0; i < 3; i++) {
		const double b = -((a[i][0] * t[0] + a[i][1] * t[1]) + a[i][2] * t[2]);
		for (int j = 0; j < 3; j++) out[4 * i + j] = (float)a[i][j];
		out[4 * i + 3] = (float)b;
	}
	for (int k = 0; k < 12; k++)
		if (!isfinite(out[k])) return 0;
	return 1;
}

/* The mesh root box (union of node 0's child boxes) mapped by (R | t): the 8 corners in double,
 * bounds widened by 1e-5 of the largest extent + 1e-6 of the largest magnitude, rounded outwards
 * to float. */
static int world_box(const float m[12], const OrNode* root, float wmn[3], float wmx[3])
{
	float mn[3], mx[3];
	for (int k = 0; k < 3; k++) {
		mn[k] = min_first(root->f[k], root->f[6 + k]);
		mx[k] = max_first(root->f[3 + k], root->f[9 + k]);
	}
	double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
	for (int c = 0; c < 8; c++) {
		const double p[3] = {(c & 1) ? mx[0] : mn[0], (c & 2) ? mx[1] : mn[1],
		                     (c & 4) ? mx[2] : mn[2]};
		for (int i = 0; i < 3; i++) {
			const double w = (((double)m[4 * i] * p[0] + (double)m[4 * i + 1] * p[1]) +
			                  (double)m[4 * i + 2] * p[2]) + (double)m[4 * i + 3];
			lo[i] = (w < lo[i]) ? w : lo[i];
			hi[i] = (hi[i] < w) ? w : hi[i];
		}
	}
	double ext = 0.0, mag = 0.0;
	for (int i = 0; i < 3; i++) {
		const double e = hi[i] - lo[i];
		ext = (ext < e) ? e : ext;
		const double a = fabs(lo[i]), b = fabs(hi[i]);
		const double g = (a < b) ? b : a;
		mag = (mag < g) ? g : mag;
	}
	const double pad = 1e-5 * ext + 1e-6 * mag;
	for (int i = 0; i < 3; i++) {
		const float a = (float)(lo[i] - pad), b = (float)(hi[i] + pad);
		wmn[i] = (double)a > lo[i] - pad ? nextafterf(a, -HUGE_VALF) : a;
		wmx[i] = (double)b < hi[i] + pad ? nextafterf(b, HUGE_VALF) : b;
		if (!isfinite(wmn[i]) || !isfinite(wmx[i])) return 0;
	}
	return 1;
}

int oracle_instance_prepare(const OrInstance* inst, float worldToObject[12], float wmn[3],
                            float wmx[3])
{
	return invert_affine(inst->objectToWorld, worldToObject) &&
	       world_box(inst->objectToWorld, &inst->nodes[0], wmn, wmx);
}

void oracle_cast_instances(const OrNode* sceneNodes, const OrTri* sceneTris,
                           const OrInstance* inst, uint32_t numInst, const OrRay* rays, uint32_t n,
                           OrHit* hits, uint32_t* ids, uint8_t* inLight)
{
	float* a = (float*)malloc(sizeof(float) * 12 * (numInst ? numInst : 1));
	float* box = (float*)malloc(sizeof(float) * 6 * (numInst ? numInst : 1));
	for (uint32_t i = 0; i < numInst; i++)
		oracle_instance_prepare(&inst[i], &a[12 * i], &box[6 * i], &box[6 * i + 3]);
	for (uint32_t r = 0; r < n; r++) {
		OrHit best = {~0u, FLT_MAX, FLT_MAX, FLT_MAX};
		uint32_t bestId = ~0u;
		int occluded = 0;
		if (sceneNodes) {
			if (inLight)
				occluded = anyhit_one(sceneNodes, sceneTris, &rays[r]);
			else
				cast_one(sceneNodes, sceneTris, &rays[r], &best, 0);
		}
		Ray w;
		make_ray(&rays[r], &w);
		for (uint32_t i = 0; i < numInst && !occluded; i++) {
			float tIn, tOut;
			const int hit = aabb(&w, &box[6 * i], &box[6 * i + 3], &tIn, &tOut);
			const float lim = inLight ? w.tMax : min_first(w.tMax, best.t);
			if (!(hit && tOut > w.tMin && (inLight ? tIn < lim : tIn <= lim))) continue;
			const float* A = &a[12 * i];
			OrRay o;
			for (int k = 0; k < 3; k++) {
				float p = A[4 * k] * w.o[0] + A[4 * k + 1] * w.o[1];
				p = p + A[4 * k + 2] * w.o[2];
				o.o[k] = p + A[4 * k + 3];
				float q = A[4 * k] * w.d[0] + A[4 * k + 1] * w.d[1];
				o.d[k] = q + A[4 * k + 2] * w.d[2];
			}
			o.tMin = w.tMin;
			o.tMax = w.tMax;
			if (inLight) {
				occluded = anyhit_one(inst[i].nodes, inst[i].tris, &o);
			} else {
				OrHit c;
				cast_one(inst[i].nodes, inst[i].tris, &o, &c, 0);
				if (c.tri != ~0u && c.t < best.t) {
					best = c;
					bestId = i;
				}
			}
		}
		if (inLight) {
			inLight[r] = (uint8_t)!occluded;
		} else {
			hits[r] = best;
			if (ids) ids[r] = bestId;
		}
	}
	free(a);
	free(box);
}
