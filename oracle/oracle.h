/*
 * oracle.h -- CPU restatement of the reference's ray-query path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only to check the HIP library's results against it (or, in the bench, as the timed CPU
 * baseline). The product (libphx) never links or calls it.
 *
 * Each function names the reference code it restates (paths relative to the reference root).
 * Parity pinning (what checks each function against something other than itself):
 *   oracle_cast_closest   - golden fixtures made by the reference's own castRay<144> compiled
 *                           from its sources (tests/golden), brute force on tiny scenes,
 *                           hand-built closed-form cases.
 *   oracle_cast_debug     - the same fixtures (counters of castDebugRay).
 *   oracle_cast_anyhit    - consistency with closest hit (occluded <=> some hit with
 *                           tMin <= t < tMax) and brute force on tiny scenes.
 *   oracle_brute_closest  - closed-form rays against hand-built triangles.
 *   oracle_primary_rays   - the reference's CPU renderer formula compiled from its sources
 *                           (golden fixture) and closed form at the image centre/corners.
 *   frame.c functions     - closed forms and distribution checks listed in frame.c's header;
 *                           the bounce stage by plane scenes with closed-form colours and
 *                           planted-mistake builds (tests/test_frame_bounce_oracle.py).
 *   oracle_pcg32          - the published pcg32 demo output (seed 42, sequence 54).
 * The structs are the byte layouts of include/phx.h, restated so the oracle stays standalone.
 */
#ifndef PHX_ORACLE_H
#define PHX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrNode {
	float f[12];
	int32_t i[4];
} OrNode;

typedef struct OrTri {
	float v[3][4];
} OrTri;

typedef struct OrRay {
	float o[3];
	float tMin;
	float d[3];
	float tMax;
} OrRay;

typedef struct OrHit {
	uint32_t tri;
	float t, u, v;
} OrHit;

/* castDebugRay's DebugRayCastData (BVHTraversal.hpp:29-35). */
typedef struct OrCounters {
	uint64_t nodesVisited;
	uint64_t trianglesIntersected;
	uint64_t triangleIntersectionTests;
	uint64_t aabbIntersectionTests;
	uint64_t aabbIntersections;
} OrCounters;

/* castRay<144> (BVHTraversal.hpp:37-128) per ray with tMin = ray.tMin, tMax = ray.tMax.
 * numThreads = 0 uses all cores; rays are handed out in chunks from an atomic counter. */
void oracle_cast_closest(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                         OrHit* hits, uint32_t numThreads);

/* castDebugRay (BVHTraversal.hpp:130-235): same hits, plus counters summed over all rays and,
 * if perRay != NULL, per ray (5 x uint32 per ray in OrCounters field order). */
void oracle_cast_debug(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                       OrHit* hits, OrCounters* total, uint32_t* perRay, uint32_t numThreads);

/* Any hit: castRay's traversal and box predicate with tCurrMax = tMax; occluded iff a triangle
 * has tMin <= t < tMax (RayCastKernel.cu:759-765). Writes inLight = !occluded. */
void oracle_cast_anyhit(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                        uint8_t* inLight, uint32_t numThreads);

/* oracle_cast_anyhit plus castDebugRay-style counters of the any-hit traversal, summed
 * (trianglesIntersected = rays occluded). */
void oracle_anyhit_debug(const OrNode* nodes, const OrTri* tris, const OrRay* rays, uint32_t n,
                         uint8_t* inLight, OrCounters* total, uint32_t numThreads);

/* Every triangle tested with Intersection.hpp:24-67 in index order; accepts t < best.t &&
 * tMin <= t <= tMax (the first index wins ties). No BVH. */
void oracle_brute_closest(const OrTri* tris, uint32_t numTris, const OrRay* rays, uint32_t n,
                          OrHit* hits);

/* Pinhole rays, CPUTracerRenderer.cpp:93-97 with CameraDef (origin, dir, dX, dY) = cam[0..11];
 * order 0 = row major, 1 = 8x8 tiles (CudaDeviceHelpers.cuh:49-62). */
void oracle_primary_rays(const float cam[12], uint32_t w, uint32_t h, float tMin, float tMax,
                         uint32_t order, OrRay* rays);

/* Intersection.hpp:24-67 on one triangle; returns hit flag. */
int oracle_triangle(const float p0[3], const float p1[3], const float p2[3], const float o[3],
                    const float d[3], float* t, float* u, float* v);

/* Intersection.hpp:77-90 on one box with invDir = 1/d; returns hit flag. */
int oracle_aabb(const float o[3], const float d[3], const float mn[3], const float mx[3],
                float* tIn, float* tOut);

/* ---- instances (two-level scenes) --------------------------------------------------------- */

/* A placed mesh: its SBVH arrays and the object -> world rows (R | t). */
typedef struct OrInstance {
	const OrNode* nodes;
	const OrTri* tris;
	float objectToWorld[12];
	const struct OrTriData* datas; /* hit records of the frame (may be NULL for casts) */
} OrInstance;

/* World -> object rows and world box of one instance (as phx_instances_build defines them). */
int oracle_instance_prepare(const OrInstance* inst, float worldToObject[12], float wmn[3],
                            float wmx[3]);
/* Closest hit (inLight NULL: hits and, if not NULL, ids written) or any hit (inLight) against a
 * static scene (sceneNodes may be NULL) and the instances, taken in index order. */
void oracle_cast_instances(const OrNode* sceneNodes, const OrTri* sceneTris,
                           const OrInstance* inst, uint32_t numInst, const OrRay* rays, uint32_t n,
                           OrHit* hits, uint32_t* ids, uint8_t* inLight);

/* ---- wavefront frame (frame.c) ---------------------------------------------------------- */

/* TriangleData (Triangle.hpp:25-37), Material (Material.hpp:17-55), SphereLight
 * (SphereLight.hpp:13-26), RayHitInfo (InterpretRayHitKernel.hpp:21-58). */
typedef struct OrTriData {
	float n0[3], n1[3], n2[3];
	float uv0[2], uv1[2], uv2[2];
	uint32_t materialIndex;
} OrTriData;

typedef struct OrMaterial {
	int32_t i[4];
	float albedo[4];
	float f2[4];
} OrMaterial;

typedef struct OrLight {
	float pos[3];
	float radius;
	float range;
	float strength[3];
	uint8_t staticShadows, dynamicShadows, pad[2];
} OrLight;

/* RawImage (rendering/RawImage.hpp:21-29): `height` rows of `pitch` bytes, `bpp` bytes a texel. */
typedef struct OrTexture {
	const uint8_t* data;
	uint32_t width, height, bpp, pitch;
} OrTexture;

typedef struct OrHitInfo {
	float d1[4]; /* position, roughness */
	float d2[4]; /* normal, metallic */
	float d3[4]; /* albedo, alpha (negative: no hit) */
} OrHitInfo;

void oracle_turn_sincos(float r, float* s, float* c);
float oracle_ggx(float nDotH, float a);
float oracle_falloff(float d, float range);
void oracle_shade(const float n[3], const float v[3], const float albedo[3], float rough,
                  float metal, const float l[3], float dist, const float strength[3], float range,
                  float out[3]);
void oracle_sample_lambert(float r1, float r2, const float n[3], float out[3]);
/* sampleGgx + clampToHemisphere (MaterialKernel.cu:220-236) around `refl` with a = roughness^2. */
void oracle_sample_ggx(float r1, float r2, float a, const float refl[3], const float n[3],
                       float out[3]);
/* The first n outputs of PCG32 seeded as pcg32_srandom_r(init, seq) (O'Neill 2014). */
void oracle_pcg32(uint64_t init, uint64_t seq, uint32_t n, uint32_t* out);
void oracle_primary_rays_jittered(const float cam[12], uint32_t w, uint32_t h, float jx, float jy,
                                  float tMin, float tMax, OrRay* rays);
void oracle_bounce_rays(const OrTriData* datas, const OrRay* rays, const OrHit* hits, uint32_t n,
                        uint64_t seed, uint32_t frame, OrRay* out);
/* Texel coordinate of a wrapping nearest-texel lookup along one axis of `dim` texels. */
uint32_t oracle_texel_coord(float c, uint32_t dim);
/* The 256 values pow(c / 255, 2.2) rounded to float. */
void oracle_linear_table(float out[256]);
void oracle_interpret(const OrRay* rays, const OrHit* hits, uint32_t n, const OrTriData* datas,
                      const OrMaterial* mats, uint32_t numMats, const OrTexture* texs,
                      uint32_t numTex, OrHitInfo* out);
void oracle_frame(const OrNode* nodes, const OrTri* tris, const OrTriData* datas,
                  const OrMaterial* mats, uint32_t numMats, const OrTexture* texs,
                  uint32_t numTex, const OrLight* lights,
                  uint32_t numLights, const float cam[12], uint32_t w, uint32_t h, float jx,
                  float jy, uint64_t seed, uint32_t frame, float* color, OrHitInfo* gbufOut,
                  OrRay* bounceOut, uint32_t threads, const OrInstance* inst, uint32_t numInst);
void oracle_velocity(const float cam[12], const float prevCam[12], const OrHitInfo* gbuf,
                     const OrRay* prim, uint32_t w, uint32_t h, float* vel);
void oracle_taa(const float* cur, const float* histIn, const float* roughness, const float* velIn,
                const float* vel, uint32_t w, uint32_t h, int reset, int clampHistory,
                float* histOut, float* out);
void oracle_halton23(uint32_t f, float* jx, float* jy);

#ifdef __cplusplus
}
#endif

#endif
