// Synthetic scenes of the benchmark workloads and the host-side camera helpers.
//
// The scenes are generated in double precision from seeded splitmix64 streams (Steele, Lea and
// Flood 2014) and rounded to float once, so the same seed gives the same triangles on any host
// with the same libm.
#include <cmath>
#include <cstring>

#include "phx_internal.hpp"

namespace phx {
namespace {

struct SplitMix64 {
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) {}
	uint64_t next()
	{
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
	double normal()  // Box-Muller
	{
		double u1 = uniform();
		double u2 = uniform();
		if (u1 < 1e-300) u1 = 1e-300;
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
	}
};

uint64_t mix64(uint64_t a, uint64_t b)
{
	SplitMix64 m(a ^ (b * 0xD6E8FEB86659FD93ull));
	return m.next();
}

struct D3 {
	double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 cross(D3 a, D3 b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline D3 normalized(D3 a)
{
	double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
	if (l == 0.0) return {0.0, 1.0, 0.0};
	return {a.x / l, a.y / l, a.z / l};
}

void putTri(PhxTriangleVertices& t, D3 a, D3 b, D3 c)
{
	t.v0[0] = float(a.x), t.v0[1] = float(a.y), t.v0[2] = float(a.z), t.v0[3] = 0.0f;
	t.v1[0] = float(b.x), t.v1[1] = float(b.y), t.v1[2] = float(b.z), t.v1[3] = 0.0f;
	t.v2[0] = float(c.x), t.v2[1] = float(c.y), t.v2[2] = float(c.z), t.v2[3] = 0.0f;
}

void putData(PhxTriangleData& d, D3 n, const float uv[6], uint32_t material)
{
	float nf[3] = {float(n.x), float(n.y), float(n.z)};
	std::memcpy(d.n0, nf, sizeof(nf));
	std::memcpy(d.n1, nf, sizeof(nf));
	std::memcpy(d.n2, nf, sizeof(nf));
	d.uv0[0] = uv[0], d.uv0[1] = uv[1];
	d.uv1[0] = uv[2], d.uv1[1] = uv[3];
	d.uv2[0] = uv[4], d.uv2[1] = uv[5];
	d.materialIndex = material;
}

// Value noise on the integer lattice with smoothstep interpolation.
double latticeValue(int64_t ix, int64_t iz, uint32_t octave, uint64_t seed)
{
	uint64_t h = mix64(seed, uint64_t(ix) * 0x9E3779B1ull + uint64_t(iz) * 0x85EBCA77ull +
	                             uint64_t(octave) * 0xC2B2AE3Dull);
	return double(h >> 11) * (1.0 / 9007199254740992.0);
}

double valueNoise(double x, double z, uint32_t octave, uint64_t seed)
{
	double fx = std::floor(x), fz = std::floor(z);
	int64_t ix = int64_t(fx), iz = int64_t(fz);
	double tx = x - fx, tz = z - fz;
	double sx = tx * tx * (3.0 - 2.0 * tx);
	double sz = tz * tz * (3.0 - 2.0 * tz);
	double v00 = latticeValue(ix, iz, octave, seed);
	double v10 = latticeValue(ix + 1, iz, octave, seed);
	double v01 = latticeValue(ix, iz + 1, octave, seed);
	double v11 = latticeValue(ix + 1, iz + 1, octave, seed);
	double a = v00 + (v10 - v00) * sx;
	double b = v01 + (v11 - v01) * sx;
	return a + (b - a) * sz;
}

// 4 octaves, lacunarity 2, gain 0.5, base frequency 4 lattice cells per unit; range about [-1.9, 1.9].
double fbm(double x, double z, uint64_t seed)
{
	double sum = 0.0, amp = 1.0, freq = 4.0;
	for (uint32_t o = 0; o < 4; o++) {
		sum += amp * (2.0 * valueNoise(x * freq, z * freq, o, seed) - 1.0);
		amp *= 0.5;
		freq *= 2.0;
	}
	return sum;
}

}  // namespace
}  // namespace phx

using namespace phx;

extern "C" phx_status phx_scene_random_triangles(uint32_t n, uint64_t seed, float eMin, float eMax,
                                                 PhxTriangleVertices* triangles,
                                                 PhxTriangleData* datas)
{
	if (!triangles || n == 0)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_random_triangles: bad arguments");
	if (!(eMin > 0.0f) || !(eMax >= eMin))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_random_triangles: bad edge range");
	SplitMix64 rng(seed);
	const double PI2 = 6.283185307179586;
	for (uint32_t i = 0; i < n; i++) {
		D3 c = {rng.uniform(), rng.uniform(), rng.uniform()};
		double a, b, e;
		do {
			a = eMin + (eMax - eMin) * rng.uniform();
			b = eMin + (eMax - eMin) * rng.uniform();
			e = eMin + (eMax - eMin) * rng.uniform();
		} while (!(a + b > e && b + e > a && a + e > b));
		// a = |BC|, b = |CA|, e = |AB|; A at the origin, B on +x, C by the law of cosines.
		double cosA = (b * b + e * e - a * a) / (2.0 * b * e);
		double sinA = std::sqrt(std::fmax(0.0, 1.0 - cosA * cosA));
		D3 P[3] = {{0.0, 0.0, 0.0}, {e, 0.0, 0.0}, {b * cosA, b * sinA, 0.0}};
		D3 g = {(P[0].x + P[1].x + P[2].x) / 3.0, (P[0].y + P[1].y + P[2].y) / 3.0, 0.0};
		// Uniform random rotation (Shoemake 1992): q = (x, y, z, w).
		double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
		double r1 = std::sqrt(1.0 - u1), r2 = std::sqrt(u1);
		double qx = r1 * std::sin(PI2 * u2), qy = r1 * std::cos(PI2 * u2);
		double qz = r2 * std::sin(PI2 * u3), qw = r2 * std::cos(PI2 * u3);
		double R[3][3] = {
			{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
			{2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
			{2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
		D3 W[3];
		for (int k = 0; k < 3; k++) {
			D3 p = sub(P[k], g);
			W[k] = {c.x + R[0][0] * p.x + R[0][1] * p.y + R[0][2] * p.z,
			        c.y + R[1][0] * p.x + R[1][1] * p.y + R[1][2] * p.z,
			        c.z + R[2][0] * p.x + R[2][1] * p.y + R[2][2] * p.z};
		}
		putTri(triangles[i], W[0], W[1], W[2]);
		if (datas) {
			const float uv[6] = {0, 0, 0, 0, 0, 0};
			putData(datas[i], normalized(cross(sub(W[1], W[0]), sub(W[2], W[0]))), uv, 0);
		}
	}
	return ok();
}

extern "C" uint32_t phx_scene_heightfield_boxes_count(uint32_t grid, uint32_t numBoxes)
{
	return 2u * grid * grid + 12u * numBoxes;
}

extern "C" phx_status phx_scene_heightfield_boxes(uint32_t grid, uint32_t numBoxes, uint64_t seed,
                                                  float amp, float boxMu, float boxSigma,
                                                  uint32_t numMaterials,
                                                  PhxTriangleVertices* triangles,
                                                  PhxTriangleData* datas)
{
	if (!triangles || grid == 0 || grid > 16384)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_heightfield_boxes: bad arguments");
	const uint64_t noiseSeed = mix64(seed, 1);
	SplitMix64 boxRng(mix64(seed, 2));
	SplitMix64 matRng(mix64(seed, 3));
	auto height = [&](double x, double z) { return double(amp) * fbm(x, z, noiseSeed); };

	const uint32_t V = grid + 1;
	std::vector<D3> verts(size_t(V) * V);
	for (uint32_t j = 0; j < V; j++)
		for (uint32_t i = 0; i < V; i++) {
			double x = double(i) / grid, z = double(j) / grid;
			verts[size_t(j) * V + i] = {x, height(x, z), z};
		}

	auto material = [&]() -> uint32_t {
		if (numMaterials == 0) return 0;
		uint32_t m = uint32_t(matRng.uniform() * numMaterials);
		return m < numMaterials ? m : numMaterials - 1;
	};

	uint32_t t = 0;
	for (uint32_t j = 0; j < grid; j++)
		for (uint32_t i = 0; i < grid; i++) {
			D3 p00 = verts[size_t(j) * V + i], p10 = verts[size_t(j) * V + i + 1];
			D3 p01 = verts[size_t(j + 1) * V + i], p11 = verts[size_t(j + 1) * V + i + 1];
			D3 tri[2][3] = {{p00, p10, p11}, {p00, p11, p01}};
			for (int h = 0; h < 2; h++, t++) {
				putTri(triangles[t], tri[h][0], tri[h][1], tri[h][2]);
				if (datas) {
					D3 n = normalized(cross(sub(tri[h][1], tri[h][0]), sub(tri[h][2], tri[h][0])));
					if (n.y < 0) n = {-n.x, -n.y, -n.z};
					float uv[6] = {float(tri[h][0].x), float(tri[h][0].z), float(tri[h][1].x),
					               float(tri[h][1].z), float(tri[h][2].x), float(tri[h][2].z)};
					putData(datas[t], n, uv, material());
				}
			}
		}

	for (uint32_t b = 0; b < numBoxes; b++) {
		double cx = boxRng.uniform(), cz = boxRng.uniform();
		double sx = std::exp(double(boxMu) + double(boxSigma) * boxRng.normal());
		double sy = std::exp(double(boxMu) + double(boxSigma) * boxRng.normal());
		double sz = std::exp(double(boxMu) + double(boxSigma) * boxRng.normal());
		double y0 = height(cx, cz);
		double lo[3] = {cx - 0.5 * sx, y0, cz - 0.5 * sz};
		double hi[3] = {cx + 0.5 * sx, y0 + sy, cz + 0.5 * sz};
		// Corner k has x = bit0, y = bit1, z = bit2.
		auto corner = [&](int k) -> D3 {
			return {(k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 4) ? hi[2] : lo[2]};
		};
		// Faces as 4 corners in order around the face, with the outward normal.
		static const int faces[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4},
		                                {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
		static const double normals[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
		                                     {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
		for (int f = 0; f < 6; f++) {
			D3 q[4] = {corner(faces[f][0]), corner(faces[f][1]), corner(faces[f][2]),
			           corner(faces[f][3])};
			D3 n = {normals[f][0], normals[f][1], normals[f][2]};
			D3 tri[2][3] = {{q[0], q[1], q[2]}, {q[0], q[2], q[3]}};
			for (int h = 0; h < 2; h++, t++) {
				putTri(triangles[t], tri[h][0], tri[h][1], tri[h][2]);
				if (datas) {
					const float uv[6] = {0, 0, 1, 0, 1, 1};
					putData(datas[t], n, uv, material());
				}
			}
		}
	}
	return ok();
}

// generateCameraDef, ray_tracer_common/GenerateRays.cpp:12-35, in float with the sfz operation
// order (normalize = v / sqrt(dot(v, v)), dot = (x*x + y*y) + z*z, Vector.inl:300-367).
namespace {
void normalize3(float v[3])
{
	float d = v[0] * v[0];
	d = d + v[1] * v[1];
	d = d + v[2] * v[2];
	float l = std::sqrt(d);
	v[0] = v[0] / l;
	v[1] = v[1] / l;
	v[2] = v[2] / l;
}
}  // namespace

extern "C" phx_status phx_camera_def(const float pos[3], const float dir[3], const float up[3],
                                     float vertFovRad, float resX, float resY,
                                     PhxCameraDef* outCam)
{
	if (!pos || !dir || !up || !outCam || !(resX > 0) || !(resY > 0))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_camera_def: bad arguments");
	float d[3] = {dir[0], dir[1], dir[2]};
	float u[3] = {up[0], up[1], up[2]};
	normalize3(d);
	normalize3(u);
	float r[3] = {d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2], d[0] * u[1] - d[1] * u[0]};
	normalize3(r);
	float aspect = resX / resY;
	float yMax = std::tan(vertFovRad * 0.5f);
	float xMax = aspect * yMax;
	for (int k = 0; k < 3; k++) {
		outCam->origin[k] = pos[k];
		outCam->dir[k] = d[k];
		outCam->dX[k] = r[k] * xMax;
		outCam->dY[k] = u[k] * yMax;
	}
	return ok();
}

extern "C" phx_status phx_camera_look_at(const float pos[3], const float target[3],
                                         const float up[3], float vertFovRad, float resX,
                                         float resY, PhxCameraDef* outCam)
{
	if (!pos || !target || !up || !outCam)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_camera_look_at: bad arguments");
	float d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
	normalize3(d);
	float s = up[0] * d[0];
	s = s + up[1] * d[1];
	s = s + up[2] * d[2];
	float u[3] = {up[0] - d[0] * s, up[1] - d[1] * s, up[2] - d[2] * s};
	normalize3(u);
	return phx_camera_def(pos, d, u, vertFovRad, resX, resY, outCam);
}

extern "C" phx_status phx_scene_materials_random(uint32_t count, uint64_t seed, PhxMaterial* out)
{
	if (count && !out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_materials_random: NULL");
	SplitMix64 rng(mix64(seed, 4));
	for (uint32_t i = 0; i < count; i++) {
		PhxMaterial& m = out[i];
		std::memset(&m, 0, sizeof(m));
		m.iData[0] = m.iData[1] = m.iData[2] = -1;
		for (int k = 0; k < 3; k++) m.albedo[k] = float(rng.uniform());
		m.albedo[3] = 1.0f;
		m.fData2[0] = float(0.05 + 0.95 * rng.uniform());
		m.fData2[1] = 0.0f;
	}
	return ok();
}

extern "C" phx_status phx_scene_lights_random(uint32_t count, uint64_t seed, PhxSphereLight* out)
{
	if (count && !out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_lights_random: NULL");
	SplitMix64 rng(mix64(seed, 5));
	for (uint32_t i = 0; i < count; i++) {
		PhxSphereLight& l = out[i];
		std::memset(&l, 0, sizeof(l));
		l.pos[0] = float(rng.uniform());
		l.pos[1] = float(0.3 + 1.2 * rng.uniform());
		l.pos[2] = float(rng.uniform());
		l.radius = 0.0f;
		l.range = 3.0f;
		for (int k = 0; k < 3; k++) l.strength[k] = float(0.5 + 1.5 * rng.uniform());
		l.staticShadows = 1;
		l.dynamicShadows = 1;
	}
	return ok();
}
