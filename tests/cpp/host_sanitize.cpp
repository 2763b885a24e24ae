// Host-code check run under the compilers' sanitizers (address + undefined behaviour, or thread):
// the SBVH builder at several sizes and thread counts (output identical for every count), the
// scene generators, and the scene-file reader and raw importer fed truncated, re-checksummed
// corrupted and inconsistent files, which must be refused with an error status and no invalid
// memory access. Links the library's host-only sources directly; no GPU and no HIP runtime.
//
// usage: host_sanitize full | threads     (prints "host_sanitize ok" and returns 0 on success)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../phantasyenginetestbed_amd/csrc/phx_internal.hpp"

static int failures = 0;

#define CHECK(cond, ...)                                           \
	do {                                                           \
		if (!(cond)) {                                             \
			std::printf("FAILED %s:%d: ", __FILE__, __LINE__);     \
			std::printf(__VA_ARGS__);                              \
			std::printf("\n");                                     \
			failures++;                                            \
		}                                                          \
	} while (0)

static uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
	const unsigned char* b = static_cast<const unsigned char*>(p);
	for (size_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 0x100000001B3ull;
	}
	return h;
}
constexpr uint64_t kBasis = 0xCBF29CE484222325ull;

// xorshift64*: deterministic corruption positions.
struct Rng {
	uint64_t s;
	uint64_t next()
	{
		s ^= s >> 12;
		s ^= s << 25;
		s ^= s >> 27;
		return s * 0x2545F4914F6CDD1Dull;
	}
	uint32_t below(uint32_t n) { return uint32_t(next() % n); }
};

struct Scene {
	std::vector<PhxTriangleVertices> tris;
	std::vector<PhxTriangleData> datas;
	phx::BuiltBVH bvh;
};

static uint64_t digest(const phx::BuiltBVH& b)
{
	uint64_t h = fnv1a(kBasis, b.nodes.data(), b.nodes.size() * sizeof(PhxBVHNode));
	h = fnv1a(h, b.triangles.data(), b.triangles.size() * sizeof(PhxTriangleVertices));
	return fnv1a(h, b.datas.data(), b.datas.size() * sizeof(PhxTriangleData));
}

static bool buildHeightfield(uint32_t grid, uint32_t boxes, uint32_t threads, Scene& s)
{
	const uint32_t n = phx_scene_heightfield_boxes_count(grid, boxes);
	s.tris.resize(n);
	s.datas.resize(n);
	phx_status st = phx_scene_heightfield_boxes(grid, boxes, 0xC0FFEE02ull, 0.1f, -3.0f, 0.5f, 4,
	                                            s.tris.data(), s.datas.data());
	CHECK(st == PHX_OK, "heightfield %u/%u: status %d", grid, boxes, int(st));
	if (st != PHX_OK) return false;
	st = phx::buildSBVH(s.tris.data(), s.datas.data(), n, threads, s.bvh);
	CHECK(st == PHX_OK, "build %u/%u threads %u: status %d", grid, boxes, threads, int(st));
	if (st != PHX_OK) return false;
	uint32_t depth = 0;
	st = phx::validateBVH(s.bvh.nodes.data(), uint32_t(s.bvh.nodes.size()), s.bvh.triangles.data(),
	                      uint32_t(s.bvh.triangles.size()), &depth);
	CHECK(st == PHX_OK && depth >= 1 && depth <= 64, "validate %u/%u: status %d depth %u", grid,
	      boxes, int(st), depth);
	return st == PHX_OK;
}

static void builderThreads(const uint32_t (*cases)[2], size_t numCases, uint32_t maxThreads)
{
	for (size_t c = 0; c < numCases; c++) {
		Scene one, many;
		if (!buildHeightfield(cases[c][0], cases[c][1], 1, one)) continue;
		if (!buildHeightfield(cases[c][0], cases[c][1], maxThreads, many)) continue;
		CHECK(digest(one.bvh) == digest(many.bvh), "grid %u: 1 and %u threads differ", cases[c][0],
		      maxThreads);
	}
}

static std::vector<unsigned char> readFile(const std::string& p)
{
	std::vector<unsigned char> v;
	FILE* f = std::fopen(p.c_str(), "rb");
	if (!f) return v;
	unsigned char buf[65536];
	size_t k;
	while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) v.insert(v.end(), buf, buf + k);
	std::fclose(f);
	return v;
}

static void writeFile(const std::string& p, const unsigned char* d, size_t n)
{
	FILE* f = std::fopen(p.c_str(), "wb");
	if (!f) return;
	if (n) std::fwrite(d, 1, n, f);
	std::fclose(f);
}

// Reads `bytes` as a scene file: returns the status; a file that loads is re-validated.
static phx_status load(const std::string& path, const std::vector<unsigned char>& bytes)
{
	writeFile(path, bytes.data(), bytes.size());
	PhxSceneFile* s = nullptr;
	phx_status st = phx_scene_file_read(path.c_str(), &s);
	if (st == PHX_OK) {
		PhxSceneArrays a{};
		CHECK(phx_scene_file_arrays(s, &a) == PHX_OK, "arrays of a loaded file");
		uint32_t depth = 0;
		CHECK(phx::validateBVH(a.nodes, a.numNodes, a.triangles, a.numTriangles, &depth) == PHX_OK,
		      "a loaded file does not validate");
		phx_scene_file_release(s);
	} else {
		CHECK(s == nullptr, "failed read left a handle");
	}
	return st;
}

static void rechecksum(std::vector<unsigned char>& b)
{
	const uint64_t h = fnv1a(kBasis, b.data() + 64, b.size() - 64);
	std::memcpy(b.data() + 48, &h, 8);
}

static void sceneFiles(const std::string& dir)
{
	Scene s;
	if (!buildHeightfield(48, 12, 2, s)) return;
	std::vector<PhxMaterial> mats(4);
	std::vector<PhxSphereLight> lights(3);
	CHECK(phx_scene_materials_random(4, 5, mats.data()) == PHX_OK, "materials");
	CHECK(phx_scene_lights_random(3, 6, lights.data()) == PHX_OK, "lights");
	PhxSceneArrays a{};
	a.nodes = s.bvh.nodes.data();
	a.numNodes = uint32_t(s.bvh.nodes.size());
	a.triangles = s.bvh.triangles.data();
	a.numTriangles = uint32_t(s.bvh.triangles.size());
	a.datas = s.bvh.datas.data();
	a.materials = mats.data();
	a.numMaterials = 4;
	a.lights = lights.data();
	a.numLights = 3;
	const std::string path = dir + "/scene.phxscene", bad = dir + "/bad.phxscene";
	phx_status st = phx_scene_file_write(path.c_str(), &a);
	CHECK(st == PHX_OK, "write: status %d", int(st));
	const std::vector<unsigned char> good = readFile(path);
	CHECK(good.size() > 64, "written file is %zu bytes", good.size());
	if (good.size() <= 64) return;
	CHECK(load(bad, good) == PHX_OK, "an intact copy must load");

	// Truncations at every 64-byte boundary and a spread of odd offsets.
	for (size_t cut = 0; cut < good.size(); cut += (cut < 256 ? 7 : good.size() / 53 + 1)) {
		std::vector<unsigned char> t(good.begin(), good.begin() + long(cut));
		CHECK(load(bad, t) == PHX_ERR_INVALID_SCENE, "truncated at %zu loaded", cut);
	}
	// Each header count off by one, and far off.
	for (int field = 0; field < 4; field++)
		for (int64_t delta : {int64_t(1), int64_t(-1), int64_t(1) << 26}) {
			std::vector<unsigned char> t = good;
			uint32_t v;
			std::memcpy(&v, t.data() + 16 + 4 * field, 4);
			v = uint32_t(int64_t(v) + delta);
			std::memcpy(t.data() + 16 + 4 * field, &v, 4);
			CHECK(load(bad, t) != PHX_OK, "count %d changed by %lld loaded", field, (long long)delta);
		}
	// Random payload damage with the checksum made to match: the structural checks must refuse
	// or accept it without reading outside the arrays; a file that loads validates.
	Rng rng{0x9E3779B97F4A7C15ull};
	const size_t nodeBytes = s.bvh.nodes.size() * sizeof(PhxBVHNode);
	int refusedCount = 0;
	for (int trial = 0; trial < 300; trial++) {
		std::vector<unsigned char> t = good;
		const int flips = 1 + int(rng.below(4));
		for (int k = 0; k < flips; k++) {
			// Mostly the index words of nodes (bytes 48-63 of a node), where damage matters.
			size_t at;
			if (rng.below(4) != 0)
				at = 64 + size_t(rng.below(uint32_t(s.bvh.nodes.size()))) * 64 + 48 + rng.below(16);
			else
				at = 64 + rng.below(uint32_t(t.size() - 64));
			t[at] ^= uint8_t(1u << rng.below(8));
		}
		rechecksum(t);
		const phx_status r = load(bad, t);
		CHECK(r == PHX_OK || r == PHX_ERR_INVALID_SCENE || r == PHX_ERR_MALFORMED_BVH,
		      "damaged file: status %d", int(r));
		refusedCount += r != PHX_OK;
	}
	CHECK(refusedCount > 100, "only %d of 300 damaged files refused", refusedCount);

	// Raw importer: whole dumps load; a partial record and damaged child indices are refused.
	const std::string rn = dir + "/n.bvhcache", rv = dir + "/v.bvhcache";
	writeFile(rn, reinterpret_cast<const unsigned char*>(a.nodes), nodeBytes);
	writeFile(rv, reinterpret_cast<const unsigned char*>(a.triangles),
	          a.numTriangles * sizeof(PhxTriangleVertices));
	PhxSceneFile* f = nullptr;
	CHECK(phx_scene_file_read_bvhcache(rn.c_str(), rv.c_str(), nullptr, &f) == PHX_OK, "raw import");
	phx_scene_file_release(f);
	writeFile(rn, reinterpret_cast<const unsigned char*>(a.nodes), nodeBytes - 8);
	f = nullptr;
	CHECK(phx_scene_file_read_bvhcache(rn.c_str(), rv.c_str(), nullptr, &f) == PHX_ERR_INVALID_SCENE,
	      "partial node record imported");
	for (int trial = 0; trial < 100; trial++) {
		std::vector<PhxBVHNode> nodes(s.bvh.nodes);
		unsigned char* raw = reinterpret_cast<unsigned char*>(nodes.data());
		raw[size_t(rng.below(uint32_t(nodes.size()))) * 64 + 48 + rng.below(16)] ^=
		    uint8_t(1u << rng.below(8));
		writeFile(rn, raw, nodeBytes);
		f = nullptr;
		const phx_status r = phx_scene_file_read_bvhcache(rn.c_str(), rv.c_str(), nullptr, &f);
		CHECK(r == PHX_OK || r == PHX_ERR_INVALID_SCENE || r == PHX_ERR_MALFORMED_BVH,
		      "damaged raw nodes: status %d", int(r));
		if (r == PHX_OK) phx_scene_file_release(f);
	}
	// Arrays of arbitrary bytes as nodes: refused, never read outside.
	for (int trial = 0; trial < 200; trial++) {
		std::vector<PhxBVHNode> nodes(1 + rng.below(40));
		for (auto& nd : nodes) {
			uint64_t w[8];
			for (auto& x : w) x = rng.next();
			std::memcpy(&nd, w, sizeof(nd));
		}
		uint32_t depth = 0;
		const phx_status r = phx::validateBVH(nodes.data(), uint32_t(nodes.size()), a.triangles,
		                                      a.numTriangles, &depth);
		CHECK(r != PHX_OK, "random nodes validated");
	}
}

int main(int argc, char** argv)
{
	const std::string mode = argc > 1 ? argv[1] : "full";
	const std::string dir = argc > 2 ? argv[2] : "/tmp";
	if (mode == "threads") {
		static const uint32_t cases[][2] = {{40, 10}, {96, 40}};
		builderThreads(cases, 2, 6);
	} else {
		static const uint32_t cases[][2] = {{8, 2}, {64, 20}, {128, 60}};
		builderThreads(cases, 3, 8);
		std::vector<PhxTriangleVertices> t(3000);
		CHECK(phx_scene_random_triangles(3000, 7, 0.001f, 0.05f, t.data(), nullptr) == PHX_OK,
		      "random triangles");
		phx::BuiltBVH b;
		CHECK(phx::buildSBVH(t.data(), nullptr, 3000, 3, b) == PHX_OK, "random build");
		sceneFiles(dir);
	}
	if (failures) {
		std::printf("host_sanitize: %d failures\n", failures);
		return 1;
	}
	std::printf("host_sanitize ok\n");
	return 0;
}
