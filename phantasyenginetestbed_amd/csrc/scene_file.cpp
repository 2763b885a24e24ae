// Versioned scene files: the reference's .bvhcache (three raw dumps of the BVH arrays,
// testbed_application/Main.cpp:105-154) made explicit as one file with a header, a checksum and
// structural validation, plus an importer for the reference's three raw files.
//
// Layout (little-endian):
//   0   char[8]  magic "PHXSCENE"
//   8   u32      version (1)
//   12  u32      header bytes (64)
//   16  u32      node count        20  u32 triangle count
//   24  u32      material count    28  u32 light count
//   32  u32      flags (bit 0: TriangleData present)
//   36  u32[3]   zero
//   48  u64      FNV-1a 64 of every byte after the header
//   56  u64      zero
//   64  sections, each starting at a 64-byte boundary: nodes (64 B each), triangles (48 B),
//       datas (64 B, if present), materials (48 B), lights (36 B)
#include <cstdio>
#include <cstring>
#include <new>

#include "runtime.hpp"

using namespace phx;

struct PhxSceneFile {
	std::vector<PhxBVHNode> nodes;
	std::vector<PhxTriangleVertices> tris;
	std::vector<PhxTriangleData> datas;
	std::vector<PhxMaterial> materials;
	std::vector<PhxSphereLight> lights;
};

namespace {

constexpr char kMagic[8] = {'P', 'H', 'X', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kHeaderBytes = 64;

struct Header {
	char magic[8];
	uint32_t version;
	uint32_t headerBytes;
	uint32_t numNodes, numTris, numMaterials, numLights;
	uint32_t flags;
	uint32_t zero[3];
	uint64_t checksum;
	uint64_t zero2;
};
static_assert(sizeof(Header) == kHeaderBytes, "scene file header");

uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
	const unsigned char* b = static_cast<const unsigned char*>(p);
	for (size_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 0x100000001B3ull;
	}
	return h;
}
constexpr uint64_t kFnvBasis = 0xCBF29CE484222325ull;

size_t pad64(size_t n) { return (64 - n % 64) % 64; }

struct Section {
	const void* data;
	size_t bytes;
};

template <class T>
phx_status readSection(FILE* f, std::vector<T>& out, uint32_t count, uint64_t& h, const char* path)
{
	out.resize(count);
	const size_t bytes = size_t(count) * sizeof(T);
	if (bytes && std::fread(out.data(), 1, bytes, f) != bytes)
		return fail(PHX_ERR_INVALID_SCENE, "%s: truncated file", path);
	h = fnv1a(h, out.data(), bytes);
	unsigned char pad[64];
	const size_t p = pad64(bytes);
	if (p && std::fread(pad, 1, p, f) != p) return fail(PHX_ERR_INVALID_SCENE, "%s: truncated file", path);
	// Padding is written as zeros. Checking it keeps a count that is one record short, whose last
	// record then reads as padding under the same checksum, from loading.
	for (size_t i = 0; i < p; i++)
		if (pad[i] != 0) return fail(PHX_ERR_INVALID_SCENE, "%s: nonzero section padding", path);
	h = fnv1a(h, pad, p);
	return PHX_OK;
}

template <class T>
phx_status readRaw(const char* path, std::vector<T>& out)
{
	FILE* f = std::fopen(path, "rb");
	if (!f) return fail(PHX_ERR_INVALID_ARGUMENT, "cannot open %s", path);
	std::fseek(f, 0, SEEK_END);
	const long size = std::ftell(f);
	std::fseek(f, 0, SEEK_SET);
	if (size < 0 || size_t(size) % sizeof(T) != 0) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: %ld bytes is not a whole number of %zu-byte records",
		            path, size, sizeof(T));
	}
	out.resize(size_t(size) / sizeof(T));
	const bool okRead = out.empty() || std::fread(out.data(), 1, size_t(size), f) == size_t(size);
	std::fclose(f);
	if (!okRead) return fail(PHX_ERR_INVALID_SCENE, "%s: read failed", path);
	return PHX_OK;
}

phx_status validateArrays(const PhxSceneFile& s, const char* what)
{
	if (s.nodes.empty() || s.tris.empty())
		return fail(PHX_ERR_INVALID_SCENE, "%s: no nodes or no triangles", what);
	if (!s.datas.empty() && s.datas.size() != s.tris.size())
		return fail(PHX_ERR_INVALID_SCENE, "%s: %zu triangle datas for %zu triangles", what,
		            s.datas.size(), s.tris.size());
	uint32_t depth = 0;
	phx_status st = validateBVH(s.nodes.data(), uint32_t(s.nodes.size()), s.tris.data(),
	                            uint32_t(s.tris.size()), &depth);
	if (st != PHX_OK) return st;
	if (!s.datas.empty() && !s.materials.empty())
		for (size_t i = 0; i < s.datas.size(); i++)
			if (s.datas[i].materialIndex >= s.materials.size())
				return fail(PHX_ERR_INVALID_SCENE, "%s: triangle %zu uses material %u of %zu", what,
				            i, s.datas[i].materialIndex, s.materials.size());
	return PHX_OK;
}

}  // namespace

extern "C" {

phx_status phx_scene_file_write(const char* path, const PhxSceneArrays* a)
{
	if (!path || !a) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_file_write: NULL argument");
	if (!a->nodes || !a->triangles || !a->numNodes || !a->numTriangles ||
	    (a->numMaterials && !a->materials) || (a->numLights && !a->lights))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_file_write: missing arrays");
	uint32_t depth = 0;
	phx_status st = validateBVH(a->nodes, a->numNodes, a->triangles, a->numTriangles, &depth);
	if (st != PHX_OK) return st;
	Header h;
	std::memset(&h, 0, sizeof(h));
	std::memcpy(h.magic, kMagic, 8);
	h.version = kVersion;
	h.headerBytes = kHeaderBytes;
	h.numNodes = a->numNodes;
	h.numTris = a->numTriangles;
	h.numMaterials = a->numMaterials;
	h.numLights = a->numLights;
	h.flags = a->datas ? 1u : 0u;
	const Section secs[5] = {
		{a->nodes, size_t(a->numNodes) * sizeof(PhxBVHNode)},
		{a->triangles, size_t(a->numTriangles) * sizeof(PhxTriangleVertices)},
		{a->datas, a->datas ? size_t(a->numTriangles) * sizeof(PhxTriangleData) : 0},
		{a->materials, size_t(a->numMaterials) * sizeof(PhxMaterial)},
		{a->lights, size_t(a->numLights) * sizeof(PhxSphereLight)}};
	const unsigned char zeros[64] = {0};
	uint64_t sum = kFnvBasis;
	for (const Section& s : secs) {
		sum = fnv1a(sum, s.data, s.bytes);
		sum = fnv1a(sum, zeros, pad64(s.bytes));
	}
	h.checksum = sum;
	FILE* f = std::fopen(path, "wb");
	if (!f) return fail(PHX_ERR_INVALID_ARGUMENT, "cannot create %s", path);
	bool good = std::fwrite(&h, 1, sizeof(h), f) == sizeof(h);
	for (const Section& s : secs) {
		if (good && s.bytes) good = std::fwrite(s.data, 1, s.bytes, f) == s.bytes;
		const size_t p = pad64(s.bytes);
		if (good && p) good = std::fwrite(zeros, 1, p, f) == p;
	}
	good = (std::fclose(f) == 0) && good;
	if (!good) return fail(PHX_ERR_INVALID_ARGUMENT, "%s: write failed", path);
	return ok();
}

phx_status phx_scene_file_read(const char* path, PhxSceneFile** out)
{
	if (!path || !out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_file_read: NULL argument");
	*out = nullptr;
	FILE* f = std::fopen(path, "rb");
	if (!f) return fail(PHX_ERR_INVALID_ARGUMENT, "cannot open %s", path);
	Header h;
	if (std::fread(&h, 1, sizeof(h), f) != sizeof(h)) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: shorter than the header", path);
	}
	if (std::memcmp(h.magic, kMagic, 8) != 0) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: not a phx scene file", path);
	}
	if (h.version != kVersion || h.headerBytes != kHeaderBytes) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: version %u (header %u bytes), this library reads "
		            "version %u", path, h.version, h.headerBytes, kVersion);
	}
	const uint64_t total = uint64_t(h.numNodes) * 64u + uint64_t(h.numTris) * (48u + 64u) +
	                       uint64_t(h.numMaterials) * 48u + uint64_t(h.numLights) * 36u;
	if (h.numNodes > (1u << 28) || h.numTris > (1u << 28) || h.numMaterials > (1u << 24) ||
	    h.numLights > (1u << 24) || total > (uint64_t(1) << 40)) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: implausible counts", path);
	}
	// The counts fix the file's size; checking it first keeps a damaged or truncated header from
	// sizing the section buffers.
	const uint64_t sectionBytes[5] = {
		uint64_t(h.numNodes) * sizeof(PhxBVHNode), uint64_t(h.numTris) * sizeof(PhxTriangleVertices),
		(h.flags & 1u) ? uint64_t(h.numTris) * sizeof(PhxTriangleData) : 0u,
		uint64_t(h.numMaterials) * sizeof(PhxMaterial), uint64_t(h.numLights) * sizeof(PhxSphereLight)};
	uint64_t expected = kHeaderBytes;
	for (uint64_t b : sectionBytes) expected += b + pad64(size_t(b));
	long fileBytes = -1;
	if (std::fseek(f, 0, SEEK_END) == 0) fileBytes = std::ftell(f);
	if (fileBytes < 0 || uint64_t(fileBytes) != expected ||
	    std::fseek(f, long(kHeaderBytes), SEEK_SET) != 0) {
		std::fclose(f);
		return fail(PHX_ERR_INVALID_SCENE, "%s: %s (%ld bytes, the header's counts give %llu)", path,
		            fileBytes >= 0 && uint64_t(fileBytes) > expected
		                ? "trailing bytes after the last section" : "truncated file",
		            fileBytes, (unsigned long long)expected);
	}
	PhxSceneFile* s = new (std::nothrow) PhxSceneFile();
	if (!s) {
		std::fclose(f);
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_scene_file_read: out of host memory");
	}
	uint64_t sum = kFnvBasis;
	phx_status st = PHX_OK;
	try {
		st = readSection(f, s->nodes, h.numNodes, sum, path);
		if (st == PHX_OK) st = readSection(f, s->tris, h.numTris, sum, path);
		if (st == PHX_OK) st = readSection(f, s->datas, (h.flags & 1u) ? h.numTris : 0u, sum, path);
		if (st == PHX_OK) st = readSection(f, s->materials, h.numMaterials, sum, path);
		if (st == PHX_OK) st = readSection(f, s->lights, h.numLights, sum, path);
	} catch (const std::bad_alloc&) {
		st = fail(PHX_ERR_OUT_OF_MEMORY, "phx_scene_file_read: out of host memory");
	}
	unsigned char extra;
	if (st == PHX_OK && std::fread(&extra, 1, 1, f) == 1)
		st = fail(PHX_ERR_INVALID_SCENE, "%s: trailing bytes after the last section", path);
	std::fclose(f);
	if (st == PHX_OK && sum != h.checksum)
		st = fail(PHX_ERR_INVALID_SCENE, "%s: checksum mismatch (file damaged)", path);
	if (st == PHX_OK) st = validateArrays(*s, path);
	if (st != PHX_OK) {
		delete s;
		return st;
	}
	registry().add(s);
	*out = s;
	return ok();
}

phx_status phx_scene_file_read_bvhcache(const char* nodesPath, const char* vertsPath,
                                        const char* datasPath, PhxSceneFile** out)
{
	if (!nodesPath || !vertsPath || !out)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_file_read_bvhcache: NULL argument");
	*out = nullptr;
	PhxSceneFile* s = new (std::nothrow) PhxSceneFile();
	if (!s) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_scene_file_read_bvhcache: out of host memory");
	phx_status st = readRaw(nodesPath, s->nodes);
	if (st == PHX_OK) st = readRaw(vertsPath, s->tris);
	if (st == PHX_OK && datasPath) st = readRaw(datasPath, s->datas);
	if (st == PHX_OK) st = validateArrays(*s, nodesPath);
	if (st != PHX_OK) {
		delete s;
		return st;
	}
	registry().add(s);
	*out = s;
	return ok();
}

phx_status phx_scene_file_arrays(const PhxSceneFile* s, PhxSceneArrays* a)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid scene file handle");
	if (!a) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_scene_file_arrays: NULL output");
	a->nodes = s->nodes.data();
	a->numNodes = uint32_t(s->nodes.size());
	a->triangles = s->tris.data();
	a->datas = s->datas.empty() ? nullptr : s->datas.data();
	a->numTriangles = uint32_t(s->tris.size());
	a->materials = s->materials.empty() ? nullptr : s->materials.data();
	a->numMaterials = uint32_t(s->materials.size());
	a->lights = s->lights.empty() ? nullptr : s->lights.data();
	a->numLights = uint32_t(s->lights.size());
	return ok();
}

void phx_scene_file_release(PhxSceneFile* s)
{
	if (s && registry().remove(s)) delete s;
}

}  // extern "C"
