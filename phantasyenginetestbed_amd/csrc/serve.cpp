// The served form: a request loop over resident scenes with continuous batching.
//
// Requests (ray batches against resident scenes) are admitted in arrival order while their
// buffers fit the HBM budget (a first-fit arena over one device allocation, so admission never
// calls hipMalloc), and admitted requests are coalesced into the next fused persistent launch, up
// to kLaunchRays rays per launch (at least one request).
// Launches run on `slots` streams, each with its own work counter and stack spill area, so a
// launch formed from requests that arrived while another one runs starts on the next stream and
// fills the running launch's tail. A dispatcher thread forms launches and retires them (outputs
// copied back, buffers returned to the arena, waiters woken). PHX_SERVE_PLAIN serves the same
// requests one launch at a time with a synchronisation after each (the per-batch pattern of
// launchRayCastKernel / launchShadowRayCastKernel, RayCastKernel.cu:806-886).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "runtime.hpp"

using namespace phx;

namespace {

using Clock = std::chrono::steady_clock;

double msSince(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

constexpr size_t kAlign = 256;

// Rays per fused launch before further admitted requests wait for the next launch. A backlog
// coalesced into one or two launches runs its host-to-device copies, traversal and copies back
// one after the other (the 143 M-ray saturated request mix: ~340 ms); in launches of a few million
// rays the slot streams overlap one launch's copies with another's traversal (~175 ms). Each
// launch of this size still fills the device many times over (4 M rays on ~1 k blocks).
constexpr uint64_t kLaunchRays = 4u << 20;
size_t alignUp(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with coalescing of neighbouring free ranges.
class Arena {
public:
	void reset(size_t size)
	{
		free_.clear();
		size_ = size;
		used_ = 0;
		if (size) free_[0] = size;
	}
	bool alloc(size_t bytes, size_t& off)
	{
		bytes = alignUp(std::max<size_t>(bytes, 1));
		for (auto it = free_.begin(); it != free_.end(); ++it) {
			if (it->second < bytes) continue;
			off = it->first;
			const size_t rest = it->second - bytes;
			free_.erase(it);
			if (rest) free_[off + bytes] = rest;
			used_ += bytes;
			return true;
		}
		return false;
	}
	void release(size_t off, size_t bytes)
	{
		bytes = alignUp(std::max<size_t>(bytes, 1));
		used_ -= bytes;
		auto next = free_.lower_bound(off);
		if (next != free_.end() && off + bytes == next->first) {
			bytes += next->second;
			next = free_.erase(next);
		}
		if (next != free_.begin()) {
			auto prev = std::prev(next);
			if (prev->first + prev->second == off) {
				prev->second += bytes;
				return;
			}
		}
		free_[off] = bytes;
	}
	size_t used() const { return used_; }
	size_t size() const { return size_; }

private:
	std::map<size_t, size_t> free_;
	size_t size_ = 0;
	size_t used_ = 0;
};

enum class ReqState { Queued, Admitted, InFlight, Done, Cancelled };

struct Request {
	uint64_t ticket = 0;
	PhxRequest req{};
	size_t rayOff = 0, rayBytes = 0;  // arena ranges (host-side rays / outputs only)
	size_t outOff = 0, outBytes = 0;
	const PhxRayIn* devRays = nullptr;
	void* devOut = nullptr;
	ReqState state = ReqState::Queued;
	Clock::time_point submitted;
	double latencyMs = 0.0;
};

struct Slot {
	hipStream_t stream = nullptr;
	bool ownStream = false;  // slot 0 launches on the context's stream
	uint32_t* work = nullptr;
	uint32_t* spill = nullptr;
	BatchDesc* descs = nullptr;
	uint32_t* offsets = nullptr;
	BatchDesc* hostDescs = nullptr;   // pinned
	uint32_t* hostOffsets = nullptr;  // pinned
	hipEvent_t done = nullptr;
	std::vector<Request*> reqs;
	bool busy = false;
};

size_t outBytesPerRay(uint32_t kind) { return kind == PHX_BATCH_ANYHIT ? 1u : sizeof(PhxRayHit); }

}  // namespace

struct PhxServer {
	PhxContext* ctx = nullptr;
	uint32_t flags = 0;
	uint64_t budget = 0;
	char* pool = nullptr;
	Arena arena;
	uint64_t sceneBytes = 0;
	std::unordered_set<const PhxScene*> scenes;
	std::vector<Slot> slots;
	TraceLaunch shape;

	std::mutex mu;
	std::condition_variable cvWork;  // dispatcher: new submissions
	std::condition_variable cvDone;  // waiters: completions
	std::deque<Request*> queue;      // submitted, not admitted
	std::deque<Request*> admitted;   // admitted, not launched
	std::unordered_map<uint64_t, Request*> byTicket;
	uint64_t nextTicket = 1;
	bool stop = false;
	bool paused = false;  // phx_server_pause: nothing new is launched (admission goes on)
	std::string error;  // first dispatcher failure
	std::thread dispatcher;
	PhxServerStats stats{};
};

namespace {

void failServer(PhxServer* s, const char* what, hipError_t e)
{
	if (s->error.empty()) s->error = std::string(what) + ": " + hipGetErrorString(e);
}

// Admits queued requests in arrival order while their buffers fit. Caller holds the lock.
void admit(PhxServer* s)
{
	while (!s->queue.empty()) {
		Request* r = s->queue.front();
		const bool devRays = (r->req.flags & PHX_MEM_DEVICE) != 0;
		size_t ro = 0, oo = 0;
		if (!devRays) {
			if (!s->arena.alloc(r->rayBytes, ro)) return;
			if (!s->arena.alloc(r->outBytes, oo)) {
				s->arena.release(ro, r->rayBytes);
				return;
			}
			r->rayOff = ro;
			r->outOff = oo;
			r->devRays = reinterpret_cast<const PhxRayIn*>(s->pool + ro);
			r->devOut = s->pool + oo;
		} else {
			r->devRays = r->req.rays;
			r->devOut = r->req.out;
		}
		r->state = ReqState::Admitted;
		s->queue.pop_front();
		s->admitted.push_back(r);
		s->stats.peakBytesInUse = std::max<uint64_t>(s->stats.peakBytesInUse, s->arena.used());
	}
}

// Forms one launch on `slot` from the admitted requests. Caller holds the lock.
hipError_t launchSlot(PhxServer* s, Slot& slot)
{
	slot.reqs.clear();
	uint64_t total = 0;
	while (!s->admitted.empty() && slot.reqs.size() < kMaxFusedBatches) {
		Request* r = s->admitted.front();
		if (total + r->req.numRays > 0x7FFFFFFFull) break;
		if (!slot.reqs.empty() && total + r->req.numRays > kLaunchRays) break;
		total += r->req.numRays;
		slot.reqs.push_back(r);
		s->admitted.pop_front();
		if (s->flags & PHX_SERVE_PLAIN) break;  // one request per launch
	}
	if (slot.reqs.empty()) return hipSuccess;
	hipError_t e = hipSuccess;
	const uint32_t nb = uint32_t(slot.reqs.size());
	uint32_t off = 0;
	for (uint32_t b = 0; b < nb && e == hipSuccess; b++) {
		Request* r = slot.reqs[b];
		r->state = ReqState::InFlight;
		if (!(r->req.flags & PHX_MEM_DEVICE) && r->req.numRays)
			e = hipMemcpyAsync(const_cast<PhxRayIn*>(r->devRays), r->req.rays, r->rayBytes,
			                   hipMemcpyHostToDevice, slot.stream);
		const PhxScene* sc = r->req.scene;
		slot.hostDescs[b] = BatchDesc{reinterpret_cast<const float4*>(sc->nodes),
		                              reinterpret_cast<const float4*>(sc->tris),
		                              reinterpret_cast<const float4*>(r->devRays), r->devOut,
		                              r->req.numRays, r->req.kind, sc->finiteBounds, 0u};
		slot.hostOffsets[b] = off;
		off += r->req.numRays;
	}
	slot.hostOffsets[nb] = off;
	if (e != hipSuccess) return e;
	if (s->flags & PHX_SERVE_PLAIN) {
		Request* r = slot.reqs[0];
		TraceLaunch L = s->shape;
		L.work = slot.work;
		L.spill = slot.spill;
		const SceneView sv = view(r->req.scene);
		if (r->req.kind == PHX_BATCH_ANYHIT)
			e = launchAnyHit(sv, r->devRays, r->req.numRays, static_cast<uint8_t*>(r->devOut), L,
			                 slot.stream);
		else
			e = launchClosest(sv, r->devRays, r->req.numRays, static_cast<PhxRayHit*>(r->devOut),
			                  L, nullptr, slot.stream);
	} else {
		e = hipMemcpyAsync(slot.descs, slot.hostDescs, sizeof(BatchDesc) * nb,
		                   hipMemcpyHostToDevice, slot.stream);
		if (e == hipSuccess)
			e = hipMemcpyAsync(slot.offsets, slot.hostOffsets, sizeof(uint32_t) * (nb + 1),
			                   hipMemcpyHostToDevice, slot.stream);
		TraceLaunch L = s->shape;
		L.work = slot.work;
		L.spill = slot.spill;
		if (e == hipSuccess)
			e = launchBatches(slot.descs, slot.offsets, nb, off, L, nullptr, 0u, nullptr,
			                  slot.stream);
	}
	for (uint32_t b = 0; b < nb && e == hipSuccess; b++) {
		Request* r = slot.reqs[b];
		if (!(r->req.flags & PHX_MEM_DEVICE) && r->req.numRays)
			e = hipMemcpyAsync(r->req.out, r->devOut, r->outBytes, hipMemcpyDeviceToHost,
			                   slot.stream);
	}
	if (e == hipSuccess) e = hipEventRecord(slot.done, slot.stream);
	if (e == hipSuccess) {
		slot.busy = true;
		s->stats.launches++;
	}
	return e;
}

// Retires a finished launch. Caller holds the lock.
void retire(PhxServer* s, Slot& slot)
{
	for (Request* r : slot.reqs) {
		r->latencyMs = msSince(r->submitted);
		r->state = ReqState::Done;
		if (!(r->req.flags & PHX_MEM_DEVICE)) {
			s->arena.release(r->rayOff, r->rayBytes);
			s->arena.release(r->outOff, r->outBytes);
		}
		s->stats.requestsDone++;
		s->stats.raysDone += r->req.numRays;
	}
	slot.reqs.clear();
	slot.busy = false;
}

void dispatchLoop(PhxServer* s)
{
	hipSetDevice(s->ctx->device);
	std::unique_lock<std::mutex> lk(s->mu);
	while (true) {
		bool anyBusy = false, progressed = false;
		for (Slot& slot : s->slots) {
			if (!slot.busy) continue;
			const hipError_t q = hipEventQuery(slot.done);
			if (q == hipSuccess) {
				retire(s, slot);
				progressed = true;
			} else if (q != hipErrorNotReady) {
				failServer(s, "server launch", q);
				retire(s, slot);
				progressed = true;
			} else {
				anyBusy = true;
			}
		}
		if (progressed) s->cvDone.notify_all();
		admit(s);
		// One launch per idle slot while admitted work is waiting; the plain mode keeps a
		// single launch in flight.
		for (Slot& slot : s->slots) {
			if (s->admitted.empty() || s->paused) break;
			if (slot.busy) continue;
			if ((s->flags & PHX_SERVE_PLAIN) && anyBusy) break;
			const hipError_t e = launchSlot(s, slot);
			if (e != hipSuccess) {
				failServer(s, "server launch", e);
				for (Request* r : slot.reqs) r->state = ReqState::Admitted;
				retire(s, slot);
				s->cvDone.notify_all();
			}
			anyBusy = anyBusy || slot.busy;
			if (s->flags & PHX_SERVE_PLAIN) break;
		}
		if (s->stop && !anyBusy && s->queue.empty() && s->admitted.empty()) break;
		if (anyBusy) {
			// poll the running launches; new submissions also wake the loop
			s->cvWork.wait_for(lk, std::chrono::microseconds(20));
		} else if (s->paused) {
			s->cvWork.wait(lk);  // submit, cancel, pause and destroy notify
		} else {
			s->cvWork.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->admitted.empty(); });
		}
	}
}

void destroyServer(PhxServer* s)
{
	for (Slot& slot : s->slots) {
		if (slot.stream) hipStreamSynchronize(slot.stream);
		hipFree(slot.work);
		hipFree(slot.spill);
		hipFree(slot.descs);
		hipFree(slot.offsets);
		if (slot.hostDescs) hipHostFree(slot.hostDescs);
		if (slot.hostOffsets) hipHostFree(slot.hostOffsets);
		if (slot.done) hipEventDestroy(slot.done);
		if (slot.stream && slot.ownStream) hipStreamDestroy(slot.stream);
	}
	hipFree(s->pool);
	for (auto& kv : s->byTicket) delete kv.second;
	delete s;
}

}  // namespace

extern "C" {

static phx_status phx_server_create_impl(PhxContext* ctx, uint64_t budgetBytes, uint32_t numSlots,
                                         uint32_t flags, PhxServer** outServer)
{
	phx_status st = checkCtx(ctx);
	if (st != PHX_OK) return st;
	if (!outServer) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_create: NULL output");
	*outServer = nullptr;
	if (numSlots == 0 || numSlots > 4)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_create: 1..4 slots, not %u", numSlots);
	size_t freeB = 0, totalB = 0;
	PHX_HIP(hipMemGetInfo(&freeB, &totalB), "phx_server_create");
	if (budgetBytes == 0 || budgetBytes > totalB)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_create: budget %llu outside 1..%zu",
		            (unsigned long long)budgetBytes, totalB);
	PhxServer* s = new (std::nothrow) PhxServer();
	if (!s) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_server_create: out of host memory");
	s->ctx = ctx;
	s->flags = flags;
	s->budget = budgetBytes;
	s->shape = traceLaunch(ctx, 0, false);
	// Continuous batching keeps every block busy: the resident grid of lane-mode blocks of the
	// fused kernel at its largest batch count, no treelet, no helper waves (which would cut the
	// blocks a CU holds), the default stack split and rays per fetch (a context's tuned values
	// are for its single casts).
	s->shape.topNodes = 0;
	s->shape.helperWaves = 0;
	s->shape.ldsDepth = 16;
	s->shape.window = 64;
	uint32_t bpc = 0;
	batchBlocksPerCU(s->shape.blockThreads, 0, 0, 16, kMaxFusedBatches, false, &bpc);
	s->shape.gridBlocks = std::max(1u, ctx->numCUs * bpc);
	s->shape.spillLanes = s->shape.gridBlocks * s->shape.blockThreads;
	s->slots.resize(numSlots);
	hipError_t e = hipSuccess;
	// Slot 0 launches on the context's stream, the others on streams of their own. HIP maps a
	// process's streams onto a few hardware queues (4 by default), and two slots on one queue run
	// one after the other: the null stream and the context's stream already hold two, so only this
	// way do 3 slots get queues of their own in a process that has created no other streams.
	for (size_t i = 0; i < s->slots.size(); i++) {
		Slot& slot = s->slots[i];
		if (i == 0) {
			slot.stream = ctx->stream;
		} else if (e == hipSuccess) {
			e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking);
			slot.ownStream = e == hipSuccess;
		}
		if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&slot.work), 64);
		if (e == hipSuccess) e = hipMemset(slot.work, 0, 64);
		if (e == hipSuccess)
			e = hipMalloc(reinterpret_cast<void**>(&slot.spill),
			              size_t(persistentSpillWords(s->shape.spillLanes)) * 4u);
		if (e == hipSuccess)
			e = hipMalloc(reinterpret_cast<void**>(&slot.descs), sizeof(BatchDesc) * kMaxFusedBatches);
		if (e == hipSuccess)
			e = hipMalloc(reinterpret_cast<void**>(&slot.offsets), 4u * (kMaxFusedBatches + 1));
		if (e == hipSuccess)
			e = hipHostMalloc(reinterpret_cast<void**>(&slot.hostDescs),
			                  sizeof(BatchDesc) * kMaxFusedBatches, hipHostMallocDefault);
		if (e == hipSuccess)
			e = hipHostMalloc(reinterpret_cast<void**>(&slot.hostOffsets),
			                  4u * (kMaxFusedBatches + 1), hipHostMallocDefault);
		if (e == hipSuccess) e = hipEventCreateWithFlags(&slot.done, hipEventDisableTiming);
	}
	if (e != hipSuccess) {
		destroyServer(s);
		return hipFail(e, "phx_server_create");
	}
	s->arena.reset(0);
	holdContext(ctx);
	registry().add(s);
	try {
		s->dispatcher = std::thread(dispatchLoop, s);
	} catch (const std::system_error& err) {
		registry().remove(s);
		destroyServer(s);
		dropContext(ctx);
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_server_create: dispatcher thread: %s", err.what());
	}
	*outServer = s;
	return ok();
}

phx_status phx_server_create(PhxContext* ctx, uint64_t budgetBytes, uint32_t numSlots,
                             uint32_t flags, PhxServer** outServer)
{
	return guarded("phx_server_create", [&] {
		return phx_server_create_impl(ctx, budgetBytes, numSlots, flags, outServer);
	});
}

void phx_server_destroy(PhxServer* s)
{
	if (!s || !registry().remove(s)) return;
	{
		std::lock_guard<std::mutex> g(s->mu);
		s->stop = true;
		s->paused = false;
	}
	s->cvWork.notify_all();
	if (s->dispatcher.joinable()) s->dispatcher.join();
	PhxContext* ctx = s->ctx;
	hipSetDevice(ctx->device);
	const std::unordered_set<const PhxScene*> held = std::move(s->scenes);  // no allocation
	destroyServer(s);
	for (const PhxScene* sc : held) dropScene(sc);
	dropContext(ctx);
}

// Makes `scene` resident for the server: its bytes count against the budget, and the request
// arena is what remains. Scenes are added before requests are submitted.
static phx_status phx_server_add_scene_impl(PhxServer* s, const PhxScene* scene)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	phx_status st = checkScene(s->ctx, scene);
	if (st != PHX_OK) return st;
	std::lock_guard<std::mutex> g(s->mu);
	if (s->scenes.count(scene)) return ok();
	if (!s->queue.empty() || !s->admitted.empty() || s->arena.used())
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_add_scene: requests in flight");
	if (s->sceneBytes + scene->bytes > s->budget)
		return fail(PHX_ERR_BUDGET, "phx_server_add_scene: scene of %zu bytes exceeds the "
		            "budget (%llu of %llu in use)", scene->bytes,
		            (unsigned long long)s->sceneBytes, (unsigned long long)s->budget);
	const uint64_t poolBytes = s->budget - s->sceneBytes - scene->bytes;
	hipSetDevice(s->ctx->device);
	char* pool = nullptr;
	if (poolBytes && hipMalloc(reinterpret_cast<void**>(&pool), poolBytes) != hipSuccess)
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_server_add_scene: request arena of %llu bytes",
		            (unsigned long long)poolBytes);
	hipFree(s->pool);
	s->pool = pool;
	s->arena.reset(poolBytes);
	s->sceneBytes += scene->bytes;
	s->scenes.insert(scene);
	holdScene(scene);
	return ok();
}

phx_status phx_server_add_scene(PhxServer* s, const PhxScene* scene)
{
	return guarded("phx_server_add_scene", [&] {
		return phx_server_add_scene_impl(s, scene);
	});
}

static phx_status phx_server_submit_impl(PhxServer* s, const PhxRequest* req, uint64_t* ticket)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	if (!req || !ticket) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_submit: NULL argument");
	if (req->kind != PHX_BATCH_CLOSEST && req->kind != PHX_BATCH_ANYHIT)
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_submit: kind %u", req->kind);
	if (req->numRays && (!req->rays || !req->out))
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_submit: NULL rays/out");
	if (req->numRays > 0x7FFFFFFFu)  // a fused launch indexes its rays with 31 bits
		return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_submit: %u rays (at most 2^31 - 1)",
		            req->numRays);
	if (!registry().has(req->scene))
		return fail(PHX_ERR_INVALID_HANDLE, "phx_server_submit: released or unknown scene");
	Request* r = new (std::nothrow) Request();
	if (!r) return fail(PHX_ERR_OUT_OF_MEMORY, "phx_server_submit: out of host memory");
	r->req = *req;
	r->submitted = Clock::now();
	r->rayBytes = size_t(req->numRays) * sizeof(PhxRayIn);
	r->outBytes = size_t(req->numRays) * outBytesPerRay(req->kind);
	std::lock_guard<std::mutex> g(s->mu);
	if (!s->error.empty()) {
		delete r;
		return fail(PHX_ERR_HIP, "server failed: %s", s->error.c_str());
	}
	if (!s->scenes.count(req->scene)) {
		delete r;
		return fail(PHX_ERR_INVALID_HANDLE, "phx_server_submit: scene not resident on this server");
	}
	const bool dev = (req->flags & PHX_MEM_DEVICE) != 0;
	if (!dev && alignUp(r->rayBytes) + alignUp(r->outBytes) > s->arena.size()) {
		delete r;
		s->stats.rejected++;
		return fail(PHX_ERR_BUDGET, "phx_server_submit: %u rays need %zu bytes, the request "
		            "budget is %zu", req->numRays, alignUp(r->rayBytes) + alignUp(r->outBytes),
		            s->arena.size());
	}
	r->ticket = s->nextTicket++;
	s->byTicket[r->ticket] = r;
	s->queue.push_back(r);
	s->stats.requestsSubmitted++;
	*ticket = r->ticket;
	s->cvWork.notify_one();
	return ok();
}

phx_status phx_server_submit(PhxServer* s, const PhxRequest* req, uint64_t* ticket)
{
	return guarded("phx_server_submit", [&] {
		return phx_server_submit_impl(s, req, ticket);
	});
}

phx_status phx_server_poll(PhxServer* s, uint64_t ticket, int* done, double* latencyMs)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	if (!done) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_poll: NULL output");
	std::lock_guard<std::mutex> g(s->mu);
	auto it = s->byTicket.find(ticket);
	if (it == s->byTicket.end()) return fail(PHX_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
	if (it->second->state == ReqState::Cancelled) {
		*done = 1;
		return fail(PHX_ERR_CANCELLED, "request %llu was cancelled", (unsigned long long)ticket);
	}
	*done = it->second->state == ReqState::Done;
	if (*done && latencyMs) *latencyMs = it->second->latencyMs;
	if (!s->error.empty()) return fail(PHX_ERR_HIP, "server failed: %s", s->error.c_str());
	return ok();
}

phx_status phx_server_cancel(PhxServer* s, uint64_t ticket, int* cancelled)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	std::lock_guard<std::mutex> g(s->mu);
	auto it = s->byTicket.find(ticket);
	if (it == s->byTicket.end())
		return fail(PHX_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
	Request* r = it->second;
	int done = 0;
	if (r->state == ReqState::Queued || r->state == ReqState::Admitted) {
		std::deque<Request*>& from = r->state == ReqState::Queued ? s->queue : s->admitted;
		from.erase(std::find(from.begin(), from.end(), r));
		if (r->state == ReqState::Admitted && !(r->req.flags & PHX_MEM_DEVICE)) {
			s->arena.release(r->rayOff, r->rayBytes);
			s->arena.release(r->outOff, r->outBytes);
		}
		r->state = ReqState::Cancelled;
		r->latencyMs = msSince(r->submitted);
		s->stats.cancelled++;
		done = 1;
		s->cvDone.notify_all();
		s->cvWork.notify_one();  // freed arena space may admit the next request
	} else if (r->state == ReqState::Cancelled) {
		done = 1;
	}
	if (cancelled) *cancelled = done;
	return ok();
}

phx_status phx_server_pause(PhxServer* s, int pause)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	{
		std::lock_guard<std::mutex> g(s->mu);
		s->paused = pause != 0;
	}
	s->cvWork.notify_all();
	return ok();
}

phx_status phx_server_wait(PhxServer* s, uint64_t ticket, double* latencyMs)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->byTicket.find(ticket);
	if (it == s->byTicket.end()) return fail(PHX_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
	Request* r = it->second;
	s->cvDone.wait(lk, [&] {
		return r->state == ReqState::Done || r->state == ReqState::Cancelled || !s->error.empty();
	});
	if (r->state == ReqState::Cancelled) {
		if (latencyMs) *latencyMs = r->latencyMs;
		s->byTicket.erase(it);
		delete r;
		return fail(PHX_ERR_CANCELLED, "request %llu was cancelled", (unsigned long long)ticket);
	}
	if (r->state != ReqState::Done) return fail(PHX_ERR_HIP, "server failed: %s", s->error.c_str());
	if (latencyMs) *latencyMs = r->latencyMs;
	s->byTicket.erase(it);
	delete r;
	if (!s->error.empty()) return fail(PHX_ERR_HIP, "server failed: %s", s->error.c_str());
	return ok();
}

phx_status phx_server_stats(PhxServer* s, PhxServerStats* out)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	if (!out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_stats: NULL output");
	std::lock_guard<std::mutex> g(s->mu);
	*out = s->stats;
	out->bytesInUse = s->arena.used();
	out->requestBudget = s->arena.size();
	out->sceneBytes = s->sceneBytes;
	out->queued = uint32_t(s->queue.size() + s->admitted.size());
	return ok();
}

static phx_status phx_server_replay_impl(PhxServer* s, const PhxRequest* reqs,
                                         const double* arrivalMs, uint32_t n, double* latencyMs,
                                         double* makespanMs)
{
	if (!s || !registry().has(s)) return fail(PHX_ERR_INVALID_HANDLE, "invalid server handle");
	if (n && (!reqs || !arrivalMs)) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_server_replay: NULL");
	std::vector<uint64_t> tickets(n, 0);
	const Clock::time_point t0 = Clock::now();
	for (uint32_t i = 0; i < n; i++) {
		const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
		                          std::chrono::duration<double, std::milli>(arrivalMs[i]));
		std::this_thread::sleep_until(due);
		phx_status st = phx_server_submit(s, &reqs[i], &tickets[i]);
		if (st != PHX_OK) {
			for (uint32_t j = 0; j < i; j++) phx_server_wait(s, tickets[j], nullptr);
			return st;
		}
	}
	phx_status worst = PHX_OK;
	std::string msg;
	for (uint32_t i = 0; i < n; i++) {
		double lat = 0.0;
		phx_status st = phx_server_wait(s, tickets[i], &lat);
		if (st != PHX_OK && worst == PHX_OK) {
			worst = st;
			msg = phx_last_error();
		}
		if (latencyMs) latencyMs[i] = lat;
	}
	if (makespanMs) *makespanMs = msSince(t0);
	if (worst != PHX_OK) return fail(worst, "%s", msg.c_str());
	return ok();
}

phx_status phx_server_replay(PhxServer* s, const PhxRequest* reqs, const double* arrivalMs,
                             uint32_t n, double* latencyMs, double* makespanMs)
{
	return guarded("phx_server_replay", [&] {
		return phx_server_replay_impl(s, reqs, arrivalMs, n, latencyMs, makespanMs);
	});
}

}  // extern "C"

extern "C" phx_status phx_host_alloc(size_t bytes, void** out)
{
	if (!out) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_host_alloc: NULL output");
	*out = nullptr;
	if (bytes == 0) return ok();
	if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
		*out = nullptr;
		return fail(PHX_ERR_OUT_OF_MEMORY, "phx_host_alloc: %zu bytes of pinned memory", bytes);
	}
	return ok();
}

extern "C" phx_status phx_host_free(void* p)
{
	if (p && hipHostFree(p) != hipSuccess) return fail(PHX_ERR_INVALID_ARGUMENT, "phx_host_free");
	return ok();
}
