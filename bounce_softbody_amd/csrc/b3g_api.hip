// C entry points of libb3g (include/b3g.h).
#include "b3g_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <shared_mutex>
#include <thread>
#include <tuple>

struct b3g_body_s {
	b3g::Body body;
};

namespace b3g {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

// The thread that enqueues a body's eager steps. With the PCG's early exit a
// step's launch sequence depends on the device (the host polls a flag between
// chunks of iterations), so b3g_step queues the whole sequence for this
// thread and returns at once, as an asynchronous step does. The thread takes
// the steps in order, so a step's first kernels are enqueued right behind
// the previous step's last ones and the stream does not run dry; up to
// kMaxQueued steps wait. Every other entry point first waits until the
// queued steps are fully enqueued (drain), so stream order is program order,
// and reports the first failure among them.
// Stream capture and the step threads exclude each other: a HIP call of a
// step thread while another thread captures a stream can invalidate that
// capture. Step threads hold this lock shared while they enqueue a step; a
// capture holds it exclusively.
static std::shared_mutex g_captureLock;

struct StepWorker {
	static constexpr size_t kMaxQueued = 4;
	std::thread thread;
	std::mutex m;
	std::condition_variable cv;
	std::deque<std::function<b3g_status()>> jobs;
	bool running = false, stop = false;
	b3g_status status = B3G_OK;
	std::string error;

	explicit StepWorker(int device) {
		thread = std::thread([this, device]() {
			(void)hipSetDevice(device);
			std::unique_lock<std::mutex> lk(m);
			for (;;) {
				cv.wait(lk, [&] { return !jobs.empty() || stop; });
				if (jobs.empty()) return;
				std::function<b3g_status()> fn = std::move(jobs.front());
				jobs.pop_front();
				running = true;
				const bool skip = status != B3G_OK;   // after a failure the queued steps are dropped
				lk.unlock();
				cv.notify_all();
				b3g_status s = B3G_OK;
				if (!skip) {
					std::shared_lock<std::shared_mutex> cap(g_captureLock);
					s = fn();
				}
				const std::string err = s != B3G_OK ? g_error : std::string();
				lk.lock();
				if (s != B3G_OK && status == B3G_OK) { status = s; error = err; }
				running = false;
				cv.notify_all();
			}
		});
	}
	// Waits until every queued step is enqueued; returns (and clears) the first failure since the last drain.
	b3g_status drain() {
		std::unique_lock<std::mutex> lk(m);
		cv.wait(lk, [&] { return jobs.empty() && !running; });
		const b3g_status s = status;
		if (s != B3G_OK) set_error(error);
		status = B3G_OK;
		error.clear();
		return s;
	}
	b3g_status post(std::function<b3g_status()> fn) {
		std::unique_lock<std::mutex> lk(m);
		if (status != B3G_OK) {   // a queued step failed: report it now
			const b3g_status s = status;
			cv.wait(lk, [&] { return jobs.empty() && !running; });
			set_error(error);
			status = B3G_OK;
			error.clear();
			return s;
		}
		cv.wait(lk, [&] { return jobs.size() < kMaxQueued; });
		jobs.push_back(std::move(fn));
		cv.notify_all();
		return B3G_OK;
	}
	~StepWorker() {
		{
			std::unique_lock<std::mutex> lk(m);
			cv.wait(lk, [&] { return jobs.empty() && !running; });
			stop = true;
			cv.notify_all();
		}
		thread.join();
	}
};

static b3g_status drain(Body& b) { return b.worker ? b.worker->drain() : B3G_OK; }

} // namespace b3g

using namespace b3g;

#define CHECK_BODY(bd)                                                  \
	do {                                                                \
		if (!(bd)) { set_error("null body handle"); return B3G_INVALID_ARGUMENT; } \
		b3g_status ds_ = drain((bd)->body);                             \
		if (ds_ != B3G_OK) return ds_;                                  \
	} while (0)
#define CHECK_BUILT(bd)                                                 \
	do {                                                                \
		CHECK_BODY(bd);                                                 \
		if (!(bd)->body.built) { set_error("body has not been built"); return B3G_NOT_COMMITTED; } \
	} while (0)
#define SET_DEVICE(bd) B3G_HIP(hipSetDevice((bd)->body.device))

namespace {

// Host copy of a list-order V3 array into creation order.
b3g_status read_v3(Body& b, const DBuf& buf, float* out) {
	std::vector<V3> tmp(b.n);
	B3G_HIP(hipMemcpyAsync(tmp.data(), buf.p, b.n * sizeof(V3), hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	for (uint32_t c = 0; c < b.n; ++c) {
		const V3& v = tmp[b.n - 1 - c];
		out[3 * c] = v.x; out[3 * c + 1] = v.y; out[3 * c + 2] = v.z;
	}
	return B3G_OK;
}

b3g_status write_v3(Body& b, DBuf& buf, const float* in, bool accumulate, const uint8_t* onlyType = nullptr) {
	std::vector<V3> tmp(b.n);
	if (accumulate) {
		B3G_HIP(hipMemcpyAsync(tmp.data(), buf.p, b.n * sizeof(V3), hipMemcpyDeviceToHost, b.stream));
		B3G_HIP(hipStreamSynchronize(b.stream));
	}
	for (uint32_t c = 0; c < b.n; ++c) {
		uint32_t i = b.n - 1 - c;
		V3 v = v3(in[3 * c], in[3 * c + 1], in[3 * c + 2]);
		if (onlyType && b.hostType[i] != *onlyType) continue;
		tmp[i] = accumulate ? tmp[i] + v : v;   // m_force += force (particle.h:249-256)
	}
	B3G_HIP(hipMemcpyAsync(buf.p, tmp.data(), b.n * sizeof(V3), hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	return B3G_OK;
}

} // namespace

extern "C" {

const char* b3g_status_string(b3g_status s) {
	switch (s) {
	case B3G_OK: return "ok";
	case B3G_INVALID_ARGUMENT: return "invalid argument";
	case B3G_HIP_ERROR: return "HIP error";
	case B3G_OUT_OF_MEMORY: return "out of device memory";
	case B3G_DEGENERATE_ELEMENT: return "degenerate element";
	case B3G_UNSUPPORTED: return "unsupported";
	case B3G_NOT_COMMITTED: return "body not built";
	case B3G_NON_POSITIVE_DIAGONAL: return "non-positive Jacobi diagonal";
	case B3G_NON_FINITE: return "non-finite state";
	}
	return "unknown status";
}

const char* b3g_last_error(void) { return g_error.c_str(); }

b3g_status b3g_device_count(int* count) {
	if (!count) return B3G_INVALID_ARGUMENT;
	hipError_t e = hipGetDeviceCount(count);
	if (e != hipSuccess) { *count = 0; set_error(hipGetErrorString(e)); return B3G_HIP_ERROR; }
	return B3G_OK;
}

b3g_status b3g_body_create(b3g_body* out, int device) {
	if (!out) { set_error("out is NULL"); return B3G_INVALID_ARGUMENT; }
	*out = nullptr;
	int count = 0;
	B3G_HIP(hipGetDeviceCount(&count));
	if (device < 0 || device >= count) { set_error("device index out of range"); return B3G_INVALID_ARGUMENT; }
	B3G_HIP(hipSetDevice(device));
	b3g_body_s* h = new (std::nothrow) b3g_body_s();
	if (!h) return B3G_OUT_OF_MEMORY;
	h->body.device = device;
	hipError_t e = hipStreamCreateWithFlags(&h->body.stream, hipStreamNonBlocking);
	if (e != hipSuccess) { set_error(hipGetErrorString(e)); delete h; return B3G_HIP_ERROR; }
	h->body.earlyExit = true;
	*out = h;
	return B3G_OK;
}

b3g_status b3g_body_destroy(b3g_body h) {
	if (!h) return B3G_OK;
	delete h->body.worker;   // waits for the step in hand to be enqueued
	h->body.worker = nullptr;
	(void)hipSetDevice(h->body.device);
	(void)hipStreamSynchronize(h->body.stream);
	drop_graph(h->body);
	dfree_all(h->body);
	free_exit_flags(h->body);
	free_persistent(h->body);
	(void)hipStreamDestroy(h->body.stream);
	delete h;
	return B3G_OK;
}

b3g_status b3g_body_build(b3g_body h, const b3g_body_desc* d) {
	CHECK_BODY(h);
	if (!d) { set_error("desc is NULL"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	return build_body(h->body, d, 1);
}

b3g_status b3g_body_build_batch(b3g_body h, const b3g_body_desc* descs, uint32_t count) {
	CHECK_BODY(h);
	if (!descs || count == 0) { set_error("a batch needs at least one description"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	return build_body(h->body, descs, count);
}

b3g_status b3g_body_rebuild(b3g_body h, const b3g_body_desc* descs, uint32_t count, const b3g_carry* carry) {
	CHECK_BUILT(h);
	if (!descs || count == 0) { set_error("a rebuild needs at least one description"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	Body& b = h->body;
	B3G_HIP(hipStreamSynchronize(b.stream));
	// the previous body's state, on the host
	const uint32_t n0 = b.n, m0 = b.tet.m, pairs0 = b.nPairs, nw0 = b.nWorld;
	std::vector<V3> x0(n0), v0(n0);
	std::vector<float4> q0(m0);
	B3G_HIP(hipMemcpy(x0.data(), b.x.p, n0 * sizeof(V3), hipMemcpyDeviceToHost));
	B3G_HIP(hipMemcpy(v0.data(), b.v.p, n0 * sizeof(V3), hipMemcpyDeviceToHost));
	if (m0) B3G_HIP(hipMemcpy(q0.data(), b.tet.q.p, m0 * sizeof(float4), hipMemcpyDeviceToHost));
	std::vector<uint32_t> act0(pairs0), stamp0(pairs0), fric0(pairs0);
	std::vector<float> nf0(pairs0);
	std::vector<V3> man0(2 * (size_t)pairs0);
	uint32_t counter0 = 0;
	if (pairs0) {
		B3G_HIP(hipMemcpy(act0.data(), b.cActive.p, pairs0 * 4, hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(stamp0.data(), b.cStamp.p, pairs0 * 4, hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(fric0.data(), b.cFriction.p, pairs0 * 4, hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(nf0.data(), b.cNormalForce.p, pairs0 * 4, hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(man0.data(), b.cManifold.p, pairs0 * 2 * sizeof(V3), hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(&counter0, b.stampCounter.p, 4, hipMemcpyDeviceToHost));
	}
	std::map<std::pair<uint32_t, uint32_t>, uint32_t> pair0;   // (sphere, world) creation indices -> pair
	std::vector<uint32_t> call0(pairs0);                          // the FindNewContacts call that created it
	for (uint32_t k = 0; k < pairs0; ++k) {
		const uint32_t sp = b.hostPairSphere[k];
		pair0[{b.hostSphereCreation[sp], nw0 - 1 - b.hostPairWorld[k]}] = k;
		call0[k] = stamp0[k] / (b.hostPairStart[sp + 1] - b.hostPairStart[sp]);
	}
	uint32_t sphereCount0 = 0;
	for (uint32_t c : b.hostSphereCreation) sphereCount0 = std::max(sphereCount0, c + 1);

	// merged counts of the new description; every map entry is checked before
	// anything is rebuilt, so a rejected call leaves the body as it was
	uint64_t np = 0, nt = 0, nsph = 0, nw = 0;
	for (uint32_t k = 0; k < count; ++k) {
		np += descs[k].n_particles; nt += descs[k].n_tet; nsph += descs[k].n_spheres; nw += descs[k].n_world;
	}
	if (carry) {
		auto bad = [](const char* what) { set_error(std::string("carry map out of range: ") + what); return B3G_INVALID_ARGUMENT; };
		auto check = [](const int64_t* map, uint64_t len, uint64_t prev) {
			for (uint64_t c = 0; map && c < len; ++c)
				if (map[c] >= (int64_t)prev) return false;
			return true;
		};
		if (!check(carry->particle_prev, np, n0)) return bad("particle");
		if (!check(carry->tet_prev, nt, m0)) return bad("tetrahedron");
		if (!check(carry->sphere_prev, nsph, sphereCount0)) return bad("sphere");
		if (!check(carry->world_prev, nw, nw0)) return bad("world");
	}

	b3g_status s = build_body(b, descs, count);
	if (s != B3G_OK) return s;
	if (!carry) return B3G_OK;
	if (carry->particle_prev) {
		std::vector<V3> x(b.n), v(b.n);
		B3G_HIP(hipMemcpy(x.data(), b.x.p, b.n * sizeof(V3), hipMemcpyDeviceToHost));
		B3G_HIP(hipMemcpy(v.data(), b.v.p, b.n * sizeof(V3), hipMemcpyDeviceToHost));
		for (uint64_t c = 0; c < np; ++c) {
			const int64_t q = carry->particle_prev[c];
			if (q < 0) continue;
			x[b.n - 1 - c] = x0[n0 - 1 - q];
			v[b.n - 1 - c] = v0[n0 - 1 - q];
		}
		B3G_HIP(hipMemcpy(b.x.p, x.data(), b.n * sizeof(V3), hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.v.p, v.data(), b.n * sizeof(V3), hipMemcpyHostToDevice));
	}
	if (carry->tet_prev && b.tet.m) {
		std::vector<float4> q(b.tet.m);
		B3G_HIP(hipMemcpy(q.data(), b.tet.q.p, b.tet.m * sizeof(float4), hipMemcpyDeviceToHost));
		for (uint64_t c = 0; c < nt; ++c) {
			const int64_t p = carry->tet_prev[c];
			if (p >= 0) q[c] = q0[p];
		}
		B3G_HIP(hipMemcpy(b.tet.q.p, q.data(), b.tet.m * sizeof(float4), hipMemcpyHostToDevice));
	}
	if (carry->sphere_prev && carry->world_prev && b.nPairs && pairs0) {
		const uint32_t P = b.nPairs;
		std::vector<uint32_t> act(P, 0), stamp(P, 0), fric(P, 0);
		std::vector<float> nf(P, 0.0f);
		std::vector<V3> man(2 * (size_t)P, zero3());
		// a carried contact keeps the FindNewContacts call that created it; its
		// stamp is re-encoded with the sphere's new pair count (call * pairs +
		// world), so the list order (newest call first, then sphere, then
		// world) is unchanged and later calls still sort after it
		for (uint32_t sp = 0; sp < b.nSpheres; ++sp) {
			const uint32_t p0 = b.hostPairStart[sp], nwp = b.hostPairStart[sp + 1] - p0;
			for (uint32_t k = p0; k < b.hostPairStart[sp + 1]; ++k) {
				const int64_t ps = carry->sphere_prev[b.hostSphereCreation[sp]];
				const int64_t pw = carry->world_prev[b.nWorld - 1 - b.hostPairWorld[k]];
				if (ps < 0 || pw < 0) continue;
				auto it = pair0.find({(uint32_t)ps, (uint32_t)pw});
				if (it == pair0.end() || !act0[it->second]) continue;
				const uint32_t o = it->second;
				act[k] = 1;
				fric[k] = fric0[o];
				nf[k] = nf0[o];
				man[2 * (size_t)k] = man0[2 * (size_t)o];
				man[2 * (size_t)k + 1] = man0[2 * (size_t)o + 1];
				stamp[k] = call0[o] * nwp + (k - p0);
			}
		}
		const uint32_t counter = counter0;
		B3G_HIP(hipMemcpy(b.cActive.p, act.data(), P * 4, hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.cStamp.p, stamp.data(), P * 4, hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.cFriction.p, fric.data(), P * 4, hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.cNormalForce.p, nf.data(), P * 4, hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.cManifold.p, man.data(), P * 2 * sizeof(V3), hipMemcpyHostToDevice));
		B3G_HIP(hipMemcpy(b.stampCounter.p, &counter, 4, hipMemcpyHostToDevice));
	}
	++b.xVersion;
	return B3G_OK;
}

b3g_status b3g_set_action_forces(b3g_body h, int enabled) {
	CHECK_BODY(h);
	SET_DEVICE(h);
	Body& b = h->body;
	B3G_HIP(hipStreamSynchronize(b.stream));
	b.trackActions = enabled != 0;
	drop_graph(b);
	if (b.built) return ensure_actions(b);
	return B3G_OK;
}

b3g_status b3g_get_action_forces(b3g_body h, uint32_t force_type, float* out, uint64_t capacity) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	Body& b = h->body;
	if (!b.trackActions) { set_error("action forces are not tracked (b3g_set_action_forces)"); return B3G_INVALID_ARGUMENT; }
	uint32_t m, per, base;
	switch (force_type) {
	case B3G_FORCE_SPRING: m = b.spring.m; per = 2; base = b.actSpring; break;
	case B3G_FORCE_STRETCH: m = b.stretch.m; per = 3; base = b.actStretch; break;
	case B3G_FORCE_SHEAR: m = b.shear.m; per = 3; base = b.actShear; break;
	case B3G_FORCE_MOUSE: m = b.mouse.m; per = 4; base = b.actMouse; break;
	default: set_error("no action forces for this force type"); return B3G_INVALID_ARGUMENT;
	}
	if (!out || capacity < 3ull * per * m) { set_error("output holds fewer than 3 * per * count floats"); return B3G_INVALID_ARGUMENT; }
	if (!m) return B3G_OK;
	std::vector<V3> planes((size_t)per * m);
	B3G_HIP(hipStreamSynchronize(b.stream));
	B3G_HIP(hipMemcpy(planes.data(), dptr<V3>(b.act) + base, planes.size() * sizeof(V3), hipMemcpyDeviceToHost));
	for (uint32_t e = 0; e < m; ++e)
		for (uint32_t i = 0; i < per; ++i) {
			const V3 f = planes[(size_t)i * m + e];
			float* o = out + 3 * ((size_t)per * e + i);
			o[0] = f.x; o[1] = f.y; o[2] = f.z;
		}
	return B3G_OK;
}

b3g_status b3g_get_instance_count(b3g_body h, uint32_t* count) {
	CHECK_BUILT(h);
	if (!count) return B3G_INVALID_ARGUMENT;
	*count = h->body.nInst;
	return B3G_OK;
}

b3g_status b3g_sdf_grid(const float* vertices, uint32_t vertex_count, const float cell_size[3], float aabb_extension,
                        float box[6], uint32_t dims[3]) {
	if (!box || !dims) { set_error("null output"); return B3G_INVALID_ARGUMENT; }
	AABB b;
	b3g_status s = sdf_grid(vertices, vertex_count, cell_size, aabb_extension, b, dims);
	if (s != B3G_OK) return s;
	box[0] = b.lo.x; box[1] = b.lo.y; box[2] = b.lo.z;
	box[3] = b.hi.x; box[4] = b.hi.y; box[5] = b.hi.z;
	return B3G_OK;
}

b3g_status b3g_build_sdf(int device, const float* vertices, uint32_t vertex_count, const uint32_t* triangles,
                         uint32_t triangle_count, uint32_t triangle_stride, const float cell_size[3],
                         float aabb_extension, float* voxels, uint64_t capacity) {
	return build_sdf(device, vertices, vertex_count, triangles, triangle_count, triangle_stride, cell_size,
	                 aabb_extension, voxels, capacity);
}

b3g_status b3g_ray_cast(b3g_body h, uint32_t instance, uint32_t count, const float* p1, const float* p2, int64_t* index,
                        float* fraction, float* normal) {
	CHECK_BUILT(h);
	if (count && (!p1 || !p2 || !index || !fraction)) { set_error("null ray or result array"); return B3G_INVALID_ARGUMENT; }
	if (instance >= h->body.nInst) { set_error("instance out of range"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	return ray_cast(h->body, instance, count, p1, p2, index, fraction, normal);
}

b3g_status b3g_ray_cast_device(b3g_body h, uint32_t instance, uint32_t count, const float* p1, const float* p2,
                               int64_t* index, float* fraction, float* normal) {
	CHECK_BUILT(h);
	if (count && (!p1 || !p2 || !index || !fraction)) { set_error("null ray or result array"); return B3G_INVALID_ARGUMENT; }
	if (instance >= h->body.nInst) { set_error("instance out of range"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	if (count) {
		// only device memory is accepted: a host pointer would fault the kernel
		const void* ptrs[5] = {p1, p2, index, fraction, normal};
		for (const void* p : ptrs) {
			if (!p) continue;
			hipPointerAttribute_t attr;
			if (hipPointerGetAttributes(&attr, p) != hipSuccess || attr.type != hipMemoryTypeDevice) {
				(void)hipGetLastError();
				set_error("b3g_ray_cast_device: an array is not device memory");
				return B3G_INVALID_ARGUMENT;
			}
		}
	}
	return ray_cast_device(h->body, instance, count, p1, p2, index, fraction, normal);
}

b3g_status b3g_step(b3g_body h, float dt, uint32_t iterations, uint32_t sub_iterations, float tolerance,
                    float sub_tolerance) {
	if (!h) { set_error("null body handle"); return B3G_INVALID_ARGUMENT; }
	if (!std::isfinite(dt)) { set_error("dt must be finite"); return B3G_INVALID_ARGUMENT; }
	float tol = tolerance > 0.0f ? tolerance : kEps;
	float stol = sub_tolerance > 0.0f ? sub_tolerance : kEps;
	if (!(stol < 1.0f)) { set_error("sub_tolerance must be in (0, 1)"); return B3G_INVALID_ARGUMENT; }   // sparse_solver.cpp:36
	Body& b = h->body;
	b.lastStep.valid = true;
	b.lastStep.dt = dt;
	b.lastStep.it = iterations;
	b.lastStep.sub = sub_iterations;
	b.lastStep.tol = tolerance;
	b.lastStep.stol = sub_tolerance;
	b3g_status s;
	if (b.earlyExit && !b.useGraph && b.built) {
		// queued for the body's thread, which also makes the step's one-time preparations; steps already
		// queued keep running meanwhile (no drain here)
		B3G_HIP(hipSetDevice(b.device));
		if (!b.worker) {
			if ((s = ensure_exit_flags(b)) != B3G_OK) return s;
			b.worker = new (std::nothrow) StepWorker(b.device);
			if (!b.worker) return B3G_OUT_OF_MEMORY;
		}
		++b.xVersion;
		return b.worker->post([&b, dt, iterations, sub_iterations, tol, stol]() {
			b3g_status st = prepare_step(b, sub_iterations);
			if (st != B3G_OK) return st;
			return enqueue_step(b, dt, iterations, sub_iterations, tol, stol, false);
		});
	}
	CHECK_BUILT(h);
	SET_DEVICE(h);
	s = prepare_step(b, sub_iterations);
	if (s != B3G_OK) return s;
	++b.xVersion;
	if (!b.useGraph) return enqueue_step(b, dt, iterations, sub_iterations, tol, stol, false);
	// Captured step: one graph per set of step parameters and modes, replayed.
	const std::vector<double> key = {(double)dt, (double)iterations, (double)sub_iterations, (double)tol, (double)stol,
	                                  (double)b.reduction, (double)b.solver, (double)b.innerIt, (double)b.innerTol,
	                                  (double)b.logStride, (double)b.trackActions};
	if (!b.graphExec || key != b.graphKey) {
		drop_graph(b);
		hipGraph_t graph = nullptr;
		std::unique_lock<std::shared_mutex> cap(g_captureLock);   // no step thread runs during the capture
		B3G_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
		b.capturing = true;   // the captured sequence is the fixed one
		s = enqueue_step(b, dt, iterations, sub_iterations, tol, stol, false);
		b.capturing = false;
		hipError_t e = hipStreamEndCapture(b.stream, &graph);
		if (s != B3G_OK) { if (graph) (void)hipGraphDestroy(graph); return s; }
		if (e != hipSuccess) { set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); return B3G_HIP_ERROR; }
		e = hipGraphInstantiate(&b.graphExec, graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		if (e != hipSuccess) {
			b.graphExec = nullptr;
			set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
			return B3G_HIP_ERROR;
		}
		b.graphKey = key;
	}
	B3G_HIP(hipGraphLaunch(b.graphExec, b.stream));
	return B3G_OK;
}

b3g_status b3g_set_graph(b3g_body h, int enabled) {
	CHECK_BODY(h);
	SET_DEVICE(h);
	B3G_HIP(hipStreamSynchronize(h->body.stream));
	h->body.useGraph = enabled != 0;
	if (!h->body.useGraph) drop_graph(h->body);
	return B3G_OK;
}

b3g_status b3g_get_state(b3g_body h, float* pos, float* vel) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	if (pos) { b3g_status s = read_v3(h->body, h->body.x, pos); if (s != B3G_OK) return s; }
	if (vel) { b3g_status s = read_v3(h->body, h->body.v, vel); if (s != B3G_OK) return s; }
	return B3G_OK;
}

b3g_status b3g_set_state(b3g_body h, const float* pos, const float* vel) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	Body& b = h->body;
	if (pos) {
		++b.xVersion;
		b3g_status s = write_v3(b, b.x, pos, false);
		if (s != B3G_OK) return s;
		// SetPosition clears the pending translation (particle.h:213-218)
		std::vector<V3> z(b.n, zero3());
		B3G_HIP(hipStreamSynchronize(b.stream));
		B3G_HIP(hipMemcpy(b.translation.p, z.data(), b.n * sizeof(V3), hipMemcpyHostToDevice));
	}
	if (vel) {
		// SetVelocity is ignored for static particles (particle.h:225-232)
		std::vector<V3> cur(b.n);
		B3G_HIP(hipStreamSynchronize(b.stream));
		B3G_HIP(hipMemcpy(cur.data(), b.v.p, b.n * sizeof(V3), hipMemcpyDeviceToHost));
		for (uint32_t c = 0; c < b.n; ++c) {
			uint32_t i = b.n - 1 - c;
			if (b.hostType[i] != kStatic) cur[i] = v3(vel[3 * c], vel[3 * c + 1], vel[3 * c + 2]);
		}
		B3G_HIP(hipMemcpy(b.v.p, cur.data(), b.n * sizeof(V3), hipMemcpyHostToDevice));
	}
	return B3G_OK;
}

b3g_status b3g_apply_external(b3g_body h, const float* force, const float* translation) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	Body& b = h->body;
	const uint8_t dyn = kDynamic;
	if (force) { b3g_status s = write_v3(b, b.force, force, true, &dyn); if (s != B3G_OK) return s; }
	if (translation) { b3g_status s = write_v3(b, b.translation, translation, true); if (s != B3G_OK) return s; }
	return B3G_OK;
}

b3g_status b3g_get_masses(b3g_body h, float* mass) {
	CHECK_BUILT(h);
	if (!mass) return B3G_INVALID_ARGUMENT;
	const Body& b = h->body;
	for (uint32_t c = 0; c < b.n; ++c) mass[c] = b.hostMass[b.n - 1 - c];
	return B3G_OK;
}

b3g_status b3g_get_stats(b3g_body h, b3g_step_stats* out) {
	CHECK_BUILT(h);
	if (!out) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	std::vector<SolverState> st(b.nInst);
	B3G_HIP(hipStreamSynchronize(b.stream));
	B3G_HIP(hipMemcpy(st.data(), b.state.p, st.size() * sizeof(SolverState), hipMemcpyDeviceToHost));
	std::vector<uint32_t> count(b.nInst, 0);
	if (b.nPairs) {
		std::vector<uint32_t> act(b.nPairs);
		B3G_HIP(hipMemcpy(act.data(), b.cActive.p, act.size() * 4, hipMemcpyDeviceToHost));
		for (uint32_t k = 0; k < b.nPairs; ++k)
			if (act[k]) count[b.hostRowInst[b.hostSphereParticle[b.hostPairSphere[k]]]] += 1;
	}
	for (uint32_t k = 0; k < b.nInst; ++k) {
		out[k].newton_iterations = st[k].statNewton;
		out[k].min_sub_iterations = st[k].statMinSub;
		out[k].max_sub_iterations = st[k].statMaxSub;
		out[k].contact_count = count[k];
		out[k].refinement_sweeps = st[k].refineOn ? st[k].statSweeps : 0u;
		out[k].relative_residual = st[k].refineOn && st[k].res0 > 0.0 ? (float)std::sqrt(st[k].res / st[k].res0) : 0.0f;
		out[k].faults = st[k].faults;
	}
	return B3G_OK;
}

b3g_status b3g_set_solver(b3g_body h, int mode, uint32_t inner_iterations, float inner_tolerance) {
	CHECK_BODY(h);
	if (mode != B3G_SOLVER_PCG && mode != B3G_SOLVER_REFINED_FP32 && mode != B3G_SOLVER_REFINED_BF16) {
		set_error("unknown solver mode");
		return B3G_INVALID_ARGUMENT;
	}
	if (mode != B3G_SOLVER_PCG && (inner_iterations == 0 || !(inner_tolerance > 0.0f && inner_tolerance < 1.0f))) {
		set_error("refinement needs inner_iterations >= 1 and inner_tolerance in (0, 1)");
		return B3G_INVALID_ARGUMENT;
	}
	Body& b = h->body;
	b.solver = mode;
	if (mode != B3G_SOLVER_PCG) {
		b.innerIt = inner_iterations;
		b.innerTol = inner_tolerance;
	}
	return B3G_OK;
}

b3g_status b3g_get_energy(b3g_body h, float* energy) {
	CHECK_BUILT(h);
	if (!energy) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	std::vector<V3> v(b.n);
	B3G_HIP(hipMemcpyAsync(v.data(), b.v.p, b.n * sizeof(V3), hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	for (uint32_t k = 0; k < b.nInst; ++k) {
		float E = 0.0f;
		for (uint32_t i = b.hostRowBegin[k]; i < b.hostRowEnd[k]; ++i) E += b.hostMass[i] * dot(v[i], v[i]);   // body.cpp:384-392
		energy[k] = 0.5f * E;
	}
	return B3G_OK;
}

b3g_status b3g_get_contacts(b3g_body h, uint32_t* count, uint32_t* particle, uint32_t* world, float* normal_force,
                            uint32_t capacity) {
	CHECK_BUILT(h);
	if (!count) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	*count = 0;
	if (!b.nPairs) return B3G_OK;
	const size_t pairs = b.nPairs;
	std::vector<uint32_t> act(pairs), stamp(pairs);
	std::vector<float> nf(pairs);
	B3G_HIP(hipStreamSynchronize(b.stream));
	B3G_HIP(hipMemcpy(act.data(), b.cActive.p, pairs * 4, hipMemcpyDeviceToHost));
	B3G_HIP(hipMemcpy(stamp.data(), b.cStamp.p, pairs * 4, hipMemcpyDeviceToHost));
	B3G_HIP(hipMemcpy(nf.data(), b.cNormalForce.p, pairs * 4, hipMemcpyDeviceToHost));
	// Per instance (creation order), the list order (contact_manager.cpp:54-62):
	// newest FindNewContacts call first; within a call the last created pair
	// (sphere list order, then world list order) is the head.
	struct Key { uint32_t inst, call, s, w; size_t k; };
	std::vector<Key> keys;
	for (size_t k = 0; k < pairs; ++k)
		if (act[k]) {
			const uint32_t s = b.hostPairSphere[k], p0 = b.hostPairStart[s], nw = b.hostPairStart[s + 1] - p0;
			keys.push_back({b.hostRowInst[b.hostSphereParticle[s]], stamp[k] / nw, s, (uint32_t)(k - p0), k});
		}
	std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& c) {
		if (a.inst != c.inst) return a.inst < c.inst;
		return std::make_tuple(a.call, a.s, a.w) > std::make_tuple(c.call, c.s, c.w);
	});
	*count = (uint32_t)keys.size();
	for (size_t i = 0; i < keys.size() && i < capacity; ++i) {
		const Key& kk = keys[i];
		if (particle) particle[i] = b.n - 1 - b.hostSphereParticle[kk.s];
		if (world) world[i] = b.nWorld - 1 - b.hostPairWorld[kk.k];
		if (normal_force) normal_force[i] = nf[kk.k];
	}
	return B3G_OK;
}

b3g_status b3g_get_tet_quats(b3g_body h, float* quats) {
	CHECK_BUILT(h);
	if (!quats) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	const uint32_t m = b.tet.m;
	if (!m) return B3G_OK;
	B3G_HIP(hipStreamSynchronize(b.stream));
	B3G_HIP(hipMemcpy(quats, b.tet.q.p, 16 * (size_t)m, hipMemcpyDeviceToHost));
	return B3G_OK;
}

b3g_status b3g_assemble(b3g_body h, float dt, uint32_t* nnz) {
	CHECK_BUILT(h);
	if (!(dt > 0.0f)) { set_error("dt must be positive"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	b3g_status s = enqueue_step(h->body, dt, 1, 0, kEps, kEps, true);
	if (s != B3G_OK) return s;
	B3G_HIP(hipStreamSynchronize(h->body.stream));
	if (nnz) *nnz = h->body.nnz;
	return B3G_OK;
}

b3g_status b3g_get_system(b3g_body h, uint32_t* rowptr, uint32_t* cols, float* vals, float* pb, float* fi) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	Body& b = h->body;
	B3G_HIP(hipStreamSynchronize(b.stream));
	if (rowptr) std::memcpy(rowptr, b.hostRowptr.data(), (b.n + 1) * 4);
	if (cols) std::memcpy(cols, b.hostCols.data(), (size_t)b.nnz * 4);
	if (vals) {
		std::vector<float> A((size_t)b.aStride * 9);
		B3G_HIP(hipMemcpy(A.data(), b.A.p, A.size() * 4, hipMemcpyDeviceToHost));
		std::vector<uint32_t> pos;
		if (b.sortedA) {   // storage position of each CSR slot
			pos.resize(b.nnz);
			B3G_HIP(hipMemcpy(pos.data(), b.slotPos.p, pos.size() * 4, hipMemcpyDeviceToHost));
		}
		for (size_t k = 0; k < b.nnz; ++k) {
			const size_t at = b.sortedA ? pos[k] : k;
			for (int c = 0; c < 9; ++c) vals[9 * k + c] = A[(size_t)c * b.aStride + at];
		}
	}
	auto rows = [&](const DBuf& buf, float* out) -> b3g_status {
		std::vector<V3> t(b.n);
		B3G_HIP(hipMemcpy(t.data(), buf.p, b.n * sizeof(V3), hipMemcpyDeviceToHost));
		for (uint32_t i = 0; i < b.n; ++i) { out[3 * i] = t[i].x; out[3 * i + 1] = t[i].y; out[3 * i + 2] = t[i].z; }
		return B3G_OK;
	};
	if (pb) { b3g_status s = rows(b.pb, pb); if (s != B3G_OK) return s; }
	if (fi) { b3g_status s = rows(b.fi, fi); if (s != B3G_OK) return s; }
	return B3G_OK;
}

b3g_status b3g_set_reduction(b3g_body h, int mode) {
	CHECK_BODY(h);
	if (mode != B3G_REDUCE_FP64 && mode != B3G_REDUCE_SEQUENTIAL && mode != B3G_REDUCE_FP64_FUSED &&
	    mode != B3G_REDUCE_FP64_SEPARATE) {
		set_error("unknown reduction mode");
		return B3G_INVALID_ARGUMENT;
	}
	h->body.reduction = mode;
	return B3G_OK;
}

} // extern "C"

namespace {
struct StateHeader {
	uint32_t magic, version, n, nTet, pairs, stampCounter, nInst, reserved;
};
constexpr uint32_t kStateMagic = 0x53473342u;   // "B3GS"
constexpr uint32_t kStateVersion = 4;            // 3: batches (per-instance solver statistics); 4: fault bits

struct StatePart {
	DBuf* buf;
	size_t bytes;
};

std::vector<StatePart> state_parts(Body& b) {
	const size_t pairs = b.nPairs;
	std::vector<StatePart> parts = {{&b.x, b.n * sizeof(V3)}, {&b.v, b.n * sizeof(V3)}, {&b.force, b.n * sizeof(V3)},
	                                {&b.translation, b.n * sizeof(V3)}, {&b.state, b.nInst * sizeof(SolverState)}};
	if (b.tet.m) parts.push_back({&b.tet.q, (size_t)b.tet.m * 16});
	if (pairs) {
		parts.push_back({&b.cActive, pairs * 4});
		parts.push_back({&b.cStamp, pairs * 4});
		parts.push_back({&b.cFriction, pairs * 4});
		parts.push_back({&b.cNormalForce, pairs * 4});
		parts.push_back({&b.cManifold, pairs * 2 * sizeof(V3)});
	}
	return parts;
}
} // namespace

extern "C" {

b3g_status b3g_save_state(b3g_body h, void* blob, uint64_t* bytes) {
	CHECK_BUILT(h);
	if (!bytes) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	std::vector<StatePart> parts = state_parts(b);
	size_t total = sizeof(StateHeader);
	for (auto& p : parts) total += p.bytes;
	if (!blob) { *bytes = total; return B3G_OK; }
	if (*bytes < total) { set_error("state blob too small"); return B3G_INVALID_ARGUMENT; }
	B3G_HIP(hipStreamSynchronize(b.stream));
	uint32_t stamps = 0;
	B3G_HIP(hipMemcpy(&stamps, b.stampCounter.p, 4, hipMemcpyDeviceToHost));
	StateHeader hd = {kStateMagic, kStateVersion, b.n, b.tet.m, b.nPairs, stamps, b.nInst, 0};
	char* out = static_cast<char*>(blob);
	std::memcpy(out, &hd, sizeof(hd));
	size_t off = sizeof(hd);
	for (auto& p : parts) {
		B3G_HIP(hipMemcpy(out + off, p.buf->p, p.bytes, hipMemcpyDeviceToHost));
		off += p.bytes;
	}
	*bytes = total;
	return B3G_OK;
}

b3g_status b3g_load_state(b3g_body h, const void* blob, uint64_t bytes) {
	CHECK_BUILT(h);
	if (!blob) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	std::vector<StatePart> parts = state_parts(b);
	size_t total = sizeof(StateHeader);
	for (auto& p : parts) total += p.bytes;
	StateHeader hd;
	if (bytes < sizeof(hd)) { set_error("state blob too small"); return B3G_INVALID_ARGUMENT; }
	std::memcpy(&hd, blob, sizeof(hd));
	if (hd.magic != kStateMagic || hd.version != kStateVersion || hd.n != b.n || hd.nTet != b.tet.m || hd.pairs != b.nPairs ||
	    hd.nInst != b.nInst || bytes != total) {
		set_error("state blob does not match this body");
		return B3G_INVALID_ARGUMENT;
	}
	B3G_HIP(hipStreamSynchronize(b.stream));
	++b.xVersion;
	const char* in = static_cast<const char*>(blob);
	size_t off = sizeof(hd);
	for (auto& p : parts) {
		B3G_HIP(hipMemcpy(p.buf->p, in + off, p.bytes, hipMemcpyHostToDevice));
		off += p.bytes;
	}
	B3G_HIP(hipMemcpy(b.stampCounter.p, &hd.stampCounter, 4, hipMemcpyHostToDevice));
	return B3G_OK;
}

b3g_status b3g_synchronize(b3g_body h) {
	CHECK_BODY(h);
	SET_DEVICE(h);
	Body& b = h->body;
	B3G_HIP(hipStreamSynchronize(b.stream));
	if (!b.built || !b.faultWord.p) return B3G_OK;
	uint32_t word = 0;
	B3G_HIP(hipMemcpy(&word, b.faultWord.p, 4, hipMemcpyDeviceToHost));
	if (!word) return B3G_OK;
	const uint32_t zero = 0;
	B3G_HIP(hipMemcpy(b.faultWord.p, &zero, 4, hipMemcpyHostToDevice));   // reported once
	std::vector<SolverState> st(b.nInst);
	B3G_HIP(hipMemcpy(st.data(), b.state.p, st.size() * sizeof(SolverState), hipMemcpyDeviceToHost));
	const uint32_t bit = (word & B3G_FAULT_NON_FINITE) ? (uint32_t)B3G_FAULT_NON_FINITE : (uint32_t)B3G_FAULT_DIAGONAL;
	uint32_t inst = 0;   // instances in creation order, as b3g_get_stats numbers them
	while (inst + 1 < b.nInst && !(st[inst].faults & bit)) ++inst;
	if (bit == B3G_FAULT_NON_FINITE) {
		set_error("a step produced a non-finite position or velocity (instance " + std::to_string(inst) + ")");
		return B3G_NON_FINITE;
	}
	set_error("a Jacobi diagonal entry of pA was not positive (instance " + std::to_string(inst) + ")");
	return B3G_NON_POSITIVE_DIAGONAL;
}

b3g_status b3g_time_spmv(b3g_body h, uint32_t launches, double* mean_ms, double* bytes_per_launch) {
	CHECK_BUILT(h);
	if (!mean_ms || launches == 0) return B3G_INVALID_ARGUMENT;
	SET_DEVICE(h);
	Body& b = h->body;
	b3g_status s = time_pcg_spmv(b, launches, mean_ms);
	if (s != B3G_OK) return s;
	// FP32 blocks: 36 + 4 (column) per block; BF16 operator (pair planes): 18 + 4 per block and a 36-byte FP32
	// diagonal per row
	const bool bf16 = b.solver == B3G_SOLVER_REFINED_BF16 && b.AbP.p;
	if (bytes_per_launch)
		*bytes_per_launch = (bf16 ? 22.0 * b.nnz + 36.0 * b.n : 40.0 * b.nnz) + 4.0 * (b.n + 1) + 24.0 * b.n;
	return B3G_OK;
}

b3g_status b3g_time_spmv_cold(b3g_body h, uint32_t launches, uint64_t evict_bytes, double* mean_ms,
                              double* bytes_per_launch) {
	CHECK_BUILT(h);
	if (!mean_ms || launches == 0) return B3G_INVALID_ARGUMENT;
	if (evict_bytes < (512ull << 20)) { set_error("evict_bytes must be at least 512 MiB"); return B3G_INVALID_ARGUMENT; }
	SET_DEVICE(h);
	Body& b = h->body;
	b3g_status s = time_pcg_spmv_cold(b, launches, (size_t)evict_bytes, mean_ms);
	if (s != B3G_OK) return s;
	const bool bf16 = b.solver == B3G_SOLVER_REFINED_BF16 && b.AbP.p;
	if (bytes_per_launch)
		*bytes_per_launch = (bf16 ? 22.0 * b.nnz + 36.0 * b.n : 40.0 * b.nnz) + 4.0 * (b.n + 1) + 24.0 * b.n;
	return B3G_OK;
}

b3g_status b3g_get_schedule(b3g_body h, b3g_schedule* out) {
	CHECK_BODY(h);
	if (!out) return B3G_INVALID_ARGUMENT;
	const Schedule& sc = h->body.sched;
	*out = {sc.tileBlocks, sc.tileRows, sc.rowBlockRows, sc.fused, sc.maxGroups, sc.persistTiles, sc.persistGroups};
	return B3G_OK;
}

namespace {
static b3g_status check_schedule(const b3g_schedule* sc) {
	if (!sc) { set_error("schedule is NULL"); return B3G_INVALID_ARGUMENT; }
	if (sc->tile_blocks < 32 || sc->tile_blocks > 1024 || sc->tile_rows < 16 || sc->tile_rows > 256 ||
	    sc->row_block_rows < 64 || sc->row_block_rows > 1024 || sc->fused < -1 || sc->fused > 1 ||
	    (sc->max_groups != 0 && sc->max_groups < 256) || sc->persist_tiles > (1u << 20)) {
		set_error("schedule out of range: tile_blocks 32..1024, tile_rows 16..256, row_block_rows 64..1024, fused -1..1, "
		          "max_groups 0 or at least 256, persist_tiles at most 2^20");
		return B3G_INVALID_ARGUMENT;
	}
	return B3G_OK;
}

// FNV-1a over the system's shape, the reduction mode and the device architecture.
static uint64_t schedule_signature(const Body& b) {
	uint64_t h = 1469598103934665603ull;
	auto mix = [&](uint64_t v) {
		for (int k = 0; k < 8; ++k) { h ^= (v >> (8 * k)) & 0xFFu; h *= 1099511628211ull; }
	};
	mix(b.n);
	mix(b.nnz);
	mix(b.nInst);
	mix((uint64_t)(int64_t)b.reduction);
	for (size_t k = 0; k < b.hostRowBegin.size() && k < b.hostRowEnd.size(); ++k) mix(b.hostRowEnd[k] - b.hostRowBegin[k]);
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, b.device) == hipSuccess)
		for (const char* c = prop.gcnArchName; *c; ++c) mix((uint8_t)*c);
	return h;
}
} // namespace

// A change of schedule re-orders pA's storage (the column sort is per tile)
// and the autotuner's timings run the body's own kernels: the assembled pA
// (in CSR slot order), pb, fi, the action forces and the per-instance solver
// states are kept across `change` and put back, pA in the new storage order.
static b3g_status reschedule(Body& b, const std::function<b3g_status()>& change) {
	B3G_HIP(hipStreamSynchronize(b.stream));
	std::vector<SolverState> st(b.nInst);
	if (!st.empty()) B3G_HIP(hipMemcpy(st.data(), b.state.p, st.size() * sizeof(SolverState), hipMemcpyDeviceToHost));
	std::vector<std::pair<DBuf*, std::vector<char>>> kept;
	for (DBuf* d : {&b.pb, &b.fi, &b.act})
		if (d->p && d->bytes) {
			kept.emplace_back(d, std::vector<char>(d->bytes));
			B3G_HIP(hipMemcpy(kept.back().second.data(), d->p, d->bytes, hipMemcpyDeviceToHost));
		}
	std::vector<float> A((size_t)b.aStride * 9), csr((size_t)b.nnz * 9);
	std::vector<uint32_t> pos(b.nnz);
	auto positions = [&]() -> b3g_status {
		if (b.sortedA) B3G_HIP(hipMemcpy(pos.data(), b.slotPos.p, pos.size() * 4, hipMemcpyDeviceToHost));
		else for (uint32_t k = 0; k < b.nnz; ++k) pos[k] = k;
		return B3G_OK;
	};
	b3g_status s;
	if (b.nnz) {
		B3G_HIP(hipMemcpy(A.data(), b.A.p, A.size() * 4, hipMemcpyDeviceToHost));
		if ((s = positions()) != B3G_OK) return s;
		for (size_t k = 0; k < b.nnz; ++k)
			for (int c = 0; c < 9; ++c) csr[(size_t)c * b.nnz + k] = A[(size_t)c * b.aStride + pos[k]];
	}
	if ((s = change()) != B3G_OK) return s;
	B3G_HIP(hipStreamSynchronize(b.stream));
	drop_graph(b);
	if (b.nnz) {
		if ((s = positions()) != B3G_OK) return s;
		for (size_t k = 0; k < b.nnz; ++k)
			for (int c = 0; c < 9; ++c) A[(size_t)c * b.aStride + pos[k]] = csr[(size_t)c * b.nnz + k];
		B3G_HIP(hipMemcpy(b.A.p, A.data(), A.size() * 4, hipMemcpyHostToDevice));
	}
	if (!st.empty()) B3G_HIP(hipMemcpy(b.state.p, st.data(), st.size() * sizeof(SolverState), hipMemcpyHostToDevice));
	for (auto& k : kept) B3G_HIP(hipMemcpy(k.first->p, k.second.data(), k.first->bytes, hipMemcpyHostToDevice));
	return B3G_OK;
}

b3g_status b3g_set_schedule(b3g_body h, const b3g_schedule* sc) {
	CHECK_BUILT(h);
	b3g_status s = check_schedule(sc);
	if (s != B3G_OK) return s;
	SET_DEVICE(h);
	Body& b = h->body;
	return reschedule(b, [&]() -> b3g_status {
		b.sched.tileBlocks = sc->tile_blocks;
		b.sched.tileRows = sc->tile_rows;
		b.sched.rowBlockRows = sc->row_block_rows;
		b.sched.fused = sc->fused;
		b.sched.maxGroups = sc->max_groups;
		b.sched.persistTiles = sc->persist_tiles;
		b.sched.persistGroups = sc->persist_groups;
		return build_tiles(b, true, true);
	});
}

b3g_status b3g_autotune(b3g_body h, const char* cache_path, b3g_schedule* chosen, double* tuned_ms,
                        double* default_ms) {
	CHECK_BUILT(h);
	SET_DEVICE(h);
	Body& b = h->body;
	B3G_HIP(hipStreamSynchronize(b.stream));
	char sig[32];
	std::snprintf(sig, sizeof(sig), "%016llx", (unsigned long long)schedule_signature(b));
	double tuned = -1.0, dflt = -1.0;
	bool hit = false;
	Schedule found;
	if (cache_path) {
		if (FILE* f = std::fopen(cache_path, "r")) {
			char line[512];
			while (std::fgets(line, sizeof(line), f)) {
				char key[32];
				b3g_schedule sc;
				double tm = -1.0, dm = -1.0;
				const int got = std::sscanf(line, "%31s %u %u %u %d %u %u %u %lf %lf", key, &sc.tile_blocks, &sc.tile_rows,
				                            &sc.row_block_rows, &sc.fused, &sc.max_groups, &sc.persist_tiles,
				                            &sc.persist_groups, &tm, &dm);
				if (got >= 8 && std::strcmp(key, sig) == 0 && check_schedule(&sc) == B3G_OK) {
					found.tileBlocks = sc.tile_blocks;
					found.tileRows = sc.tile_rows;
					found.rowBlockRows = sc.row_block_rows;
					found.fused = sc.fused;
					found.maxGroups = sc.max_groups;
					found.persistTiles = sc.persist_tiles;
					found.persistGroups = sc.persist_groups;
					tuned = got >= 9 ? tm : -1.0;
					dflt = got >= 10 ? dm : -1.0;
					hit = true;
				}
			}
			std::fclose(f);
		}
	}
	b3g_status s;
	if (hit) {
		if ((s = reschedule(b, [&]() -> b3g_status { b.sched = found; return build_tiles(b, true, true); })) != B3G_OK) return s;
	} else {
		// after a step: whole steps with the last step's parameters from the
		// current state, which is saved first and loaded before every timing
		// and at the end; before any step: PCG iterations on the system
		std::function<b3g_status(double*)> timer;
		std::vector<char> blob;
		const uint32_t lastConverged = b.lastConverged;
		if (b.lastStep.valid) {
			uint64_t bytes = 0;
			if ((s = b3g_save_state(h, nullptr, &bytes)) != B3G_OK) return s;
			blob.resize(bytes);
			if ((s = b3g_save_state(h, blob.data(), &bytes)) != B3G_OK) return s;
			const auto ls = b.lastStep;
			timer = [&, ls](double* ms) -> b3g_status {
				auto run = [&](uint32_t steps, double* el) -> b3g_status {
					b3g_status r = b3g_load_state(h, blob.data(), blob.size());
					if (r != B3G_OK) return r;
					b.lastConverged = lastConverged;
					const auto t0 = std::chrono::steady_clock::now();
					for (uint32_t k = 0; k < steps; ++k)
						if ((r = b3g_step(h, ls.dt, ls.it, ls.sub, ls.tol, ls.stol)) != B3G_OK) return r;
					if ((r = b3g_synchronize(h)) != B3G_OK) return r;
					*el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
					return B3G_OK;
				};
				double probe = 0.0, el = 0.0;
				b3g_status r = run(1, &probe);   // warm-up (and the capture of a captured step)
				if (r != B3G_OK) return r;
				const uint32_t steps = (uint32_t)std::min(20.0, std::max(2.0, 40.0 / std::max(probe, 1e-3)));
				if ((r = run(steps, &el)) != B3G_OK) return r;
				*ms = el / steps;
				return B3G_OK;
			};
		}
		s = reschedule(b, [&]() -> b3g_status { return autotune(b, timer, &tuned, &dflt); });
		if (!blob.empty()) {
			b3g_status r = b3g_load_state(h, blob.data(), blob.size());
			b.lastConverged = lastConverged;
			if (s == B3G_OK) s = r;
		}
		if (s != B3G_OK) return s;
		if (cache_path) {
			FILE* f = std::fopen(cache_path, "a");
			if (!f) { set_error(std::string("cannot append to the schedule cache ") + cache_path); return B3G_INVALID_ARGUMENT; }
			std::fprintf(f, "%s %u %u %u %d %u %u %u %.6f %.6f # ms per %s; rows %u blocks %u instances %u\n", sig,
			             b.sched.tileBlocks, b.sched.tileRows, b.sched.rowBlockRows, b.sched.fused, b.sched.maxGroups,
			             b.sched.persistTiles, b.sched.persistGroups, tuned, dflt,
			             blob.empty() ? "PCG iteration" : "step", b.n, b.nnz, b.nInst);
			std::fclose(f);
		}
	}
	if (chosen)
		*chosen = {b.sched.tileBlocks, b.sched.tileRows, b.sched.rowBlockRows, b.sched.fused, b.sched.maxGroups,
		           b.sched.persistTiles, b.sched.persistGroups};
	if (tuned_ms) *tuned_ms = tuned;
	if (default_ms) *default_ms = dflt;
	return B3G_OK;
}

b3g_status b3g_set_early_exit(b3g_body h, int enabled) {
	CHECK_BODY(h);
	h->body.earlyExit = enabled != 0;
	return B3G_OK;
}

b3g_status b3g_get_system_size(b3g_body h, uint32_t* rows, uint32_t* blocks) {
	CHECK_BUILT(h);
	if (rows) *rows = h->body.n;
	if (blocks) *blocks = h->body.nnz;
	return B3G_OK;
}

// Pattern of pA (list-order CSR, reference row order) from a description, host only.
b3g_status b3g_pattern(const b3g_body_desc* d, uint32_t* nnz, uint32_t* rowptr, uint32_t* cols, uint32_t capacity) {
	if (!d || !nnz) return B3G_INVALID_ARGUMENT;
	std::vector<uint32_t> rp, cl;
	b3g_status s = host_pattern(*d, rp, cl);
	if (s != B3G_OK) return s;
	*nnz = (uint32_t)cl.size();
	if (rowptr) std::memcpy(rowptr, rp.data(), rp.size() * 4);
	if (cols && capacity >= cl.size()) std::memcpy(cols, cl.data(), cl.size() * 4);
	return B3G_OK;
}

// ----------------------------------------------------------------------------- operator-level entry points
namespace {
// A body that holds only a matrix (for b3g_solve_cg / b3g_spmv).
b3g_status matrix_body(Body& b, int device, uint32_t n, const uint32_t* rowptr, const uint32_t* cols,
                       const float* vals) {
	B3G_HIP(hipSetDevice(device));
	b.device = device;
	B3G_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
	b.n = n;
	b.nnz = rowptr[n];
	b.hostRowptr.assign(rowptr, rowptr + n + 1);
	b.hostCols.assign(cols, cols + b.nnz);
	for (uint32_t i = 0; i < n; ++i) {
		if (rowptr[i + 1] < rowptr[i]) { set_error("rowptr must be non-decreasing"); return B3G_INVALID_ARGUMENT; }
		for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
			if (cols[k] >= n) { set_error("column index out of range"); return B3G_INVALID_ARGUMENT; }
	}
	b.aStride = plane_stride(b.nnz);
	std::vector<float> planes((size_t)b.aStride * 9, 0.0f);
	for (size_t k = 0; k < b.nnz; ++k)
		for (int c = 0; c < 9; ++c) planes[(size_t)c * b.aStride + k] = vals[9 * k + c];
	std::vector<uint32_t> colsPad(b.aStride, 0u);
	std::copy(cols, cols + b.nnz, colsPad.begin());
	b3g_status s;
	if ((s = dupload(b, b.A, planes.data(), planes.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, b.rowptr, rowptr, (n + 1) * 4)) != B3G_OK) return s;
	if ((s = dupload(b, b.cols, colsPad.data(), colsPad.size() * 4)) != B3G_OK) return s;
	for (DBuf* buf : {&b.pb, &b.invM, &b.px, &b.r, &b.s, &b.q, &b.d0, &b.d1})
		if ((s = dalloc(b, *buf, (size_t)n * sizeof(V3))) != B3G_OK) return s;
	if ((s = dalloc(b, b.state, sizeof(SolverState))) != B3G_OK) return s;
	return build_tiles(b, false);   // the caller's CSR storage
}

void matrix_body_free(Body& b) {
	(void)hipStreamSynchronize(b.stream);
	dfree_all(b);
	if (b.stream) (void)hipStreamDestroy(b.stream);
}

__global__ void k_set_state(SolverState* st, uint32_t maxIt, float eps2) {
	uint32_t* words = reinterpret_cast<uint32_t*>(st);
	for (size_t k = 0; k < sizeof(SolverState) / 4; ++k) words[k] = 0;
	st->maxIt = maxIt;
	st->eps2_sub = eps2;
	st->newtonActive = 1;
}

__global__ void k_invm(uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* A, uint32_t stride, V3* invM) {
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float axx = 0.0f, ayy = 0.0f, azz = 0.0f;
	for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)   // b3SparseMat33::operator() const: first match from the head
		if (cols[k] == i) { axx = A[k]; ayy = A[(size_t)4 * stride + k]; azz = A[(size_t)8 * stride + k]; break; }
	invM[i] = v3(1.0f / axx, 1.0f / ayy, 1.0f / azz);
}
} // namespace

b3g_status b3g_solve_cg(int device, uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals,
                        const float* b_in, float* x, uint32_t max_iterations, float tolerance, uint32_t* iterations) {
	if (!rowptr || !cols || !vals || !b_in || !x || n == 0) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	if (!(tolerance > 0.0f && tolerance < 1.0f)) { set_error("tolerance must be in (0, 1)"); return B3G_INVALID_ARGUMENT; }
	Body b;
	b3g_status s = matrix_body(b, device, n, rowptr, cols, vals);
	if (s == B3G_OK) {
		auto body = [&]() -> b3g_status {
			B3G_HIP(hipMemcpyAsync(b.pb.p, b_in, n * sizeof(V3), hipMemcpyHostToDevice, b.stream));
			B3G_HIP(hipMemcpyAsync(b.px.p, x, n * sizeof(V3), hipMemcpyHostToDevice, b.stream));
			hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, b.stream, dptr<SolverState>(b.state), max_iterations,
			                   tolerance * tolerance);
			hipLaunchKernelGGL(k_invm, dim3((n + 255) / 256), dim3(256), 0, b.stream, n, dptr<uint32_t>(b.rowptr),
			                   dptr<uint32_t>(b.cols), dptr<float>(b.A), b.aStride, dptr<V3>(b.invM));
			B3G_HIP(hipGetLastError());
			b3g_status s2 = enqueue_solve_cg(b, max_iterations, tolerance);
			if (s2 != B3G_OK) return s2;
			B3G_HIP(hipMemcpyAsync(x, b.px.p, n * sizeof(V3), hipMemcpyDeviceToHost, b.stream));
			SolverState st;
			B3G_HIP(hipMemcpyAsync(&st, b.state.p, sizeof(st), hipMemcpyDeviceToHost, b.stream));
			B3G_HIP(hipStreamSynchronize(b.stream));
			if (iterations) *iterations = st.it;
			return B3G_OK;
		};
		s = body();
	}
	matrix_body_free(b);
	return s;
}

b3g_status b3g_spmv(int device, uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals,
                    const float* x, float* y) {
	if (!rowptr || !cols || !vals || !x || !y || n == 0) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	Body b;
	b3g_status s = matrix_body(b, device, n, rowptr, cols, vals);
	if (s == B3G_OK) {
		auto body = [&]() -> b3g_status {
			B3G_HIP(hipMemcpyAsync(b.px.p, x, n * sizeof(V3), hipMemcpyHostToDevice, b.stream));
			b3g_status s2 = enqueue_spmv(b, b.px, b.q);
			if (s2 != B3G_OK) return s2;
			B3G_HIP(hipMemcpyAsync(y, b.q.p, n * sizeof(V3), hipMemcpyDeviceToHost, b.stream));
			B3G_HIP(hipStreamSynchronize(b.stream));
			return B3G_OK;
		};
		s = body();
	}
	matrix_body_free(b);
	return s;
}

} // extern "C"
