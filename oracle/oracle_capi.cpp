/*
 * TEST INFRASTRUCTURE (oracle). C entry points over the CPU restatement
 * (cpu_ref.hpp), loaded by tests/, smoke() and bench.py's CPU baseline leg.
 * Arrays crossing this boundary are in creation order unless stated.
 */
#include "cpu_ref.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

using namespace b3o;

namespace {

struct Handle {
	int precision;
	std::unique_ptr<Body<float>> f;
	std::unique_ptr<Body<double>> d;
	// last assembly capture (list order)
	SpMat<float> capA{0};
	DVec<float> capB, capF;
	SpMat<double> capA64{0};
	DVec<double> capB64, capF64;
	// creation index of each tet / list index
	std::vector<uint32_t> tetCreationToList;
};

template <class T, class O> void GetState(Body<T>& b, O* x, O* v) {
	const uint32_t n = (uint32_t)b.particles.size();
	for (uint32_t c = 0; c < n; ++c) {
		const Particle<T>& p = b.particles[n - 1 - c];
		if (x) { x[3 * c] = O(p.x.x); x[3 * c + 1] = O(p.x.y); x[3 * c + 2] = O(p.x.z); }
		if (v) { v[3 * c] = O(p.v.x); v[3 * c + 1] = O(p.v.y); v[3 * c + 2] = O(p.v.z); }
	}
}

} // namespace

extern "C" {

void* b3o_create(const b3s_scene* sc, int precision) {
	Handle* h = new Handle();
	h->precision = precision;
	if (precision == 64) { h->d.reset(new Body<double>()); h->d->Build(*sc); }
	else { h->f.reset(new Body<float>()); h->f->Build(*sc); }
	// list index of each tet created k-th: forces are listed in reverse creation
	uint32_t nt = 0;
	for (uint32_t c = 0; c < sc->n_forces; ++c) nt += sc->f_type[c] == B3S_FORCE_TET_ELEMENT;
	h->tetCreationToList.resize(nt);
	for (uint32_t k = 0; k < nt; ++k) h->tetCreationToList[k] = nt - 1 - k;
	return h;
}

void b3o_destroy(void* p) { delete (Handle*)p; }

/* A new handle holding a copy of the body and its whole dynamic state. */
void* b3o_clone(void* p) {
	const Handle* h = (const Handle*)p;
	Handle* c = new Handle();
	c->precision = h->precision;
	if (h->f) c->f.reset(new Body<float>(*h->f));
	if (h->d) c->d.reset(new Body<double>(*h->d));
	c->tetCreationToList = h->tetCreationToList;
	return c;
}

/* 1: the solve's dot products and Newton error accumulated in double (VDotWide); 0: as the reference. */
void b3o_set_wide_dots(void* p, int wide) {
	Handle* h = (Handle*)p;
	if (h->f) h->f->wideDots = wide != 0;
	if (h->d) h->d->wideDots = wide != 0;
}

void b3o_step(void* p, float dt, uint32_t it, uint32_t sub, float tol, float subTol) {
	Handle* h = (Handle*)p;
	if (h->precision == 64) h->d->Step(double(dt), it, sub, double(tol), double(subTol));
	else h->f->Step(dt, it, sub, tol, subTol);
}

/* Steps `steps` times and returns the wall time in seconds of those steps. */
double b3o_step_timed(void* p, uint32_t steps, float dt, uint32_t it, uint32_t sub, float tol, float subTol) {
	auto t0 = std::chrono::steady_clock::now();
	for (uint32_t s = 0; s < steps; ++s) b3o_step(p, dt, it, sub, tol, subTol);
	auto t1 = std::chrono::steady_clock::now();
	return std::chrono::duration<double>(t1 - t0).count();
}

void b3o_get_state(void* p, float* x, float* v) {
	Handle* h = (Handle*)p;
	if (h->precision == 64) GetState(*h->d, x, v);
	else GetState(*h->f, x, v);
}

void b3o_get_state64(void* p, double* x, double* v) {
	Handle* h = (Handle*)p;
	if (h->precision == 64) GetState(*h->d, x, v);
	else GetState(*h->f, x, v);
}

/* Overwrites positions/velocities (creation order). */
void b3o_set_state(void* p, const float* x, const float* v) {
	Handle* h = (Handle*)p;
	auto set = [&](auto& b) {
		const uint32_t n = (uint32_t)b.particles.size();
		for (uint32_t c = 0; c < n; ++c) {
			auto& q = b.particles[n - 1 - c];
			using T = decltype(q.mass);
			if (x) q.x = MakeV3(T(x[3 * c]), T(x[3 * c + 1]), T(x[3 * c + 2]));
			if (v) q.v = MakeV3(T(v[3 * c]), T(v[3 * c + 1]), T(v[3 * c + 2]));
		}
	};
	if (h->precision == 64) set(*h->d);
	else set(*h->f);
}

/* External force / translation accumulators (b3Particle::ApplyForce / ApplyTranslation), creation order. */
void b3o_apply_external(void* p, const float* force, const float* translation) {
	Handle* h = (Handle*)p;
	auto app = [&](auto& b) {
		const uint32_t n = (uint32_t)b.particles.size();
		for (uint32_t c = 0; c < n; ++c) {
			auto& q = b.particles[n - 1 - c];
			using T = decltype(q.mass);
			if (force && q.type == kDynamic) q.force += MakeV3(T(force[3 * c]), T(force[3 * c + 1]), T(force[3 * c + 2]));
			if (translation) q.translation += MakeV3(T(translation[3 * c]), T(translation[3 * c + 1]), T(translation[3 * c + 2]));
		}
	};
	if (h->precision == 64) app(*h->d);
	else app(*h->f);
}

void b3o_get_masses(void* p, float* m) {
	Handle* h = (Handle*)p;
	if (h->precision == 64) {
		const uint32_t n = (uint32_t)h->d->particles.size();
		for (uint32_t c = 0; c < n; ++c) m[c] = float(h->d->particles[n - 1 - c].mass);
	} else {
		const uint32_t n = (uint32_t)h->f->particles.size();
		for (uint32_t c = 0; c < n; ++c) m[c] = h->f->particles[n - 1 - c].mass;
	}
}

void b3o_get_stats(void* p, uint32_t* out3) {
	Handle* h = (Handle*)p;
	StepStats s = h->precision == 64 ? h->d->stats : h->f->stats;
	out3[0] = s.newtonIterations;
	out3[1] = s.minSub;
	out3[2] = s.maxSub;
}

/* Restarts the running statistics (b3_forceSolver* globals) as a fresh process would hold them. */
void b3o_reset_stats(void* p) {
	Handle* h = (Handle*)p;
	if (h->f) h->f->stats = StepStats();
	if (h->d) h->d->stats = StepStats();
}

double b3o_get_energy(void* p) {
	Handle* h = (Handle*)p;
	return h->precision == 64 ? h->d->GetEnergy() : double(h->f->GetEnergy());
}

/* Contact list in list order: creation index of the sphere fixture's particle and of the world fixture. */
uint32_t b3o_get_contacts(void* p, uint32_t* particle, uint32_t* world, float* normal_force, uint32_t cap) {
	Handle* h = (Handle*)p;
	auto get = [&](auto& b) {
		uint32_t k = 0;
		const uint32_t n = (uint32_t)b.particles.size();
		const uint32_t nw = (uint32_t)b.worlds.size();
		for (const auto& c : b.contacts) {
			if (k < cap) {
				particle[k] = n - 1 - b.spheres[c.sphere].p;
				world[k] = nw - 1 - c.world;
				if (normal_force) normal_force[k] = float(c.normalForce);
			}
			++k;
		}
		return k;
	};
	return h->precision == 64 ? get(*h->d) : get(*h->f);
}

/* Warm-start quaternions (x y z s) of the tetrahedron elements, creation order. */
void b3o_get_tet_quats(void* p, float* q) {
	Handle* h = (Handle*)p;
	auto get = [&](auto& b) {
		const uint32_t nt = (uint32_t)b.tets.size();
		for (uint32_t k = 0; k < nt; ++k) {
			const auto& t = b.tets[nt - 1 - k];
			q[4 * k] = float(t.q.v.x); q[4 * k + 1] = float(t.q.v.y); q[4 * k + 2] = float(t.q.v.z); q[4 * k + 3] = float(t.q.s);
		}
	};
	if (h->precision == 64) get(*h->d);
	else get(*h->f);
}

/*
 * Runs the contact update and the first Newton linearisation of a step on a
 * copy of the body and keeps pA, pb and the internal force fi (list order).
 * Returns the number of blocks of pA.
 */
uint32_t b3o_assemble(void* p, float dt) {
	Handle* h = (Handle*)p;
	uint32_t nnz = 0;
	if (h->precision == 64) {
		Body<double> copy = *h->d;
		copy.capture.want = true;
		copy.capture.pA = &h->capA64;
		copy.capture.pb = &h->capB64;
		copy.capture.fi = &h->capF64;
		copy.Step(double(dt), 1, 0);
		for (auto& r : h->capA64.rows) nnz += (uint32_t)r.size();
	} else {
		Body<float> copy = *h->f;
		copy.capture.want = true;
		copy.capture.pA = &h->capA;
		copy.capture.pb = &h->capB;
		copy.capture.fi = &h->capF;
		copy.Step(dt, 1, 0);
		for (auto& r : h->capA.rows) nnz += (uint32_t)r.size();
	}
	return nnz;
}

/* pA rows in list order, each row head-to-tail (the reference SpMV order); values column-major. */
void b3o_get_assembled(void* p, uint32_t* rowptr, uint32_t* cols, float* vals, float* pb, float* fi) {
	Handle* h = (Handle*)p;
	auto get = [&](auto& A, auto& b, auto& f) {
		uint32_t k = 0;
		rowptr[0] = 0;
		for (uint32_t i = 0; i < A.n(); ++i) {
			const auto& r = A.rows[i];
			for (size_t e = r.size(); e-- > 0;) {
				cols[k] = r[e].col;
				const auto* m = &r[e].val.x.x;
				for (int t = 0; t < 9; ++t) vals[9 * k + t] = float(m[t]);
				++k;
			}
			rowptr[i + 1] = k;
			for (int t = 0; t < 3; ++t) {
				pb[3 * i + t] = float(b[i][t]);
				fi[3 * i + t] = float(f[i][t]);
			}
		}
	};
	if (h->precision == 64) get(h->capA64, h->capB64, h->capF64);
	else get(h->capA, h->capB, h->capF);
}

/* The same in double (FP64 oracle: the assembled system without rounding to float). */
void b3o_get_assembled64(void* p, uint32_t* rowptr, uint32_t* cols, double* vals, double* pb, double* fi) {
	Handle* h = (Handle*)p;
	auto get = [&](auto& A, auto& b, auto& f) {
		uint32_t k = 0;
		rowptr[0] = 0;
		for (uint32_t i = 0; i < A.n(); ++i) {
			const auto& r = A.rows[i];
			for (size_t e = r.size(); e-- > 0;) {
				cols[k] = r[e].col;
				const auto* m = &r[e].val.x.x;
				for (int t = 0; t < 9; ++t) vals[9 * k + t] = double(m[t]);
				++k;
			}
			rowptr[i + 1] = k;
			for (int t = 0; t < 3; ++t) {
				pb[3 * i + t] = double(b[i][t]);
				fi[3 * i + t] = double(f[i][t]);
			}
		}
	};
	if (h->precision == 64) get(h->capA64, h->capB64, h->capF64);
	else get(h->capA, h->capB, h->capF);
}

/* Restated b3StretchForceDef::Initialize (stretch_force.cpp:36-60) for test comparison. */
void b3o_uv_initialize(const float* A, const float* B, const float* C, float* out6) {
	UVInitialize<float>(MakeV3(A[0], A[1], A[2]), MakeV3(B[0], B[1], B[2]), MakeV3(C[0], C[1], C[2]), out6);
}

/*
 * Standalone restatement of b3SparseSolveCG (sparse_solver.cpp:28-109) on a
 * matrix given in list-order CSR. x is the warm start on entry and the result
 * on exit. Returns the iteration count.
 */
uint32_t b3o_solve_cg(uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals, const float* b,
                      float* x, uint32_t maxIt, float eps) {
	SpMat<float> A(n);
	for (uint32_t i = 0; i < n; ++i) {
		// rows arrive head-to-tail; the vector form keeps the head at back()
		for (uint32_t e = rowptr[i + 1]; e-- > rowptr[i];) {
			SpMat<float>::Entry en;
			en.col = cols[e];
			std::memcpy(&en.val.x.x, vals + 9 * e, 9 * sizeof(float));
			A.rows[i].push_back(en);
		}
	}
	DVec<float> bv(n), xv(n);
	for (uint32_t i = 0; i < n; ++i) {
		bv[i] = MakeV3(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
		xv[i] = MakeV3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
	}
	uint32_t it = b3o::SolveCGFree<float>(A, bv, maxIt, eps, xv);
	for (uint32_t i = 0; i < n; ++i) { x[3 * i] = xv[i].x; x[3 * i + 1] = xv[i].y; x[3 * i + 2] = xv[i].z; }
	return it;
}

/* Restated SpMV (sparse_mat33.h:494-512) on list-order CSR. */
void b3o_spmv(uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals, const float* x, float* y) {
	for (uint32_t i = 0; i < n; ++i) {
		V3<float> acc;
		acc.SetZero();
		for (uint32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
			M33<float> a;
			std::memcpy(&a.x.x, vals + 9 * e, 9 * sizeof(float));
			acc += a * MakeV3(x[3 * cols[e]], x[3 * cols[e] + 1], x[3 * cols[e] + 2]);
		}
		y[3 * i] = acc.x; y[3 * i + 1] = acc.y; y[3 * i + 2] = acc.z;
	}
}

// The state a body keeps across topology edits, from `src` into `dst` (built
// from the edited description): per object of dst its creation index in src
// or -1 (new). Particle positions and velocities, tetrahedron warm-start
// quaternions, and the contacts whose sphere and world fixture both carry
// over, in src's list order (the contact list is edited only by removal).
void b3o_carry(void* dst, void* src, const int64_t* particle_prev, const int64_t* tet_prev, const int64_t* sphere_prev,
               const int64_t* world_prev) {
	Body<float>& D = *((Handle*)dst)->f;
	const Body<float>& S = *((Handle*)src)->f;
	const size_t nD = D.particles.size(), nS = S.particles.size();
	for (size_t c = 0; c < nD; ++c) {
		const int64_t q = particle_prev[c];
		if (q < 0) continue;
		D.particles[nD - 1 - c].x = S.particles[nS - 1 - (size_t)q].x;
		D.particles[nD - 1 - c].v = S.particles[nS - 1 - (size_t)q].v;
	}
	const size_t mD = D.tets.size(), mS = S.tets.size();
	for (size_t c = 0; c < mD; ++c)
		if (tet_prev[c] >= 0) D.tets[mD - 1 - c].q = S.tets[mS - 1 - (size_t)tet_prev[c]].q;
	// creation index -> list position, and old -> new creation index
	uint32_t maxIndex = 0;
	for (const auto& sp : D.spheres) maxIndex = std::max(maxIndex, sp.index);
	for (const auto& sp : S.spheres) maxIndex = std::max(maxIndex, sp.index);
	std::vector<int64_t> sphereNew(maxIndex + 1, -1), sphereListD(maxIndex + 1, -1);
	std::vector<int64_t> worldNew(S.worlds.size() + 1, -1);
	for (size_t k = 0; k < D.spheres.size(); ++k) {
		const uint32_t c = D.spheres[k].index;
		sphereListD[c] = (int64_t)k;
		if (sphere_prev[c] >= 0) sphereNew[sphere_prev[c]] = c;
	}
	const size_t wD = D.worlds.size(), wS = S.worlds.size();
	for (size_t c = 0; c < wD; ++c)
		if (world_prev[c] >= 0) worldNew[world_prev[c]] = (int64_t)c;
	D.contacts.clear();
	D.contactKeys.clear();
	for (const Contact<float>& ct : S.contacts) {
		const int64_t sc = sphereNew[S.spheres[ct.sphere].index];
		const int64_t wc = worldNew[wS - 1 - ct.world];
		if (sc < 0 || wc < 0) continue;
		Contact<float> n = ct;
		n.sphere = (uint32_t)sphereListD[sc];
		n.world = (uint32_t)(wD - 1 - (size_t)wc);
		D.contacts.push_back(n);
		D.contactKeys.insert({n.sphere, n.world});
	}
}

// Action forces (m_f1.. of spring / stretch / shear / mouse forces, e.g.
// spring_force.h:155-163) of the forces of one type in creation order: 2, 3,
// 3 or 4 vectors per force. Returns the force count.
uint32_t b3o_get_action_forces(void* p, uint32_t type, float* out) {
	Handle* h = (Handle*)p;
	Body<float>& b = *h->f;
	auto put = [&](const auto& table, int k) {
		const uint32_t m = (uint32_t)table.size();
		for (uint32_t c = 0; c < m; ++c) {   // tables are in list order: reverse creation within the type
			const auto& e = table[m - 1 - c];
			for (int i = 0; i < k; ++i) {
				out[3 * ((size_t)k * c + i)] = e.act[i].x;
				out[3 * ((size_t)k * c + i) + 1] = e.act[i].y;
				out[3 * ((size_t)k * c + i) + 2] = e.act[i].z;
			}
		}
		return m;
	};
	switch (type) {
	case B3S_FORCE_SPRING: return put(b.springs, 2);
	case B3S_FORCE_STRETCH: return put(b.stretches, 3);
	case B3S_FORCE_SHEAR: return put(b.shears, 3);
	case B3S_FORCE_MOUSE: return put(b.mice, 4);
	default: return 0;
	}
}

// n rays p1[i] -> p2[i] against the FP32 body's triangle fixtures:
// index[i] = creation index of the fixture hit (-1: none), fraction, normal.
void b3o_ray_cast(void* p, uint32_t n, const float* p1, const float* p2, int64_t* index, float* fraction, float* normal) {
	Handle* h = (Handle*)p;
	for (uint32_t i = 0; i < n; ++i) {
		float f = 0.0f;
		V3<float> nn = MakeV3(0.0f, 0.0f, 0.0f);
		const V3<float> a = MakeV3(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]);
		const V3<float> b = MakeV3(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]);
		index[i] = h->f->RayCastSingle(a, b, &f, &nn);
		fraction[i] = index[i] >= 0 ? f : 0.0f;
		normal[3 * i] = index[i] >= 0 ? nn.x : 0.0f;
		normal[3 * i + 1] = index[i] >= 0 ? nn.y : 0.0f;
		normal[3 * i + 2] = index[i] >= 0 ? nn.z : 0.0f;
	}
}

// b3BuildSDF (sdf_builder.cpp:155-174): the grid (box lower/upper, voxel
// counts) and, when `voxels` is non-null and holds `capacity` floats, the
// voxel values. Returns the voxel count.
uint64_t b3o_build_sdf(const float* vertices, uint32_t nv, const uint32_t* triangles, uint32_t nt, uint32_t stride,
                       const float* cell, float extension, float* box6, uint32_t* dims, float* voxels, uint64_t capacity) {
	AABB<float> box;
	SdfGrid(vertices, nv, cell, extension, &box, dims);
	box6[0] = box.lo.x; box6[1] = box.lo.y; box6[2] = box.lo.z;
	box6[3] = box.hi.x; box6[4] = box.hi.y; box6[5] = box.hi.z;
	const uint64_t count = (uint64_t)dims[0] * dims[1] * dims[2];
	if (voxels && capacity >= count) BuildSdf(vertices, triangles, nt, stride, box, dims, voxels);
	return count;
}

// b3SparseSolveBE (sparse_force_solver.cpp:104-232) with a caller's force
// model. The model gets the iterate (x, v: 3 n floats), a zeroed f to fill
// and the two matrices as opaque pointers for b3o_matrix_add. M and S are
// 9 n floats (column-major diagonal blocks); z may be null (zero). stats:
// in/out running min / max of the PCG counts. wide: the widened-accumulator
// dot products. Returns the Newton iteration count.
typedef void (*b3o_model_fn)(void* user, uint32_t n, const float* x, const float* v, float* f, void* dfdx, void* dfdv);

// (*m)(rows[k], cols[k]) += blocks[9 k..] (sign > 0) or -= (sign < 0) for k in order, the reference's
// b3SparseMat33::operator() (sparse_mat33.h:268-288) then b3Mat33::operator+= / -=.
void b3o_matrix_add(void* m, uint32_t count, const uint32_t* rows, const uint32_t* cols, const float* blocks, int sign) {
	SpMat<float>& A = *(SpMat<float>*)m;
	for (uint32_t k = 0; k < count; ++k) {
		M33<float> K;
		float* d = &K.x.x;
		for (int t = 0; t < 9; ++t) d[t] = blocks[9 * (size_t)k + t];
		if (sign > 0) A.at(rows[k], cols[k]) += K;
		else A.at(rows[k], cols[k]) -= K;
	}
}

uint32_t b3o_solve_be(uint32_t n, float h, const float* x0, const float* v0, const float* fe, const float* M,
                      const float* y, const float* S, const float* z, uint32_t maxIt, float tol, uint32_t maxSub,
                      float subTol, b3o_model_fn fn, void* user, int wide, float* x, float* v, float* error,
                      uint32_t* subMinMax) {
	auto vec = [n](const float* p) {
		DVec<float> o(n);
		for (uint32_t i = 0; i < n; ++i) o[i] = p ? MakeV3(p[3 * i], p[3 * i + 1], p[3 * i + 2]) : MakeV3(0.0f, 0.0f, 0.0f);
		return o;
	};
	auto mat = [n](const float* p) {
		DMat<float> o(n);
		for (uint32_t i = 0; i < n; ++i) {
			float* d = &o[i].x.x;
			for (int t = 0; t < 9; ++t) d[t] = p[9 * (size_t)i + t];
		}
		return o;
	};
	DVec<float> X0 = vec(x0), V0 = vec(v0), FE = vec(fe), Y = vec(y), Z = vec(z), XO, VO;
	DMat<float> MM = mat(M), SS = mat(S);
	std::vector<float> xb(3 * (size_t)n), vb(3 * (size_t)n), fb(3 * (size_t)n);
	auto model = [&](const DVec<float>& xx, const DVec<float>& vv, DVec<float>& f, SpMat<float>& dfdx, SpMat<float>& dfdv) {
		for (uint32_t i = 0; i < n; ++i) {
			xb[3 * i] = xx[i].x; xb[3 * i + 1] = xx[i].y; xb[3 * i + 2] = xx[i].z;
			vb[3 * i] = vv[i].x; vb[3 * i + 1] = vv[i].y; vb[3 * i + 2] = vv[i].z;
		}
		std::fill(fb.begin(), fb.end(), 0.0f);
		fn(user, n, xb.data(), vb.data(), fb.data(), &dfdx, &dfdv);
		for (uint32_t i = 0; i < n; ++i) f[i] = MakeV3(fb[3 * i], fb[3 * i + 1], fb[3 * i + 2]);
	};
	StepStats st;
	st.minSub = subMinMax[0];
	st.maxSub = subMinMax[1];
	float err = 0.0f;
	const uint32_t it = SolveBEGeneric(h, maxIt, tol, maxSub, subTol, X0, V0, FE, MM, Y, SS, Z, XO, VO, model, wide != 0,
	                                   st, (Capture<float>*)nullptr, &err);
	for (uint32_t i = 0; i < n; ++i) {
		x[3 * i] = XO[i].x; x[3 * i + 1] = XO[i].y; x[3 * i + 2] = XO[i].z;
		v[3 * i] = VO[i].x; v[3 * i + 1] = VO[i].y; v[3 * i + 2] = VO[i].z;
	}
	*error = err;
	subMinMax[0] = st.minSub;
	subMinMax[1] = st.maxSub;
	return it;
}

} // extern "C"
