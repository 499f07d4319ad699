/*
 * TEST INFRASTRUCTURE. Harness over the *unmodified* reference library
 * (compiled from /root/reference/src by oracle/Makefile into oracle/_ref/).
 * It builds a b3Body from a b3s_scene through the reference's own API,
 * steps it and reads its state, so that goldens and the CPU baseline come
 * from the real reference code.
 *
 * Two harness-only extensions need access to internals (reference members
 * are private); this translation unit alone widens access after the
 * standard headers are in:
 *   - fast set-up: links fixtures directly and calls ResetMass once, instead
 *     of the O(N^2) duplicate scans of CreateSphere/Triangle/Tetrahedron
 *     (body.cpp:151-308); the resulting body is identical;
 *   - refh_step_tol: the b3Body::Step sequence (body.cpp:552-604) with the
 *     Newton/PCG tolerances exposed, since b3ForceSolver::Solve keeps the
 *     B3_EPSILON defaults (force_solver.cpp:142-155).
 *
 * The same file builds both the reference as shipped (scalar = float,
 * _ref/libb3ref.so) and the reference with scalar = double
 * (_ref/libb3ref64.so, settings_f64.h forced in): arrays crossing this
 * boundary are float either way, converted to and from scalar, and
 * refh_get_state64 returns the state unrounded.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <new>
#include <vector>
#include <algorithm>
#include <cstdio>
#include <fcntl.h>
#include <unistd.h>

#define private public
#define protected public
#include <bounce_softbody/bounce_softbody.h>
#include <bounce_softbody/dynamics/contacts/sphere_shape_contact.h>
#include <bounce_softbody/dynamics/time_step.h>
#include <bounce_softbody/sparse/sparse_force_solver.h>
#include <bounce_softbody/sparse/sparse_solver.h>
#include <bounce_softbody/sparse/sparse_mat33.h>
#include <bounce_softbody/sparse/dense_vec3.h>
#include <bounce_softbody/sparse/diag_mat33.h>
#undef private
#undef protected

#include "scene.h"

extern uint32 b3_forceSolverIterations;
extern uint32 b3_forceSolverMinSubIterations;
extern uint32 b3_forceSolverMaxSubIterations;

namespace {

// A b3Mesh with its own vertex and triangle storage (the mesh holds raw pointers).
struct HarnessMesh {
	std::vector<b3Vec3> v;
	std::vector<b3Triangle> t;
	b3Mesh mesh;
};

struct RefBody {
	b3Body body;
	std::vector<HarnessMesh*> meshes;   // deleted after the body (its shapes point at them)
	std::vector<b3Particle*> particles;   // creation order
	std::vector<b3Force*> tets;           // tet element forces, creation order
	std::vector<b3SDF*> sdfs;
	uint32_t nworld = 0;
	std::vector<b3WorldFixture*> worlds;  // creation order
	std::vector<b3TriangleFixture*> triangles;   // creation order (a duplicate maps to the fixture it repeats)
	std::vector<b3Force*> byType[7];             // forces of each type, index = position in the type's table
	// captured linear system (list order)
	std::vector<uint32_t> capRowptr, capCols;
	std::vector<float> capVals, capB, capF;
};

b3Vec3 V(const float* p) { return b3Vec3(p[0], p[1], p[2]); }

// The force model of force_solver.cpp:55-86, rebuilt here because that class is file-local.
struct HarnessModel : public b3SparseForceModel {
	std::vector<b3Particle*>* ps;
	std::vector<b3Force*>* fs;
	std::vector<b3SphereAndShapeContact*>* cs;
	void ApplyForces(const b3SparseForceSolverData* data) {
		for (b3Particle* p : *ps) p->ApplyForces(data);
		for (b3Force* f : *fs) f->ApplyForces(data);
		for (b3SphereAndShapeContact* c : *cs) c->ApplyForces(data);
	}
};

struct Lists {
	std::vector<b3Particle*> ps;
	std::vector<b3Force*> fs;
	std::vector<b3SphereAndShapeContact*> cs;
};

Lists Gather(b3Body& body) {   // body.cpp:533-546 (solverId = list position)
	Lists l;
	for (b3Particle* p = body.m_particleList; p; p = p->m_next) { p->m_solverId = (uint32)l.ps.size(); l.ps.push_back(p); }
	for (b3Force* f = body.m_forceList; f; f = f->m_next) l.fs.push_back(f);
	for (b3SphereAndShapeContact* c = body.m_contactManager.m_contactList; c; c = c->m_next) l.cs.push_back(c);
	return l;
}

} // namespace

extern "C" {

void* refh_create(const b3s_scene* sc, int fast) {
	RefBody* rb = new RefBody();
	b3Body& body = rb->body;
	for (uint32_t c = 0; c < sc->n_particles; ++c) {
		b3ParticleDef pd;
		pd.type = (b3ParticleType)sc->p_type[c];
		pd.position = V(sc->p_pos + 3 * c);
		pd.velocity = V(sc->p_vel + 3 * c);
		pd.damping = sc->p_damping[c];
		pd.userIndex = c;
		rb->particles.push_back(body.CreateParticle(pd));
	}
	for (uint32_t c = 0; c < sc->n_spheres; ++c) {
		b3SphereFixtureDef sd;
		sd.p = rb->particles[sc->s_particle[c]];
		sd.radius = sc->s_radius[c];
		sd.friction = sc->s_friction[c];
		sd.userIndex = c;
		if (!fast) { body.CreateSphere(sd); continue; }
		bool dup = false;   // fast path keeps the first fixture of a particle, as CreateSphere does
		for (b3SphereFixture* s = body.m_sphereList; s && !dup && c < 64; s = s->m_next) dup = s->m_p == sd.p;
		if (dup) continue;
		void* mem = body.m_blockAllocator.Allocate(sizeof(b3SphereFixture));
		b3SphereFixture* s = new (mem) b3SphereFixture(sd, &body);
		s->m_prev = nullptr; s->m_next = body.m_sphereList;
		if (body.m_sphereList) body.m_sphereList->m_prev = s;
		body.m_sphereList = s; ++body.m_sphereCount;
	}
	for (uint32_t c = 0; c < sc->n_trifix; ++c) {
		b3TriangleFixtureDef td;
		td.p1 = rb->particles[sc->tf_p[3 * c]];
		td.p2 = rb->particles[sc->tf_p[3 * c + 1]];
		td.p3 = rb->particles[sc->tf_p[3 * c + 2]];
		td.v1 = V(sc->tf_v + 9 * c); td.v2 = V(sc->tf_v + 9 * c + 3); td.v3 = V(sc->tf_v + 9 * c + 6);
		td.density = sc->tf_density[c]; td.radius = sc->tf_radius[c]; td.friction = sc->tf_friction[c];
		td.userIndex = c;
		if (!fast) { rb->triangles.push_back(body.CreateTriangle(td)); continue; }
		void* mem = body.m_blockAllocator.Allocate(sizeof(b3TriangleFixture));
		b3TriangleFixture* t = new (mem) b3TriangleFixture(td, &body);
		rb->triangles.push_back(t);
		t->m_proxyId = body.m_tree.CreateProxy(t->ComputeAABB(), t);
		t->m_prev = nullptr; t->m_next = body.m_triangleList;
		if (body.m_triangleList) body.m_triangleList->m_prev = t;
		body.m_triangleList = t; ++body.m_triangleCount;
	}
	for (uint32_t c = 0; c < sc->n_tetfix; ++c) {
		b3TetrahedronFixtureDef td;
		td.p1 = rb->particles[sc->ttf_p[4 * c]];
		td.p2 = rb->particles[sc->ttf_p[4 * c + 1]];
		td.p3 = rb->particles[sc->ttf_p[4 * c + 2]];
		td.p4 = rb->particles[sc->ttf_p[4 * c + 3]];
		const float* v = sc->ttf_v + 12 * c;
		td.v1 = V(v); td.v2 = V(v + 3); td.v3 = V(v + 6); td.v4 = V(v + 9);
		td.density = sc->ttf_density[c]; td.radius = sc->ttf_radius[c]; td.friction = sc->ttf_friction[c];
		td.userIndex = c;
		if (!fast) { body.CreateTetrahedron(td); continue; }
		void* mem = body.m_blockAllocator.Allocate(sizeof(b3TetrahedronFixture));
		b3TetrahedronFixture* t = new (mem) b3TetrahedronFixture(td, &body);
		t->m_prev = nullptr; t->m_next = body.m_tetrahedronList;
		if (body.m_tetrahedronList) body.m_tetrahedronList->m_prev = t;
		body.m_tetrahedronList = t; ++body.m_tetrahedronCount;
	}
	if (fast) body.ResetMass();
	auto keep = [&](uint32_t type, uint32_t k, b3Force* f) {
		if (rb->byType[type].size() <= k) rb->byType[type].resize(k + 1, nullptr);
		rb->byType[type][k] = f;
	};
	for (uint32_t c = 0; c < sc->n_forces; ++c) {
		uint32_t k = sc->f_index[c];
		switch (sc->f_type[c]) {
		case B3S_FORCE_TET_ELEMENT: {
			b3TetrahedronElementForceDef d;
			d.p1 = rb->particles[sc->tet_p[4 * k]]; d.p2 = rb->particles[sc->tet_p[4 * k + 1]];
			d.p3 = rb->particles[sc->tet_p[4 * k + 2]]; d.p4 = rb->particles[sc->tet_p[4 * k + 3]];
			const float* v = sc->tet_v + 12 * k;
			d.v1 = V(v); d.v2 = V(v + 3); d.v3 = V(v + 6); d.v4 = V(v + 9);
			d.youngModulus = sc->tet_param[3 * k]; d.poissonRatio = sc->tet_param[3 * k + 1];
			d.stiffnessDamping = sc->tet_param[3 * k + 2];
			d.userIndex = c;
			rb->tets.push_back(body.CreateForce(d));
			keep(B3S_FORCE_TET_ELEMENT, k, rb->tets.back());
			break;
		}
		case B3S_FORCE_TRI_ELEMENT: {
			b3TriangleElementForceDef d;
			d.p1 = rb->particles[sc->tri_p[3 * k]]; d.p2 = rb->particles[sc->tri_p[3 * k + 1]];
			d.p3 = rb->particles[sc->tri_p[3 * k + 2]];
			const float* v = sc->tri_v + 9 * k;
			d.v1 = V(v); d.v2 = V(v + 3); d.v3 = V(v + 6);
			const float* q = sc->tri_param + 6 * k;
			d.youngModulusX = q[0]; d.youngModulusY = q[1]; d.shearModulus = q[2];
			d.poissonRationXY = q[3]; d.poissonRationYX = q[4]; d.stiffnessDamping = q[5];
			d.userIndex = c;
			keep(B3S_FORCE_TRI_ELEMENT, k, body.CreateForce(d));
			break;
		}
		case B3S_FORCE_STRETCH: {
			b3StretchForceDef d;
			d.p1 = rb->particles[sc->st_p[3 * k]]; d.p2 = rb->particles[sc->st_p[3 * k + 1]];
			d.p3 = rb->particles[sc->st_p[3 * k + 2]];
			const float* q = sc->st_param + 12 * k;
			d.u1 = q[0]; d.v1 = q[1]; d.u2 = q[2]; d.v2 = q[3]; d.u3 = q[4]; d.v3 = q[5];
			d.stiffness_u = q[6]; d.damping_stiffness_u = q[7]; d.b_u = q[8];
			d.stiffness_v = q[9]; d.damping_stiffness_v = q[10]; d.b_v = q[11];
			d.userIndex = c;
			keep(B3S_FORCE_STRETCH, k, body.CreateForce(d));
			break;
		}
		case B3S_FORCE_SHEAR: {
			b3ShearForceDef d;
			d.p1 = rb->particles[sc->sh_p[3 * k]]; d.p2 = rb->particles[sc->sh_p[3 * k + 1]];
			d.p3 = rb->particles[sc->sh_p[3 * k + 2]];
			const float* q = sc->sh_param + 8 * k;
			d.u1 = q[0]; d.v1 = q[1]; d.u2 = q[2]; d.v2 = q[3]; d.u3 = q[4]; d.v3 = q[5];
			d.stiffness = q[6]; d.dampingStiffness = q[7];
			d.userIndex = c;
			keep(B3S_FORCE_SHEAR, k, body.CreateForce(d));
			break;
		}
		case B3S_FORCE_SPRING: {
			b3SpringForceDef d;
			d.p1 = rb->particles[sc->sp_p[2 * k]]; d.p2 = rb->particles[sc->sp_p[2 * k + 1]];
			const float* q = sc->sp_param + 3 * k;
			d.length = q[0]; d.stiffness = q[1]; d.dampingStiffness = q[2];
			d.userIndex = c;
			keep(B3S_FORCE_SPRING, k, body.CreateForce(d));
			break;
		}
		case B3S_FORCE_MOUSE: {
			b3MouseForceDef d;
			d.p1 = rb->particles[sc->mo_p[4 * k]]; d.p2 = rb->particles[sc->mo_p[4 * k + 1]];
			d.p3 = rb->particles[sc->mo_p[4 * k + 2]]; d.p4 = rb->particles[sc->mo_p[4 * k + 3]];
			const float* q = sc->mo_param + 6 * k;
			d.w2 = q[0]; d.w3 = q[1]; d.w4 = q[2]; d.length = q[3]; d.stiffness = q[4]; d.dampingStiffness = q[5];
			d.userIndex = c;
			keep(B3S_FORCE_MOUSE, k, body.CreateForce(d));
			break;
		}
		default: break;
		}
	}
	for (uint32_t c = 0; c < sc->n_world; ++c) {
		const float* q = sc->w_param + B3S_WORLD_PARAMS * c;
		b3WorldFixtureDef wd;
		wd.friction = sc->w_friction[c];
		b3SphereShape sphere; b3CapsuleShape capsule; b3TriangleShape tri; b3BoxShape box; b3SDFShape sdf;
		b3MeshShape meshShape;
		switch (sc->w_type[c]) {
		case B3S_SHAPE_SPHERE: sphere.m_radius = q[0]; sphere.m_center = V(q + 1); wd.shape = &sphere; break;
		case B3S_SHAPE_CAPSULE: capsule.m_radius = q[0]; capsule.m_center1 = V(q + 1); capsule.m_center2 = V(q + 4); wd.shape = &capsule; break;
		case B3S_SHAPE_TRIANGLE:
			tri.Set(V(q + 1), V(q + 4), V(q + 7));
			tri.m_radius = q[0];
			tri.m_hasE1Vertex = q[10] != 0.0f; tri.m_hasE2Vertex = q[11] != 0.0f; tri.m_hasE3Vertex = q[12] != 0.0f;
			tri.m_e1Vertex = V(q + 13); tri.m_e2Vertex = V(q + 16); tri.m_e3Vertex = V(q + 19);
			wd.shape = &tri;
			break;
		case B3S_SHAPE_BOX:
			box.m_radius = q[0]; box.m_extents = V(q + 1);
			box.m_xf.translation = V(q + 4);
			box.m_xf.rotation.v = V(q + 7); box.m_xf.rotation.s = q[10];
			wd.shape = &box;
			break;
		case B3S_SHAPE_SDF: {
			b3SDF* s = new b3SDF();
			b3AABB aabb;
			aabb.lowerBound = V(q + 1); aabb.upperBound = V(q + 4);
			uint32 W = (uint32)q[7], H = (uint32)q[8], D = (uint32)q[9];
			s->voxelGrid.Create(aabb, W, H, D);
			for (uint32 i = 0; i < W * H * D; ++i) s->voxelGrid.SetVoxel(i, sc->w_voxels[c][i]);
			rb->sdfs.push_back(s);
			sdf.m_radius = q[0]; sdf.m_sdf = s;
			sdf.m_xf.translation = V(q + 10);
			sdf.m_xf.rotation.v = V(q + 13); sdf.m_xf.rotation.s = q[16];
			sdf.m_invert = q[17] != 0.0f;
			wd.shape = &sdf;
			break;
		}
		case B3S_SHAPE_MESH: {   // the scene carries the wing vertices; the tree is the reference's own
			HarnessMesh* hm = new HarnessMesh();
			const uint32_t nv = sc->w_mesh_counts[2 * c], nt = sc->w_mesh_counts[2 * c + 1];
			for (uint32_t i = 0; i < nv; ++i) hm->v.push_back(V(sc->w_mesh_vertices[c] + 3 * i));
			for (uint32_t i = 0; i < nt; ++i) {
				const uint32_t* t = sc->w_mesh_triangles[c] + 6 * (size_t)i;
				b3Triangle tr;
				tr.v1 = t[0]; tr.v2 = t[1]; tr.v3 = t[2]; tr.u1 = t[3]; tr.u2 = t[4]; tr.u3 = t[5];
				hm->t.push_back(tr);
			}
			hm->mesh.vertexCount = nv;
			hm->mesh.vertices = hm->v.data();
			hm->mesh.triangleCount = nt;
			hm->mesh.triangles = hm->t.data();
			hm->mesh.BuildTree();
			rb->meshes.push_back(hm);
			meshShape.m_radius = q[0];
			meshShape.m_mesh = &hm->mesh;
			meshShape.m_scale = V(q + 1);
			meshShape.m_xf.translation = V(q + 4);
			meshShape.m_xf.rotation.v = V(q + 7); meshShape.m_xf.rotation.s = q[10];
			wd.shape = &meshShape;
			break;
		}
		default: break;
		}
		rb->worlds.push_back(body.CreateFixture(wd));
	}
	rb->nworld = sc->n_world;
	body.SetGravity(b3Vec3(sc->gravity[0], sc->gravity[1], sc->gravity[2]));
	return rb;
}

void refh_destroy(void* p) {
	RefBody* rb = (RefBody*)p;
	std::vector<b3SDF*> sdfs = rb->sdfs;
	std::vector<HarnessMesh*> meshes = rb->meshes;
	delete rb;
	for (b3SDF* s : sdfs) delete s;
	for (HarnessMesh* m : meshes) delete m;
}

/* The public entry point: b3Body::Step (body.cpp:552-604). */
void refh_step(void* p, float dt, uint32_t it, uint32_t sub) { ((RefBody*)p)->body.Step(dt, it, sub); }

/*
 * b3Body::Step with explicit tolerances: the same sequence as body.cpp:552-604,
 * body_solver.cpp:67-93 and force_solver.cpp:88-180, with
 * b3SolveBEInput::tolerance / subTolerance set from the arguments.
 */
void refh_step_tol(void* p, float dt, uint32_t iters, uint32_t subIters, float tol, float subTol) {
	b3Body& body = ((RefBody*)p)->body;
	body.m_contactManager.UpdateContacts();
	for (b3Force* f = body.m_forceList; f; f = f->m_next) f->ClearForces();
	if (dt > 0.0f) {
		Lists l = Gather(body);
		uint32 n = (uint32)l.ps.size();
		b3DenseVec3 x0(n), v0(n), fe(n), y(n), x(n), v(n), z(n);
		b3DiagMat33 M(n), S(n);
		b3Vec3 gravity = body.m_gravity;
		for (uint32 i = 0; i < n; ++i) {
			b3Particle* q = l.ps[i];
			x0[i] = q->m_position; v0[i] = q->m_velocity; fe[i] = q->m_force; y[i] = q->m_translation;
			z[i].SetZero();
			if (q->m_type == e_dynamicParticle) {
				M[i] = b3Mat33Diagonal(q->m_mass);
				fe[i] += q->m_mass * gravity;
				S[i].SetIdentity();
			} else {
				M[i] = b3Mat33Diagonal(scalar(1));
				S[i].SetZero();
			}
		}
		HarnessModel model;
		model.ps = &l.ps; model.fs = &l.fs; model.cs = &l.cs;
		b3SolveBEInput in;
		in.forceModel = &model;
		in.h = dt;
		in.inv_h = scalar(1) / scalar(dt);
		in.dofCount = n;
		in.x0 = &x0; in.v0 = &v0; in.fe = &fe; in.M = &M; in.y = &y; in.S = &S; in.z = &z;
		in.maxIterations = iters;
		in.maxSubIterations = subIters;
		if (tol > 0.0f) in.tolerance = tol;
		if (subTol > 0.0f) in.subTolerance = subTol;
		b3SolveBEOutput out;
		out.x = &x; out.v = &v;
		out.minSubIterations = b3_forceSolverMinSubIterations;
		out.maxSubIterations = b3_forceSolverMaxSubIterations;
		b3SparseSolveBE(&out, &in);
		b3_forceSolverIterations = out.iterations;
		b3_forceSolverMinSubIterations = out.minSubIterations;
		b3_forceSolverMaxSubIterations = out.maxSubIterations;
		for (uint32 i = 0; i < n; ++i) { l.ps[i]->m_position = x[i]; l.ps[i]->m_velocity = v[i]; }
		b3TimeStep step;
		step.dt = dt; step.inv_dt = scalar(1) / scalar(dt); step.forceIterations = iters; step.forceSubIterations = subIters;
		for (b3SphereAndShapeContact* c : l.cs) c->ApplyFriction(step, gravity);
	}
	for (b3Particle* q = body.m_particleList; q; q = q->m_next) { q->m_force.SetZero(); q->m_translation.SetZero(); }
	for (b3TriangleFixture* t = body.m_triangleList; t; t = t->m_next) {
		b3Vec3 vel = (t->m_p1->m_velocity + t->m_p2->m_velocity + t->m_p3->m_velocity) / scalar(3);
		t->Synchronize(dt * vel);
	}
	body.m_contactManager.FindNewContacts();
}

double refh_step_timed(void* p, uint32_t steps, float dt, uint32_t it, uint32_t sub, float tol, float subTol) {
	auto t0 = std::chrono::steady_clock::now();
	for (uint32_t s = 0; s < steps; ++s) {
		if (tol > 0.0f || subTol > 0.0f) refh_step_tol(p, dt, it, sub, tol, subTol);
		else refh_step(p, dt, it, sub);
	}
	auto t1 = std::chrono::steady_clock::now();
	return std::chrono::duration<double>(t1 - t0).count();
}

void refh_get_state(void* p, float* x, float* v) {
	RefBody* rb = (RefBody*)p;
	for (size_t c = 0; c < rb->particles.size(); ++c) {
		b3Vec3 a = rb->particles[c]->GetPosition(), b = rb->particles[c]->GetVelocity();
		if (x) { x[3 * c] = a.x; x[3 * c + 1] = a.y; x[3 * c + 2] = a.z; }
		if (v) { v[3 * c] = b.x; v[3 * c + 1] = b.y; v[3 * c + 2] = b.z; }
	}
}

/* The same in double: the FP64 build's state without rounding. */
void refh_get_state64(void* p, double* x, double* v) {
	RefBody* rb = (RefBody*)p;
	for (size_t c = 0; c < rb->particles.size(); ++c) {
		b3Vec3 a = rb->particles[c]->GetPosition(), b = rb->particles[c]->GetVelocity();
		if (x) { x[3 * c] = a.x; x[3 * c + 1] = a.y; x[3 * c + 2] = a.z; }
		if (v) { v[3 * c] = b.x; v[3 * c + 1] = b.y; v[3 * c + 2] = b.z; }
	}
}

void refh_set_state(void* p, const float* x, const float* v) {
	RefBody* rb = (RefBody*)p;
	for (size_t c = 0; c < rb->particles.size(); ++c) {
		if (x) rb->particles[c]->m_position = b3Vec3(x[3 * c], x[3 * c + 1], x[3 * c + 2]);
		if (v) rb->particles[c]->m_velocity = b3Vec3(v[3 * c], v[3 * c + 1], v[3 * c + 2]);
	}
}

void refh_apply_external(void* p, const float* force, const float* translation) {
	RefBody* rb = (RefBody*)p;
	for (size_t c = 0; c < rb->particles.size(); ++c) {
		if (force) rb->particles[c]->ApplyForce(b3Vec3(force[3 * c], force[3 * c + 1], force[3 * c + 2]));
		if (translation) rb->particles[c]->ApplyTranslation(b3Vec3(translation[3 * c], translation[3 * c + 1], translation[3 * c + 2]));
	}
}

void refh_get_masses(void* p, float* m) {
	RefBody* rb = (RefBody*)p;
	for (size_t c = 0; c < rb->particles.size(); ++c) m[c] = rb->particles[c]->GetMass();
}

void refh_get_stats(void* p, uint32_t* out3) {
	(void)p;
	out3[0] = b3_forceSolverIterations;
	out3[1] = b3_forceSolverMinSubIterations;
	out3[2] = b3_forceSolverMaxSubIterations;
}

void refh_reset_stats(void) {
	b3_forceSolverIterations = 0;
	b3_forceSolverMinSubIterations = B3_MAX_U32;
	b3_forceSolverMaxSubIterations = 0;
}

double refh_get_energy(void* p) { return ((RefBody*)p)->body.GetEnergy(); }

uint32_t refh_get_contacts(void* p, uint32_t* particle, uint32_t* world, float* normal_force, uint32_t cap) {
	RefBody* rb = (RefBody*)p;
	uint32_t k = 0;
	for (b3SphereAndShapeContact* c = rb->body.m_contactManager.m_contactList; c; c = c->m_next) {
		if (k < cap) {
			particle[k] = c->m_fixture1->m_p->m_userIndex;
			uint32_t w = 0;
			while (w < rb->worlds.size() && rb->worlds[w] != c->m_fixture2) ++w;
			world[k] = w;
			if (normal_force) normal_force[k] = c->m_normalForce;
		}
		++k;
	}
	return k;
}

void refh_get_tet_quats(void* p, float* q) {
	RefBody* rb = (RefBody*)p;
	for (size_t k = 0; k < rb->tets.size(); ++k) {
		b3TetrahedronElementForce* t = (b3TetrahedronElementForce*)rb->tets[k];
		q[4 * k] = t->m_q.v.x; q[4 * k + 1] = t->m_q.v.y; q[4 * k + 2] = t->m_q.v.z; q[4 * k + 3] = t->m_q.s;
	}
}

/*
 * First Newton linearisation of a step, with the reference's own operators
 * (sparse_force_solver.cpp:146-174), on the current state; the body is not
 * advanced except for the contact update and the tet warm starts, which
 * ApplyForces writes as the reference does. Returns nnz of pA.
 */
uint32_t refh_assemble(void* p, float dt) {
	RefBody* rb = (RefBody*)p;
	b3Body& body = rb->body;
	body.m_contactManager.UpdateContacts();
	Lists l = Gather(body);
	uint32 n = (uint32)l.ps.size();
	b3DenseVec3 x0(n), v0(n), fe(n), y(n), z(n);
	b3DiagMat33 M(n), S(n);
	for (uint32 i = 0; i < n; ++i) {
		b3Particle* q = l.ps[i];
		x0[i] = q->m_position; v0[i] = q->m_velocity; fe[i] = q->m_force; y[i] = q->m_translation;
		z[i].SetZero();
		if (q->m_type == e_dynamicParticle) {
			M[i] = b3Mat33Diagonal(q->m_mass);
			fe[i] += q->m_mass * body.m_gravity;
			S[i].SetIdentity();
		} else {
			M[i] = b3Mat33Diagonal(scalar(1));
			S[i].SetZero();
		}
	}
	HarnessModel model;
	model.ps = &l.ps; model.fs = &l.fs; model.cs = &l.cs;
	scalar h = dt;
	b3DiagMat33 ST(n);
	b3Transpose(ST, S);
	b3DiagMat33 I(n);
	I.SetIdentity();
	b3DenseVec3 x = x0;
	b3DenseVec3 v = v0;
	b3DenseVec3 fi(n);
	fi.SetZero();
	b3SparseMat33 dfdx(n);
	b3SparseMat33 dfdv(n);
	b3SparseForceSolverData sd;
	sd.h = h; sd.inv_h = scalar(1) / h; sd.x = &x; sd.v = &v; sd.f = &fi; sd.dfdx = &dfdx; sd.dfdv = &dfdv;
	model.ApplyForces(&sd);
	b3SparseMat33 A = M - h * dfdv - (h * h) * dfdx;
	b3DenseVec3 b = M * (v0 - v) + h * (fe + fi) + h * dfdx * (x0 - x + h * v + y);
	b3SparseMat33 pA = S * A * ST + I - S;
	b3DenseVec3 pb = S * (b - A * z);
	rb->capRowptr.assign(1, 0);
	rb->capCols.clear();
	rb->capVals.clear();
	rb->capB.resize(3 * n);
	rb->capF.resize(3 * n);
	for (uint32 i = 0; i < n; ++i) {
		for (b3RowEntry* e = pA.rows[i].head; e; e = e->next) {
			rb->capCols.push_back(e->column);
			const scalar* m = &e->value.x.x;
			for (int t = 0; t < 9; ++t) rb->capVals.push_back(m[t]);
		}
		rb->capRowptr.push_back((uint32_t)rb->capCols.size());
		for (int t = 0; t < 3; ++t) { rb->capB[3 * i + t] = pb[i][t]; rb->capF[3 * i + t] = fi[i][t]; }
	}
	return (uint32_t)rb->capCols.size();
}

void refh_get_assembled(void* p, uint32_t* rowptr, uint32_t* cols, float* vals, float* pb, float* fi) {
	RefBody* rb = (RefBody*)p;
	std::memcpy(rowptr, rb->capRowptr.data(), rb->capRowptr.size() * 4);
	std::memcpy(cols, rb->capCols.data(), rb->capCols.size() * 4);
	std::memcpy(vals, rb->capVals.data(), rb->capVals.size() * 4);
	std::memcpy(pb, rb->capB.data(), rb->capB.size() * 4);
	std::memcpy(fi, rb->capF.data(), rb->capF.size() * 4);
}

/* b3StretchForceDef::Initialize (stretch_force.cpp:36-60). */
void refh_uv_initialize(const float* A, const float* B, const float* C, float* out6) {
	b3StretchForceDef d;
	d.Initialize(V(A), V(B), V(C));
	out6[0] = d.u1; out6[1] = d.v1; out6[2] = d.u2; out6[3] = d.v2; out6[4] = d.u3; out6[5] = d.v3;
}

/* b3SparseSolveCG (sparse_solver.cpp:28-109) on a list-order CSR matrix. */
uint32_t refh_solve_cg(uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals, const float* b,
                       float* x, uint32_t maxIt, float eps) {
	b3SparseMat33 A(n);
	for (uint32_t i = 0; i < n; ++i)
		for (uint32_t e = rowptr[i + 1]; e-- > rowptr[i];) {   // insert tail first so the head ends first
			b3RowEntry* en = (b3RowEntry*)b3Alloc(sizeof(b3RowEntry));
			en->column = cols[e];
			scalar* m = &en->value.x.x;
			for (int t = 0; t < 9; ++t) m[t] = vals[9 * e + t];
			A.rows[i].Insert(en);
		}
	b3DenseVec3 bv(n), xv(n);
	for (uint32_t i = 0; i < n; ++i) { bv[i] = V(b + 3 * i); xv[i] = V(x + 3 * i); }
	b3SolveCGInput in;
	in.A = &A; in.b = &bv; in.maxIterations = maxIt; in.tolerance = eps;
	b3SolveCGOutput out;
	out.x = &xv;
	b3SparseSolveCG(&out, &in);
	for (uint32_t i = 0; i < n; ++i) { x[3 * i] = xv[i].x; x[3 * i + 1] = xv[i].y; x[3 * i + 2] = xv[i].z; }
	return out.iterations;
}

// b3SparseSolveBE (sparse_force_solver.cpp:104-232) itself, with a force
// model that hands the iterate to a callback: the callback fills f (zeroed)
// and adds blocks to the two b3SparseMat33 through refh_matrix_add. M and S
// are 9 n floats (column-major diagonal blocks); z may be null (zero).
// subMinMax: in/out, the output's minSubIterations / maxSubIterations.
typedef void (*refh_model_fn)(void* user, uint32_t n, const float* x, const float* v, float* f, void* dfdx, void* dfdv);

// (*m)(rows[k], cols[k]) += / -= blocks[9 k..] for k in order (sign > 0: +=).
void refh_matrix_add(void* m, uint32_t count, const uint32_t* rows, const uint32_t* cols, const float* blocks, int sign) {
	b3SparseMat33& A = *(b3SparseMat33*)m;
	for (uint32_t k = 0; k < count; ++k) {
		b3Mat33 K;
		scalar* d = &K.x.x;
		for (int t = 0; t < 9; ++t) d[t] = blocks[9 * (size_t)k + t];
		if (sign > 0) A(rows[k], cols[k]) += K;
		else A(rows[k], cols[k]) -= K;
	}
}

namespace {
struct CallbackModel : public b3SparseForceModel {
	refh_model_fn fn;
	void* user;
	uint32_t n;
	std::vector<float> x, v, f;
	void ApplyForces(const b3SparseForceSolverData* data) override {
		for (uint32_t i = 0; i < n; ++i) {
			const b3Vec3 xi = (*data->x)[i], vi = (*data->v)[i];
			x[3 * i] = (float)xi.x; x[3 * i + 1] = (float)xi.y; x[3 * i + 2] = (float)xi.z;
			v[3 * i] = (float)vi.x; v[3 * i + 1] = (float)vi.y; v[3 * i + 2] = (float)vi.z;
		}
		std::fill(f.begin(), f.end(), 0.0f);
		fn(user, n, x.data(), v.data(), f.data(), data->dfdx, data->dfdv);
		for (uint32_t i = 0; i < n; ++i) (*data->f)[i] = V(f.data() + 3 * i);
	}
};
} // namespace

uint32_t refh_solve_be(uint32_t n, float h, const float* x0, const float* v0, const float* fe, const float* M,
                       const float* y, const float* S, const float* z, uint32_t maxIt, float tol, uint32_t maxSub,
                       float subTol, refh_model_fn fn, void* user, float* x, float* v, float* error,
                       uint32_t* subMinMax) {
	b3DenseVec3 X0(n), V0(n), FE(n), Y(n), Z(n), X(n), VV(n);
	b3DiagMat33 MM(n), SS(n);
	for (uint32_t i = 0; i < n; ++i) {
		X0[i] = V(x0 + 3 * i); V0[i] = V(v0 + 3 * i); FE[i] = V(fe + 3 * i); Y[i] = V(y + 3 * i);
		Z[i] = z ? V(z + 3 * i) : b3Vec3(scalar(0), scalar(0), scalar(0));
		scalar* m = &MM[i].x.x;
		scalar* s = &SS[i].x.x;
		for (int t = 0; t < 9; ++t) { m[t] = M[9 * (size_t)i + t]; s[t] = S[9 * (size_t)i + t]; }
	}
	CallbackModel model;
	model.fn = fn; model.user = user; model.n = n;
	model.x.resize(3 * (size_t)n); model.v.resize(3 * (size_t)n); model.f.resize(3 * (size_t)n);
	b3SolveBEInput in;
	in.h = h;
	in.inv_h = h != scalar(0) ? scalar(1) / scalar(h) : scalar(0);
	in.forceModel = &model;
	in.dofCount = n;
	in.x0 = &X0; in.v0 = &V0; in.fe = &FE; in.M = &MM; in.y = &Y; in.S = &SS; in.z = &Z;
	in.maxIterations = maxIt; in.tolerance = tol;
	in.maxSubIterations = maxSub; in.subTolerance = subTol;
	b3SolveBEOutput out;
	out.x = &X; out.v = &VV;
	out.iterations = 0; out.error = scalar(0);
	out.minSubIterations = subMinMax[0];
	out.maxSubIterations = subMinMax[1];
	b3SparseSolveBE(&out, &in);
	for (uint32_t i = 0; i < n; ++i) {
		x[3 * i] = (float)X[i].x; x[3 * i + 1] = (float)X[i].y; x[3 * i + 2] = (float)X[i].z;
		v[3 * i] = (float)VV[i].x; v[3 * i + 1] = (float)VV[i].y; v[3 * i + 2] = (float)VV[i].z;
	}
	*error = (float)out.error;
	subMinMax[0] = out.minSubIterations;
	subMinMax[1] = out.maxSubIterations;
	return out.iterations;
}

// Topology edits on the live reference body (body.cpp:62-354), creation
// indices as the scene numbers them; new objects take the next index.
void refh_destroy_force(void* p, uint32_t type, uint32_t k) {
	RefBody* rb = (RefBody*)p;
	rb->body.DestroyForce(rb->byType[type][k]);
	rb->byType[type][k] = nullptr;
}

uint32_t refh_create_particle(void* p, uint32_t type, const float* pos, const float* vel, float damping) {
	RefBody* rb = (RefBody*)p;
	b3ParticleDef pd;
	pd.type = (b3ParticleType)type;
	pd.position = V(pos);
	pd.velocity = V(vel);
	pd.damping = damping;
	pd.userIndex = (uint32)rb->particles.size();
	rb->particles.push_back(rb->body.CreateParticle(pd));
	return (uint32_t)rb->particles.size() - 1;
}

void refh_create_sphere(void* p, uint32_t particle, float radius, float friction) {
	RefBody* rb = (RefBody*)p;
	b3SphereFixtureDef sd;
	sd.p = rb->particles[particle];
	sd.radius = radius;
	sd.friction = friction;
	rb->body.CreateSphere(sd);
}

uint32_t refh_create_spring(void* p, uint32_t p1, uint32_t p2, float length, float ks, float kd) {
	RefBody* rb = (RefBody*)p;
	b3SpringForceDef d;
	d.p1 = rb->particles[p1];
	d.p2 = rb->particles[p2];
	d.length = length;
	d.stiffness = ks;
	d.dampingStiffness = kd;
	rb->byType[B3S_FORCE_SPRING].push_back(rb->body.CreateForce(d));
	return (uint32_t)rb->byType[B3S_FORCE_SPRING].size() - 1;
}

// Action forces of the forces of one type, table order: spring
// GetActionForce / GetReactionForce (spring_force.h:155-163), stretch and
// shear GetActionForce1..3, mouse GetActionForce1..4. Returns the count.
uint32_t refh_get_action_forces(void* p, uint32_t type, float* out) {
	RefBody* rb = (RefBody*)p;
	const std::vector<b3Force*>& fs = rb->byType[type];
	auto put = [&](size_t slot, const b3Vec3& v) { out[3 * slot] = v.x; out[3 * slot + 1] = v.y; out[3 * slot + 2] = v.z; };
	for (size_t c = 0; c < fs.size(); ++c) {
		if (!fs[c]) {   // destroyed
			const int per = type == B3S_FORCE_SPRING ? 2 : type == B3S_FORCE_MOUSE ? 4 : 3;
			for (int i = 0; i < per; ++i) put(per * c + i, b3Vec3(0.0f, 0.0f, 0.0f));
			continue;
		}
		switch (type) {
		case B3S_FORCE_SPRING: {
			b3SpringForce* f = (b3SpringForce*)fs[c];
			put(2 * c, f->GetActionForce()); put(2 * c + 1, f->GetReactionForce());
			break;
		}
		case B3S_FORCE_STRETCH: {
			b3StretchForce* f = (b3StretchForce*)fs[c];
			put(3 * c, f->GetActionForce1()); put(3 * c + 1, f->GetActionForce2()); put(3 * c + 2, f->GetActionForce3());
			break;
		}
		case B3S_FORCE_SHEAR: {
			b3ShearForce* f = (b3ShearForce*)fs[c];
			put(3 * c, f->GetActionForce1()); put(3 * c + 1, f->GetActionForce2()); put(3 * c + 2, f->GetActionForce3());
			break;
		}
		case B3S_FORCE_MOUSE: {
			b3MouseForce* f = (b3MouseForce*)fs[c];
			put(4 * c, f->GetActionForce1()); put(4 * c + 1, f->GetActionForce2());
			put(4 * c + 2, f->GetActionForce3()); put(4 * c + 3, f->GetActionForce4());
			break;
		}
		default: return 0;
		}
	}
	return (uint32_t)fs.size();
}

// b3Body::RayCastSingle (body.cpp:497-521) for n rays: index = creation
// index of the triangle fixture hit (-1: none), fraction, normal.
void refh_ray_cast(void* p, uint32_t n, const float* p1, const float* p2, int64_t* index, float* fraction, float* normal) {
	RefBody* rb = (RefBody*)p;
	for (uint32_t i = 0; i < n; ++i) {
		b3BodyRayCastSingleOutput out;
		const bool hit = rb->body.RayCastSingle(&out, V(p1 + 3 * i), V(p2 + 3 * i));
		index[i] = -1;
		fraction[i] = 0.0f;
		normal[3 * i] = normal[3 * i + 1] = normal[3 * i + 2] = 0.0f;
		if (!hit) continue;
		for (size_t c = 0; c < rb->triangles.size(); ++c)
			if (rb->triangles[c] == out.triangle) { index[i] = (int64_t)c; break; }
		fraction[i] = out.fraction;
		normal[3 * i] = out.normal.x; normal[3 * i + 1] = out.normal.y; normal[3 * i + 2] = out.normal.z;
	}
}

// b3BuildSDF (sdf_builder.cpp:155-174) on a b3Mesh of the given arrays
// (triangles: `stride` words each, vertex indices first): box lower/upper,
// voxel counts and, when `voxels` holds `capacity` floats, the voxel values
// in the grid's order. Returns the voxel count.
uint64_t refh_build_sdf(const float* vertices, uint32_t nv, const uint32_t* triangles, uint32_t nt, uint32_t stride,
                        const float* cell, float extension, float* box6, uint32_t* dims, float* voxels, uint64_t capacity) {
	HarnessMesh hm;
	for (uint32_t i = 0; i < nv; ++i) hm.v.push_back(V(vertices + 3 * i));
	for (uint32_t i = 0; i < nt; ++i) {
		b3Triangle tr;
		tr.v1 = triangles[(size_t)stride * i]; tr.v2 = triangles[(size_t)stride * i + 1]; tr.v3 = triangles[(size_t)stride * i + 2];
		tr.u1 = tr.u2 = tr.u3 = B3_NULL_VERTEX;
		hm.t.push_back(tr);
	}
	hm.mesh.vertexCount = nv; hm.mesh.vertices = hm.v.data();
	hm.mesh.triangleCount = nt; hm.mesh.triangles = hm.t.data();
	b3SDF sdf;
	// b3BuildSDF logs its progress on stdout (b3Log): keep it quiet
	fflush(stdout);
	const int saved = dup(1), devnull = open("/dev/null", O_WRONLY);
	if (saved >= 0 && devnull >= 0) dup2(devnull, 1);
	b3BuildSDF(&sdf, &hm.mesh, V(cell), extension);
	fflush(stdout);
	if (saved >= 0 && devnull >= 0) dup2(saved, 1);
	if (devnull >= 0) close(devnull);
	if (saved >= 0) close(saved);
	const b3AABB& box = sdf.GetAABB();
	box6[0] = box.lowerBound.x; box6[1] = box.lowerBound.y; box6[2] = box.lowerBound.z;
	box6[3] = box.upperBound.x; box6[4] = box.upperBound.y; box6[5] = box.upperBound.z;
	dims[0] = sdf.voxelGrid.GetWidth(); dims[1] = sdf.voxelGrid.GetHeight(); dims[2] = sdf.voxelGrid.GetDepth();
	const uint64_t count = (uint64_t)dims[0] * dims[1] * dims[2];
	if (voxels && capacity >= count)
		for (uint64_t i = 0; i < count; ++i) voxels[i] = sdf.voxelGrid.GetVoxel((uint32)i);
	return count;
}

} // extern "C"
