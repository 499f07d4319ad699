/*
 * TEST INFRASTRUCTURE (oracle). CPU restatement of the reference implicit
 * step b3Body::Step (dynamics/body.cpp:552-604) and everything it calls on
 * the hot path, on flat arrays, templated on the scalar type:
 *   - Body<float>  is the FP32 oracle; it reproduces the reference bit for bit
 *     (pinned against goldens written by the compiled reference, oracle/_ref).
 *   - Body<double> is the FP64 oracle (reference semantics with scalar=double).
 *
 * Nothing in the product links this file. Objects are kept in the reference's
 * list order (reverse creation, body.cpp:67-75), so particle i here is the
 * particle with m_solverId == i (body_solver.cpp:51-55).
 */
#ifndef B3O_CPU_REF_HPP
#define B3O_CPU_REF_HPP

#include "cpu_ref_math.hpp"
#include "cpu_ref_sparse.hpp"
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <list>
#include <set>
#include <utility>
#include <vector>

namespace b3o {

enum { kStatic = 0, kKinematic = 1, kDynamic = 2 };

template <class T> struct Particle {
	uint8_t type;
	V3<T> x, v, force, translation;
	T mass, invMass, damping;
};

template <class T> struct TetForce {          // tetrahedron_element_force.h:125-157
	uint32_t p[4];
	V3<T> x[4];
	M33<T> invE;
	T E, nu, kd;
	M33<T> K[16];
	Q4<T> q;
};
template <class T> struct TriForce {          // triangle_element_force.h
	uint32_t p[3];
	V3<T> v[3];
	V2<T> x[3];
	M22<T> invS;
	T Ex, Ey, Es, nuxy, nuyx, kd;
	M22<T> K[9];
};
template <class T> struct StretchForce {      // stretch_force.h
	uint32_t p[3];
	T alpha, du1, dv1, du2, dv2, inv_det;
	V3<T> dwudx, dwvdx;
	T ks_u, kd_u, b_u, ks_v, kd_v, b_v;
	V3<T> act[3];                             // m_f1..m_f3
};
template <class T> struct ShearForce {        // shear_force.h
	uint32_t p[3];
	T alpha, du1, dv1, du2, dv2, inv_det;
	V3<T> dwudx, dwvdx;
	T ks, kd;
	V3<T> act[3];                             // m_f1..m_f3: accumulated over a step's ApplyForces calls
};
template <class T> struct SpringForce {       // spring_force.h
	uint32_t p[2];
	T L0, ks, kd;
	V3<T> act[2];                             // m_f1, m_f2
};
template <class T> struct MouseForce {        // mouse_force.h
	uint32_t p[4];
	T w2, w3, w4, ks, kd, L0;
	V3<T> act[4];                             // m_f1..m_f4
};

template <class T> struct SphereFixture { uint32_t p; T radius, friction; uint32_t index; };   // index: creation order
template <class T> struct TriFixture { uint32_t p[3]; T area, density; uint32_t index; };   // index: creation order
template <class T> struct TetFixture { uint32_t p[4]; T volume, density; };

template <class T> struct WorldFixture {
	uint32_t type;
	T radius, friction;
	V3<T> a, b, c;                // sphere center / capsule centers / triangle vertices / box extents
	bool hasE[3];
	V3<T> e[3];
	Xf<T> xf;                     // box and sdf transform
	AABB<T> gridBox;              // sdf voxel grid aabb
	uint32_t W, H, D;             // sdf voxel counts
	bool invert;
	std::vector<T> vox;
	// mesh: b3MeshShape (mesh_shape.h:30-50) over a b3Mesh (mesh.h:52-87)
	V3<T> scale;
	std::vector<V3<T>> meshV;
	std::vector<uint32_t> meshT;  // 6 per triangle: v1 v2 v3 u1 u2 u3
};

template <class T> struct Manifold { V3<T> point, normal; };

template <class T> struct Contact {   // sphere_shape_contact.h:30-53
	uint32_t sphere, world;           // indices into the sphere / world fixture lists
	Manifold<T> manifold;
	T normalForce;
	bool applyFriction;
};

struct StepStats {
	uint32_t newtonIterations = 0;
	uint32_t minSub = 0xFFFFFFFFu;    // force_solver.cpp:33-34 (running min / max)
	uint32_t maxSub = 0;
};

// Per-element hooks to expose intermediate state (used by parity tests).
template <class T> struct Capture {
	bool want = false;
	SpMat<T>* pA = nullptr;
	DVec<T>* pb = nullptr;
	DVec<T>* fi = nullptr;
};

template <class T> class Body {
public:
	std::vector<Particle<T>> particles;
	std::vector<std::pair<uint32_t, uint32_t>> forces;   // (type, index) in list order
	std::vector<TetForce<T>> tets;
	std::vector<TriForce<T>> tris;
	std::vector<StretchForce<T>> stretches;
	std::vector<ShearForce<T>> shears;
	std::vector<SpringForce<T>> springs;
	std::vector<MouseForce<T>> mice;
	std::vector<SphereFixture<T>> spheres;     // list order
	std::vector<TriFixture<T>> triFixtures;
	std::vector<TetFixture<T>> tetFixtures;
	std::vector<WorldFixture<T>> worlds;       // list order
	std::list<Contact<T>> contacts;            // list order (front = head)
	std::set<std::pair<uint32_t, uint32_t>> contactKeys;
	V3<T> gravity;
	StepStats stats;
	Capture<T> capture;
	// false: b3Dot / b3LengthSquared as the reference accumulates them (in T);
	// true: accumulated in double (VDotWide), the device's FP64 reduction mode
	bool wideDots = false;

	// Build from a creation-ordered scene (mimics the b3Body::Create* calls).
	void Build(const b3s_scene& sc);

	// body.cpp:552-604 with optional tolerances (tol <= 0 -> B3_EPSILON as in
	// sparse_force_solver.h:55-61).
	void Step(T dt, uint32_t iters, uint32_t subIters, T tol = T(-1), T subTol = T(-1));

	// body.cpp:394-466
	void ResetMass();

	T GetEnergy() const {  // body.cpp:384-392 (list order)
		T E = T(0);
		for (const auto& p : particles) E += p.mass * Dot(p.v, p.v);
		return T(0.5) * E;
	}

	// Force model callback: force_solver.cpp:70-86
	void ApplyForces(T h, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);

	bool Collide(const WorldFixture<T>& w, V3<T> center, T radius, Manifold<T>* m) const;
	// b3Body::RayCastSingle (body.cpp:497-521) by brute force over the
	// triangle fixtures: the smallest hit fraction; among equal fractions the
	// lowest creation index. Returns the fixture's creation index or -1.
	int64_t RayCastSingle(V3<T> p1, V3<T> p2, T* fraction, V3<T>* normal) const;
	AABB<T> WorldAABB(const WorldFixture<T>& w) const;

private:
	void ApplyTet(TetForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplyTri(TriForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplyStretch(StretchForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplyShear(ShearForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplySpring(SpringForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplyMouse(MouseForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv);
	void ApplyContact(Contact<T>& c, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx);
	void Solve(T dt, uint32_t iters, uint32_t subIters, T tol, T subTol);
	void SolveBE(T h, uint32_t maxIt, T eps, uint32_t maxSub, T subEps,
	             const DVec<T>& x0, const DVec<T>& v0, const DVec<T>& fe, const DMat<T>& M,
	             const DVec<T>& y, const DMat<T>& S, const DVec<T>& z, DVec<T>& xo, DVec<T>& vo);
	uint32_t SolveCG(const SpMat<T>& A, const DVec<T>& b, uint32_t maxIt, T eps, DVec<T>& x);
	void UpdateContacts();
	void FindNewContacts();
	void ApplyFriction(Contact<T>& c, T dt);
	AABB<T> SphereAABB(const SphereFixture<T>& s) const {  // sphere_fixture.cpp:32-35
		return AABB<T>::FromCenterRadius(particles[s.p].x, s.radius);
	}
};

// ======================================================================
// Element set-up (definitions -> solver data)
// ======================================================================

// tetrahedron_element_force.cpp:37-57
template <class T> void ComputeD(T out[36], T E, T nu) {
	T lambda = (nu * E) / ((T(1) + nu) * (T(1) - T(2) * nu));
	T mu = E / (T(2) * (T(1) + nu));
	T D[36] = {lambda + 2 * mu, lambda, lambda, 0, 0, 0,
	           lambda, lambda + 2 * mu, lambda, 0, 0, 0,
	           lambda, lambda, lambda + 2 * mu, 0, 0, 0,
	           0, 0, 0, mu, 0, 0,
	           0, 0, 0, 0, mu, 0,
	           0, 0, 0, 0, 0, mu};
	for (int i = 0; i < 36; ++i) out[i] = D[i];
}

// tetrahedron_element_force.cpp:121-159
template <class T> void ComputeB(T out[72], const M33<T>& invE) {
	T dN2dx = invE.x.x, dN3dx = invE.x.y, dN4dx = invE.x.z;
	T dN1dx = -dN2dx - dN3dx - dN4dx;
	T dN2dy = invE.y.x, dN3dy = invE.y.y, dN4dy = invE.y.z;
	T dN1dy = -dN2dy - dN3dy - dN4dy;
	T dN2dz = invE.z.x, dN3dz = invE.z.y, dN4dz = invE.z.z;
	T dN1dz = -dN2dz - dN3dz - dN4dz;
	T B[72] = {dN1dx, 0, 0, dN1dy, 0, dN1dz,
	           0, dN1dy, 0, dN1dx, dN1dz, 0,
	           0, 0, dN1dz, 0, dN1dy, dN1dx,
	           dN2dx, 0, 0, dN2dy, 0, dN2dz,
	           0, dN2dy, 0, dN2dx, dN2dz, 0,
	           0, 0, dN2dz, 0, dN2dy, dN2dx,
	           dN3dx, 0, 0, dN3dy, 0, dN3dz,
	           0, dN3dy, 0, dN3dx, dN3dz, 0,
	           0, 0, dN3dz, 0, dN3dy, dN3dx,
	           dN4dx, 0, 0, dN4dy, 0, dN4dz,
	           0, dN4dy, 0, dN4dx, dN4dz, 0,
	           0, 0, dN4dz, 0, dN4dy, dN4dx};
	for (int i = 0; i < 72; ++i) out[i] = B[i];
}

// tetrahedron_element_force.cpp:219-265 (ResetElementData)
template <class T> void TetReset(TetForce<T>& t) {
	V3<T> e1 = t.x[1] - t.x[0];
	V3<T> e2 = t.x[2] - t.x[0];
	V3<T> e3 = t.x[3] - t.x[0];
	M33<T> E = {e1, e2, e3};
	T det = Det(e1, e2, e3);
	T V = Abs(det) / T(6);
	det = T(1) / det;
	t.invE = det * Adjugate(E);
	T D[36], B[72], BT[72], BT_D[72], BT_D_B[144];
	ComputeD(D, t.E, t.nu);
	ComputeB(B, t.invE);
	GTranspose(BT, B, 6, 12);
	GMul(BT_D, BT, 12, 6, D, 6, 6);
	GMul(BT_D_B, BT_D, 12, 6, B, 6, 12);
	for (int i = 0; i < 144; ++i) BT_D_B[i] *= V;
	// b3SetK (tetrahedron_element_force.cpp:163-192)
	for (int i = 0; i < 12; ++i)
		for (int j = 0; j < 12; ++j) t.K[(i / 3) + 4 * (j / 3)].at(i % 3, j % 3) = BT_D_B[i + 12 * j];
}

// triangle_element_force.cpp:30-50
template <class T> M33<T> ComputeC(T Ex, T Ey, T Es, T nu_xy, T nu_yx) {
	T s = T(1) - nu_xy * nu_yx;
	M33<T> C;
	C.x.x = Ex / s;
	C.x.y = Ey * nu_xy / s;
	C.x.z = T(0);
	C.y.x = Ex * nu_yx / s;
	C.y.y = Ey / s;
	C.y.z = T(0);
	C.z.x = T(0);
	C.z.y = T(0);
	C.z.z = Es;
	return C;
}

// triangle_element_force.cpp:208-279
template <class T> void TriReset(TriForce<T>& t) {
	V3<T> p1 = t.v[0], p2 = t.v[1], p3 = t.v[2];
	V3<T> n = Cross(p2 - p1, p3 - p1);
	n.Normalize();
	V3<T> px, py;
	ComputeBasis(n, px, py);
	T P[6] = {px.x, py.x, px.y, py.y, px.z, py.z};
	V2<T> x1, x2, x3;
	GMul(&x1.x, P, 2, 3, &p1.x, 3, 1);
	GMul(&x2.x, P, 2, 3, &p2.x, 3, 1);
	GMul(&x3.x, P, 2, 3, &p3.x, 3, 1);
	t.x[0] = x1;
	t.x[1] = x2;
	t.x[2] = x3;
	V2<T> e1 = x2 - x1, e2 = x3 - x1;
	M22<T> S = {e1, e2};
	T det = Det(e1, e2);
	T A = T(0.5) * Abs(det);
	t.invS = det * Adjugate(S);   // quirk: scaled by det, not 1/det (triangle_element_force.cpp:252)
	M33<T> C = ComputeC(t.Ex, t.Ey, t.Es, t.nuxy, t.nuyx);
	T dN2dx = t.invS.x.x, dN3dx = t.invS.x.y, dN1dx = -dN2dx - dN3dx;
	T dN2dy = t.invS.y.x, dN3dy = t.invS.y.y, dN1dy = -dN2dy - dN3dy;
	T B[18] = {dN1dx, 0, dN1dy, 0, dN1dy, dN1dx,
	           dN2dx, 0, dN2dy, 0, dN2dy, dN2dx,
	           dN3dx, 0, dN3dy, 0, dN3dy, dN3dx};
	T BT[18], BT_C[18], K[36];
	GTranspose(BT, B, 3, 6);
	GMul(BT_C, BT, 6, 3, &C.x.x, 3, 3);
	GMul(K, BT_C, 6, 3, B, 3, 6);
	for (int i = 0; i < 36; ++i) K[i] *= A;
	for (int i = 0; i < 6; ++i)
		for (int j = 0; j < 6; ++j) t.K[(i / 2) + 3 * (j / 2)].at(i % 2, j % 2) = K[i + 6 * j];
}

// stretch_force.cpp:62-107 and shear_force.cpp:51-92 (same (u,v) set-up)
template <class T, class F> void UVReset(F& f, T u1, T v1, T u2, T v2, T u3, T v3) {
	T du1 = u2 - u1, dv1 = v2 - v1, du2 = u3 - u1, dv2 = v3 - v1;
	f.du1 = du1; f.dv1 = dv1; f.du2 = du2; f.dv2 = dv2;
	T det = du1 * dv2 - du2 * dv1;
	f.inv_det = T(1) / det;
	f.dwudx.x = f.inv_det * (dv1 - dv2);
	f.dwudx.y = f.inv_det * dv2;
	f.dwudx.z = -f.inv_det * dv1;
	f.dwvdx.x = f.inv_det * (du2 - du1);
	f.dwvdx.y = -f.inv_det * du2;
	f.dwvdx.z = f.inv_det * du1;
	f.alpha = T(0.5) * Abs(det);
}

// stretch_force.cpp:36-60 (b3StretchForceDef::Initialize, identical in shear_force.cpp:25-49)
template <class T> void UVInitialize(V3<T> A, V3<T> B, V3<T> C, T out[6]) {
	V3<T> AB = B - A, AC = C - A;
	T u1 = T(0), v1 = T(0);
	T u2 = Length(AB), v2 = T(0);
	V3<T> n_AB = AB / u2;
	T a2 = Length(Cross(AB, AC));
	T u3 = Dot(AC, n_AB);
	T v3 = a2 / u2;
	out[0] = u1; out[1] = v1; out[2] = u2; out[3] = v2; out[4] = u3; out[5] = v3;
}

// ======================================================================
// Build
// ======================================================================
template <class T> static V3<T> LoadV3(const float* p) { return MakeV3(T(p[0]), T(p[1]), T(p[2])); }

template <class T> void Body<T>::Build(const b3s_scene& sc) {
	const uint32_t n = sc.n_particles;
	auto L = [n](uint32_t c) { return n - 1 - c; };  // creation index -> list index
	particles.assign(n, Particle<T>());
	for (uint32_t c = 0; c < n; ++c) {
		Particle<T>& p = particles[L(c)];
		p.type = sc.p_type[c];
		p.x = LoadV3<T>(sc.p_pos + 3 * c);
		p.v = LoadV3<T>(sc.p_vel + 3 * c);
		p.force.SetZero();
		p.translation.SetZero();
		p.damping = T(sc.p_damping[c]);
		// particle.cpp:40-49
		p.mass = p.type == kDynamic ? T(1) : T(0);
		p.invMass = p.mass;
	}
	gravity = MakeV3(T(sc.gravity[0]), T(sc.gravity[1]), T(sc.gravity[2]));

	// Sphere fixtures: body.cpp:151-176 (an existing fixture on the particle wins)
	{
		std::vector<SphereFixture<T>> created;
		std::vector<char> has(n, 0);
		for (uint32_t c = 0; c < sc.n_spheres; ++c) {
			uint32_t p = L(sc.s_particle[c]);
			if (has[p]) continue;
			has[p] = 1;
			created.push_back({p, T(sc.s_radius[c]), T(sc.s_friction[c]), c});
		}
		spheres.assign(created.rbegin(), created.rend());
	}
	// Triangle fixtures: body.cpp:205-242 + triangle_fixture.cpp:24-41
	{
		std::vector<TriFixture<T>> created;
		std::set<std::vector<uint32_t>> seen;
		for (uint32_t c = 0; c < sc.n_trifix; ++c) {
			std::vector<uint32_t> key = {sc.tf_p[3 * c], sc.tf_p[3 * c + 1], sc.tf_p[3 * c + 2]};
			std::sort(key.begin(), key.end());
			if (!seen.insert(key).second) continue;
			TriFixture<T> t;
			for (int k = 0; k < 3; ++k) t.p[k] = L(sc.tf_p[3 * c + k]);
			V3<T> A = LoadV3<T>(sc.tf_v + 9 * c), B = LoadV3<T>(sc.tf_v + 9 * c + 3), C = LoadV3<T>(sc.tf_v + 9 * c + 6);
			V3<T> N = Cross(B - A, C - A);
			t.area = T(0.5) * Length(N);
			t.density = T(sc.tf_density[c]);
			t.index = c;
			created.push_back(t);
		}
		triFixtures.assign(created.rbegin(), created.rend());
	}
	// Tetrahedron fixtures: body.cpp:274-308 + tetrahedron_fixture.cpp:23-49
	{
		std::vector<TetFixture<T>> created;
		std::set<std::vector<uint32_t>> seen;
		for (uint32_t c = 0; c < sc.n_tetfix; ++c) {
			std::vector<uint32_t> key = {sc.ttf_p[4 * c], sc.ttf_p[4 * c + 1], sc.ttf_p[4 * c + 2], sc.ttf_p[4 * c + 3]};
			std::sort(key.begin(), key.end());
			if (!seen.insert(key).second) continue;
			TetFixture<T> t;
			for (int k = 0; k < 4; ++k) t.p[k] = L(sc.ttf_p[4 * c + k]);
			const float* v = sc.ttf_v + 12 * c;
			V3<T> A = LoadV3<T>(v), B = LoadV3<T>(v + 3), C = LoadV3<T>(v + 6), D = LoadV3<T>(v + 9);
			V3<T> E1 = B - A, E2 = C - A, E3 = D - A;
			const T inv6 = T(1) / T(6);
			T det = Det(E1, E2, E3);
			T sign = Sign(det);
			t.volume = inv6 * sign * det;
			t.density = T(sc.ttf_density[c]);
			created.push_back(t);
		}
		tetFixtures.assign(created.rbegin(), created.rend());
	}
	// Forces in list order: body.cpp:109-125 (head insertion)
	forces.clear();
	for (uint32_t c = sc.n_forces; c-- > 0;) {
		uint32_t type = sc.f_type[c], k = sc.f_index[c];
		switch (type) {
		case B3S_FORCE_TET_ELEMENT: {
			TetForce<T> t;
			for (int a = 0; a < 4; ++a) { t.p[a] = L(sc.tet_p[4 * k + a]); t.x[a] = LoadV3<T>(sc.tet_v + 12 * k + 3 * a); }
			t.E = T(sc.tet_param[3 * k]); t.nu = T(sc.tet_param[3 * k + 1]); t.kd = T(sc.tet_param[3 * k + 2]);
			t.q.SetIdentity();
			TetReset(t);
			forces.push_back({type, (uint32_t)tets.size()});
			tets.push_back(t);
			break;
		}
		case B3S_FORCE_TRI_ELEMENT: {
			TriForce<T> t;
			for (int a = 0; a < 3; ++a) { t.p[a] = L(sc.tri_p[3 * k + a]); t.v[a] = LoadV3<T>(sc.tri_v + 9 * k + 3 * a); }
			const float* q = sc.tri_param + 6 * k;
			t.Ex = T(q[0]); t.Ey = T(q[1]); t.Es = T(q[2]); t.nuxy = T(q[3]); t.nuyx = T(q[4]); t.kd = T(q[5]);
			TriReset(t);
			forces.push_back({type, (uint32_t)tris.size()});
			tris.push_back(t);
			break;
		}
		case B3S_FORCE_STRETCH: {
			StretchForce<T> s;
			for (int a = 0; a < 3; ++a) s.p[a] = L(sc.st_p[3 * k + a]);
			const float* q = sc.st_param + 12 * k;
			UVReset<T>(s, T(q[0]), T(q[1]), T(q[2]), T(q[3]), T(q[4]), T(q[5]));
			s.ks_u = T(q[6]); s.kd_u = T(q[7]); s.b_u = T(q[8]);
			s.ks_v = T(q[9]); s.kd_v = T(q[10]); s.b_v = T(q[11]);
			forces.push_back({type, (uint32_t)stretches.size()});
			stretches.push_back(s);
			break;
		}
		case B3S_FORCE_SHEAR: {
			ShearForce<T> s;
			for (int a = 0; a < 3; ++a) s.p[a] = L(sc.sh_p[3 * k + a]);
			const float* q = sc.sh_param + 8 * k;
			UVReset<T>(s, T(q[0]), T(q[1]), T(q[2]), T(q[3]), T(q[4]), T(q[5]));
			s.ks = T(q[6]); s.kd = T(q[7]);
			forces.push_back({type, (uint32_t)shears.size()});
			shears.push_back(s);
			break;
		}
		case B3S_FORCE_SPRING: {
			SpringForce<T> s;
			for (int a = 0; a < 2; ++a) s.p[a] = L(sc.sp_p[2 * k + a]);
			const float* q = sc.sp_param + 3 * k;
			s.L0 = T(q[0]); s.ks = T(q[1]); s.kd = T(q[2]);
			forces.push_back({type, (uint32_t)springs.size()});
			springs.push_back(s);
			break;
		}
		case B3S_FORCE_MOUSE: {
			MouseForce<T> s;
			for (int a = 0; a < 4; ++a) s.p[a] = L(sc.mo_p[4 * k + a]);
			const float* q = sc.mo_param + 6 * k;
			s.w2 = T(q[0]); s.w3 = T(q[1]); s.w4 = T(q[2]); s.L0 = T(q[3]); s.ks = T(q[4]); s.kd = T(q[5]);
			forces.push_back({type, (uint32_t)mice.size()});
			mice.push_back(s);
			break;
		}
		default: break;
		}
	}
	for (auto& e : stretches) for (auto& a : e.act) a.SetZero();
	for (auto& e : shears) for (auto& a : e.act) a.SetZero();
	for (auto& e : springs) for (auto& a : e.act) a.SetZero();
	for (auto& e : mice) for (auto& a : e.act) a.SetZero();
	// World fixtures: body.cpp:337-354 (head insertion), world_fixture.cpp:39-44
	worlds.clear();
	for (uint32_t c = sc.n_world; c-- > 0;) {
		const float* q = sc.w_param + B3S_WORLD_PARAMS * c;
		WorldFixture<T> w;
		w.type = sc.w_type[c];
		w.friction = T(sc.w_friction[c]);
		w.radius = T(q[0]);
		w.invert = false;
		w.hasE[0] = w.hasE[1] = w.hasE[2] = false;
		w.W = w.H = w.D = 0;
		switch (w.type) {
		case B3S_SHAPE_SPHERE: w.a = LoadV3<T>(q + 1); break;
		case B3S_SHAPE_CAPSULE: w.a = LoadV3<T>(q + 1); w.b = LoadV3<T>(q + 4); break;
		case B3S_SHAPE_TRIANGLE:
			w.a = LoadV3<T>(q + 1); w.b = LoadV3<T>(q + 4); w.c = LoadV3<T>(q + 7);
			for (int k = 0; k < 3; ++k) { w.hasE[k] = q[10 + k] != 0.0f; w.e[k] = LoadV3<T>(q + 13 + 3 * k); }
			break;
		case B3S_SHAPE_BOX:
			w.a = LoadV3<T>(q + 1);
			w.xf.p = LoadV3<T>(q + 4);
			w.xf.q.v = LoadV3<T>(q + 7); w.xf.q.s = T(q[10]);
			break;
		case B3S_SHAPE_SDF: {
			w.gridBox.lo = LoadV3<T>(q + 1); w.gridBox.hi = LoadV3<T>(q + 4);
			w.W = (uint32_t)q[7]; w.H = (uint32_t)q[8]; w.D = (uint32_t)q[9];
			w.xf.p = LoadV3<T>(q + 10);
			w.xf.q.v = LoadV3<T>(q + 13); w.xf.q.s = T(q[16]);
			w.invert = q[17] != 0.0f;
			size_t nv = (size_t)w.W * w.H * w.D;
			w.vox.resize(nv);
			for (size_t i = 0; i < nv; ++i) w.vox[i] = T(sc.w_voxels[c][i]);
			break;
		}
		case B3S_SHAPE_MESH: {
			w.scale = LoadV3<T>(q + 1);
			w.xf.p = LoadV3<T>(q + 4);
			w.xf.q.v = LoadV3<T>(q + 7); w.xf.q.s = T(q[10]);
			const uint32_t nv = sc.w_mesh_counts[2 * c], nt = sc.w_mesh_counts[2 * c + 1];
			for (uint32_t i = 0; i < nv; ++i) w.meshV.push_back(LoadV3<T>(sc.w_mesh_vertices[c] + 3 * i));
			w.meshT.assign(sc.w_mesh_triangles[c], sc.w_mesh_triangles[c] + 6 * (size_t)nt);
			break;
		}
		default: break;
		}
		worlds.push_back(w);
	}
	contacts.clear();
	contactKeys.clear();
	ResetMass();
	// A body is built before its first step; pairs exist only after a step
	// (contact_manager.cpp:65-82 runs at the end of b3Body::Step).
}

// body.cpp:394-466
template <class T> void Body<T>::ResetMass() {
	for (const auto& t : triFixtures) for (int k = 0; k < 3; ++k) particles[t.p[k]].mass = T(0);
	for (const auto& t : tetFixtures) for (int k = 0; k < 4; ++k) particles[t.p[k]].mass = T(0);
	const T inv3 = T(1) / T(3);
	for (const auto& t : triFixtures) {
		T mass = t.density * t.area;
		for (int k = 0; k < 3; ++k) particles[t.p[k]].mass += inv3 * mass;
	}
	const T inv4 = T(1) / T(4);
	for (const auto& t : tetFixtures) {
		T mass = t.density * t.volume;
		for (int k = 0; k < 4; ++k) particles[t.p[k]].mass += inv4 * mass;
	}
	for (auto& p : particles) {
		if (p.type == kStatic || p.type == kKinematic) { p.mass = T(0); p.invMass = T(0); continue; }
		if (p.mass > T(0)) p.invMass = T(1) / p.mass;
		else { p.mass = T(1); p.invMass = T(1); }
	}
}

// ======================================================================
// Forces (ApplyForces of each element type)
// ======================================================================

// "A Robust Method to Extract the Rotational Part of Deformations":
// tetrahedron_element_force.cpp:273-311
template <class T> Q4<T> ExtractRotation(const M33<T>& A, const Q4<T>& q0, uint32_t maxIterations = 32) {
	Q4<T> q = q0;
	for (uint32_t it = 0; it < maxIterations; ++it) {
		M33<T> R = q.GetRotationMatrix();
		T den = Abs(Dot(R.x, A.x) + Dot(R.y, A.y) + Dot(R.z, A.z));
		if (den == T(0)) break;
		V3<T> torque = Cross(R.x, A.x) + Cross(R.y, A.y) + Cross(R.z, A.z);
		const T kTol = T(1.0e-9);
		T inv_den = T(1) / den + kTol;
		V3<T> omega = inv_den * torque;
		T w = Length(omega);
		if (w < kTol) break;
		Q4<T> qw;
		qw.SetAxisAngle(omega / w, w);
		q = QMul(qw, q);
		q.Normalize();
	}
	return q;
}

// tetrahedron_element_force.cpp:313-464
template <class T>
void Body<T>::ApplyTet(TetForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	uint32_t i1 = e.p[0], i2 = e.p[1], i3 = e.p[2], i4 = e.p[3];
	V3<T> p1 = x[i1], p2 = x[i2], p3 = x[i3], p4 = x[i4];
	V3<T> v1 = v[i1], v2 = v[i2], v3 = v[i3], v4 = v[i4];
	M33<T> E = {p2 - p1, p3 - p1, p4 - p1};
	M33<T> F = E * e.invE;
	Q4<T> q = ExtractRotation(F, e.q);
	e.q = q;
	M33<T> R = q.GetRotationMatrix();
	M33<T> RT = Transpose(R);
	M33<T> K[16];
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) K[i + 4 * j] = R * e.K[i + 4 * j] * RT;
	uint32_t is[4] = {i1, i2, i3, i4};
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) dfdx.at(is[i], is[j]) -= K[i + 4 * j];
	V3<T> us[4];
	us[0] = RT * p1 - e.x[0];
	us[1] = RT * p2 - e.x[1];
	us[2] = RT * p3 - e.x[2];
	us[3] = RT * p4 - e.x[3];
	V3<T> fs[4];
	for (int i = 0; i < 4; ++i) {
		fs[i].SetZero();
		for (int j = 0; j < 4; ++j) fs[i] += e.K[i + 4 * j] * us[j];
	}
	for (int i = 0; i < 4; ++i) fs[i] = R * fs[i];
	for (int i = 0; i < 4; ++i) f[is[i]] -= fs[i];
	if (e.kd > T(0)) {
		V3<T> vs[4] = {v1, v2, v3, v4};
		V3<T> fds[4];
		for (int i = 0; i < 4; ++i) {
			fds[i].SetZero();
			for (int j = 0; j < 4; ++j) fds[i] += e.kd * K[i + 4 * j] * vs[j];
		}
		for (int i = 0; i < 4; ++i) f[is[i]] -= fds[i];
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) dfdv.at(is[i], is[j]) -= e.kd * K[i + 4 * j];
	}
}

// Shoemake-Duff 2D polar: triangle_element_force.cpp:131-150
template <class T> M22<T> ExtractRotation2(const M22<T>& M) {
	T m11 = M.x.x, m12 = M.y.x;
	T m21 = M.x.y, m22 = M.y.y;
	T det = m11 * m22 - m12 * m21;
	M22<T> A;
	A.x.x = M.y.y;
	A.x.y = -M.y.x;
	A.y.x = -M.x.y;
	A.y.y = M.x.x;
	M22<T> Q = M + Sign(det) * A;
	Q.x.Normalize();
	Q.y.Normalize();
	return Q;
}

// triangle_element_force.cpp:285-468
template <class T>
void Body<T>::ApplyTri(TriForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	uint32_t i1 = e.p[0], i2 = e.p[1], i3 = e.p[2];
	V3<T> p1 = x[i1], p2 = x[i2], p3 = x[i3];
	V3<T> v1 = v[i1], v2 = v[i2], v3 = v[i3];
	V3<T> n = Cross(p2 - p1, p3 - p1);
	n.Normalize();
	V3<T> px, py;
	ComputeBasis(n, px, py);
	T P[6] = {px.x, py.x, px.y, py.y, px.z, py.z};
	T PT[6] = {px.x, px.y, px.z, py.x, py.y, py.z};
	V2<T> x1, x2, x3;
	GMul(&x1.x, P, 2, 3, &p1.x, 3, 1);
	GMul(&x2.x, P, 2, 3, &p2.x, 3, 1);
	GMul(&x3.x, P, 2, 3, &p3.x, 3, 1);
	M22<T> Tm;
	Tm.x = x2 - x1;
	Tm.y = x3 - x1;
	M22<T> F = Tm * e.invS;
	M22<T> R = ExtractRotation2(F);
	M22<T> RT = Transpose(R);
	V2<T> us[3];
	us[0] = RT * x1 - e.x[0];
	us[1] = RT * x2 - e.x[1];
	us[2] = RT * x3 - e.x[2];
	V2<T> fs[3];
	for (int i = 0; i < 3; ++i) {
		fs[i].SetZero();
		for (int j = 0; j < 3; ++j) fs[i] = fs[i] + e.K[i + 3 * j] * us[j];
	}
	for (int i = 0; i < 3; ++i) fs[i] = R * fs[i];
	V3<T> f3[3];
	for (int i = 0; i < 3; ++i) GMul(&f3[i].x, PT, 3, 2, &fs[i].x, 2, 1);
	uint32_t is[3] = {i1, i2, i3};
	for (int i = 0; i < 3; ++i) f[is[i]] -= f3[i];
	M33<T> K[9];
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			M22<T> k = R * e.K[i + 3 * j] * RT;
			T PT_k[6];
			GMul(PT_k, PT, 3, 2, &k.x.x, 2, 2);
			M33<T> Ke;
			GMul(&Ke.x.x, PT_k, 3, 2, P, 2, 3);
			K[i + 3 * j] = Ke;
		}
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) dfdx.at(is[i], is[j]) -= K[i + 3 * j];
	if (e.kd > T(0)) {
		V3<T> vs[3] = {v1, v2, v3};
		V3<T> fds[3];
		for (int i = 0; i < 3; ++i) {
			fds[i].SetZero();
			for (int j = 0; j < 3; ++j) fds[i] += e.kd * K[i + 3 * j] * vs[j];
		}
		for (int i = 0; i < 3; ++i) f[is[i]] -= fds[i];
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j) dfdv.at(is[i], is[j]) -= e.kd * K[i + 3 * j];
	}
}

// Adds a 3x3 pattern of blocks in the reference's order (i-major).
template <class T> static void Add3x3(SpMat<T>& m, const uint32_t is[3], const M33<T> K[3][3]) {
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) m.at(is[i], is[j]) += K[i][j];
}

// stretch_force.cpp:121-389
template <class T>
void Body<T>::ApplyStretch(StretchForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	T alpha = e.alpha, du1 = e.du1, dv1 = e.dv1, du2 = e.du2, dv2 = e.dv2, inv_det = e.inv_det;
	V3<T> dwudx = e.dwudx, dwvdx = e.dwvdx;
	uint32_t is[3] = {e.p[0], e.p[1], e.p[2]};
	V3<T> x1 = x[is[0]], x2 = x[is[1]], x3 = x[is[2]];
	V3<T> vs[3] = {v[is[0]], v[is[1]], v[is[2]]};
	M33<T> I = M33<T>::Identity();
	V3<T> dx1 = x2 - x1, dx2 = x3 - x1;
	V3<T> wu = inv_det * (dv2 * dx1 - dv1 * dx2);
	T len_wu = Length(wu);
	V3<T> wv = inv_det * (-du2 * dx1 + du1 * dx2);
	T len_wv = Length(wv);
	// The u and v parts are identical up to (w, dwdx, ks, kd, b).
	struct Part { V3<T> w; T len; V3<T> dwdx; T ks, kd, b; };
	Part parts[2] = {{wu, len_wu, dwudx, e.ks_u, e.kd_u, e.b_u}, {wv, len_wv, dwvdx, e.ks_v, e.kd_v, e.b_v}};
	for (const Part& pt : parts) {
		if (!(pt.len > T(0))) continue;
		T inv_len = T(1) / pt.len;
		V3<T> nw = inv_len * pt.w;
		V3<T> dCdx[3];
		for (int i = 0; i < 3; ++i) dCdx[i] = alpha * pt.dwdx[i] * nw;
		if (pt.ks > T(0)) {
			T C = alpha * (pt.len - pt.b);
			V3<T> fs[3];
			for (int i = 0; i < 3; ++i) fs[i] = -pt.ks * C * dCdx[i];
			for (int i = 0; i < 3; ++i) f[is[i]] += fs[i];
			for (int i = 0; i < 3; ++i) e.act[i] += fs[i];
			M33<T> K[3][3];
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j) {
					M33<T> Kij = Outer(dCdx[i], dCdx[j]);
					if (pt.len > pt.b) {
						M33<T> d2 = (alpha * inv_len * pt.dwdx[i] * pt.dwdx[j]) * (I - Outer(nw, nw));
						Kij += C * d2;
					}
					K[i][j] = -pt.ks * Kij;
				}
			Add3x3(dfdx, is, K);
		}
		if (pt.kd > T(0)) {
			T dCdt = T(0);
			for (int i = 0; i < 3; ++i) dCdt += Dot(dCdx[i], vs[i]);
			V3<T> fs[3];
			for (int i = 0; i < 3; ++i) fs[i] = -pt.kd * dCdt * dCdx[i];
			for (int i = 0; i < 3; ++i) f[is[i]] += fs[i];
			for (int i = 0; i < 3; ++i) e.act[i] += fs[i];
			M33<T> K[3][3];
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j) K[i][j] = -pt.kd * Outer(dCdx[i], dCdx[j]);
			Add3x3(dfdv, is, K);
		}
	}
}

// shear_force.cpp:106-245
template <class T>
void Body<T>::ApplyShear(ShearForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	T alpha = e.alpha, du1 = e.du1, dv1 = e.dv1, du2 = e.du2, dv2 = e.dv2, inv_det = e.inv_det;
	V3<T> dwudx = e.dwudx, dwvdx = e.dwvdx;
	uint32_t is[3] = {e.p[0], e.p[1], e.p[2]};
	V3<T> x1 = x[is[0]], x2 = x[is[1]], x3 = x[is[2]];
	V3<T> vs[3] = {v[is[0]], v[is[1]], v[is[2]]};
	V3<T> dx1 = x2 - x1, dx2 = x3 - x1;
	V3<T> wu = inv_det * (dv2 * dx1 - dv1 * dx2);
	V3<T> wv = inv_det * (-du2 * dx1 + du1 * dx2);
	V3<T> dCdx[3];
	for (int i = 0; i < 3; ++i) dCdx[i] = alpha * (dwudx[i] * wv + dwvdx[i] * wu);
	if (e.ks > T(0)) {
		T C = alpha * Dot(wu, wv);
		V3<T> fs[3];
		for (int i = 0; i < 3; ++i) fs[i] = -e.ks * C * dCdx[i];
		for (int i = 0; i < 3; ++i) f[is[i]] += fs[i];
		for (int i = 0; i < 3; ++i) e.act[i] += fs[i];
		M33<T> K[3][3];
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j) K[i][j] = -e.ks * Outer(dCdx[i], dCdx[j]);
		Add3x3(dfdx, is, K);
	}
	if (e.kd > T(0)) {
		T dCdt = T(0);
		for (int i = 0; i < 3; ++i) dCdt += Dot(dCdx[i], vs[i]);
		V3<T> fs[3];
		for (int i = 0; i < 3; ++i) fs[i] = -e.kd * dCdt * dCdx[i];
		for (int i = 0; i < 3; ++i) f[is[i]] += fs[i];
		for (int i = 0; i < 3; ++i) e.act[i] += fs[i];
		M33<T> K[3][3];
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j) K[i][j] = -e.kd * Outer(dCdx[i], dCdx[j]);
		Add3x3(dfdv, is, K);
	}
}

// spring_force.cpp:61-143
template <class T>
void Body<T>::ApplySpring(SpringForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	uint32_t i1 = e.p[0], i2 = e.p[1];
	V3<T> x1 = x[i1], v1 = v[i1], x2 = x[i2], v2 = v[i2];
	M33<T> I = M33<T>::Identity();
	V3<T> dx = x1 - x2;
	T L = Length(dx);
	if (!(L > T(0))) return;
	V3<T> n = dx / L;
	if (e.ks > T(0)) {
		if (L > e.L0) {
			T C = L - e.L0;
			V3<T> f1 = -e.ks * C * n;
			V3<T> f2 = -f1;
			f[i1] += f1;
			f[i2] += f2;
			e.act[0] += f1;
			e.act[1] += f2;
			M33<T> K11 = -e.ks * (Outer(dx, dx) + (T(1) - e.L0 / L) * (I - Outer(dx, dx)));
			M33<T> K12 = -K11;
			M33<T> K21 = K12;
			M33<T> K22 = K11;
			dfdx.at(i1, i1) += K11;
			dfdx.at(i1, i2) += K12;
			dfdx.at(i2, i1) += K21;
			dfdx.at(i2, i2) += K22;
		}
	}
	if (e.kd > T(0)) {
		T dCdt = Dot(n, v1 - v2);
		V3<T> f1 = -e.kd * dCdt * n;
		V3<T> f2 = -f1;
		f[i1] += f1;
		f[i2] += f2;
		e.act[0] += f1;
		e.act[1] += f2;
		M33<T> K11 = -e.kd * Outer(n, n);
		M33<T> K12 = -K11;
		M33<T> K21 = K12;
		M33<T> K22 = K11;
		dfdv.at(i1, i1) += K11;
		dfdv.at(i1, i2) += K12;
		dfdv.at(i2, i1) += K21;
		dfdv.at(i2, i2) += K22;
	}
}

// mouse_force.cpp:58-246
template <class T>
void Body<T>::ApplyMouse(MouseForce<T>& e, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	uint32_t is[4] = {e.p[0], e.p[1], e.p[2], e.p[3]};
	V3<T> x1 = x[is[0]], x2 = x[is[1]], x3 = x[is[2]], x4 = x[is[3]];
	V3<T> v1 = v[is[0]], v2 = v[is[1]], v3 = v[is[2]], v4 = v[is[3]];
	M33<T> I = M33<T>::Identity();
	T w2 = e.w2, w3 = e.w3, w4 = e.w4;
	V3<T> c2 = w2 * x2 + w3 * x3 + w4 * x4;
	V3<T> d = x1 - c2;
	T L = Length(d);
	if (!(L > T(0))) return;
	T inv_L = T(1) / L;
	V3<T> n = inv_L * d;
	V3<T> dCdx[4];
	dCdx[0] = n;
	dCdx[1] = -w2 * n;
	dCdx[2] = -w3 * n;
	dCdx[3] = -w4 * n;
	if (e.ks > T(0)) {
		if (L > e.L0) {
			T C = L - e.L0;
			V3<T> fs[4];
			for (int i = 0; i < 4; ++i) fs[i] = -e.ks * C * dCdx[i];
			for (int i = 0; i < 4; ++i) f[is[i]] += fs[i];
			for (int i = 0; i < 4; ++i) e.act[i] += fs[i];
			T L3 = L * L * L;
			T inv_L3 = L3 > T(0) ? T(1) / L3 : T(0);
			M33<T> d2[4][4];
			M33<T> A = inv_L * I - inv_L3 * Outer(d, d);
			d2[0][0] = A;
			d2[0][1] = -w2 * A;
			d2[0][2] = -w3 * A;
			d2[0][3] = -w4 * A;
			M33<T> B = inv_L3 * Outer(d, d) - inv_L * I;
			d2[1][0] = w2 * B;
			d2[1][1] = -w2 * w2 * B;
			d2[1][2] = -w2 * w3 * B;
			d2[1][3] = -w2 * w4 * B;
			d2[2][0] = w3 * B;
			d2[2][1] = -w2 * w3 * B;
			d2[2][2] = -w3 * w3 * B;
			d2[2][3] = -w3 * w4 * B;
			d2[3][0] = w4 * B;
			d2[3][1] = -w2 * w4 * B;
			d2[3][2] = -w3 * w4 * B;
			d2[3][3] = -w4 * w4 * B;
			for (int i = 0; i < 4; ++i)
				for (int j = 0; j < 4; ++j) dfdx.at(is[i], is[j]) += -e.ks * (Outer(dCdx[i], dCdx[j]) + C * d2[i][j]);
		}
	}
	if (e.kd > T(0)) {
		V3<T> vc2 = w2 * v2 + w3 * v3 + w4 * v4;
		T dCdt = Dot(v1 - vc2, n);
		V3<T> fs[4];
		for (int i = 0; i < 4; ++i) fs[i] = -e.kd * dCdt * dCdx[i];
		for (int i = 0; i < 4; ++i) f[is[i]] += fs[i];
		for (int i = 0; i < 4; ++i) e.act[i] += fs[i];
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) dfdv.at(is[i], is[j]) += -e.kd * Outer(dCdx[i], dCdx[j]);
	}
}

// ======================================================================
// Collision (narrow phase of world shapes against a sphere)
// ======================================================================

// geometry.h:27-40, 45-68
template <class T> void Barycentric2(T out[3], V3<T> A, V3<T> B, V3<T> Q) {
	V3<T> AB = B - A, QA = A - Q, QB = B - Q;
	T divisor = Dot(AB, AB);
	out[0] = Dot(QB, AB);
	out[1] = -Dot(QA, AB);
	out[2] = divisor;
}
template <class T> void Barycentric3(T out[4], V3<T> A, V3<T> B, V3<T> C, V3<T> Q) {
	V3<T> AB = B - A, AC = C - A;
	V3<T> QA = A - Q, QB = B - Q, QC = C - Q;
	V3<T> QB_x_QC = Cross(QB, QC), QC_x_QA = Cross(QC, QA), QA_x_QB = Cross(QA, QB);
	V3<T> AB_x_AC = Cross(AB, AC);
	T divisor = Dot(AB_x_AC, AB_x_AC);
	out[0] = Dot(QB_x_QC, AB_x_AC);
	out[1] = Dot(QC_x_QA, AB_x_AC);
	out[2] = Dot(QA_x_QB, AB_x_AC);
	out[3] = divisor;
}

// Voxel grid helpers: voxel_grid.h, regular_grid_indexer.h, index_3d.h
template <class T> struct Grid {
	const WorldFixture<T>& w;
	V3<T> Dims() const { return MakeV3(T(w.W - 1), T(w.H - 1), T(w.D - 1)); }          // cells per axis
	V3<T> CellSize() const { return w.gridBox.Dimensions() / Dims(); }                 // :78-81
	void CellIndex(V3<T> p, int64_t c[3]) const {                                       // :104-110
		V3<T> cs = CellSize();
		V3<T> lp = p - w.gridBox.lo;
		V3<T> cp = lp / cs;
		c[0] = (int64_t)std::floor(cp.x);
		c[1] = (int64_t)std::floor(cp.y);
		c[2] = (int64_t)std::floor(cp.z);
	}
	bool ContainsCell(const int64_t c[3]) const {                                      // :51-57
		return c[0] >= 0 && c[0] < int64_t(w.W - 1) && c[1] >= 0 && c[1] < int64_t(w.H - 1) && c[2] >= 0 &&
		       c[2] < int64_t(w.D - 1);
	}
	AABB<T> CellAABB(const int64_t c[3]) const {                                        // :89-101
		V3<T> cs = CellSize();
		V3<T> lp = MakeV3(T(c[0]) * cs.x, T(c[1]) * cs.y, T(c[2]) * cs.z);
		V3<T> center = w.gridBox.lo + lp + cs / T(2);
		V3<T> radius = CellSize() / T(2);
		return AABB<T>::FromCenterRadius(center, radius);
	}
	T Voxel(int64_t i, int64_t j, int64_t k) const {                                    // index_3d.h:41-46
		return w.vox[(uint32_t)(i + j * int64_t(w.W) + k * int64_t(w.W) * int64_t(w.H))];
	}
	void CellVoxels(const int64_t c[3], T v[8]) const {                                 // voxel_grid.h:220-236
		int64_t i = c[0], j = c[1], k = c[2];
		v[0] = Voxel(i, j, k);
		v[1] = Voxel(i, j, k + 1);
		v[2] = Voxel(i, j + 1, k);
		v[3] = Voxel(i, j + 1, k + 1);
		v[4] = Voxel(i + 1, j, k);
		v[5] = Voxel(i + 1, j, k + 1);
		v[6] = Voxel(i + 1, j + 1, k);
		v[7] = Voxel(i + 1, j + 1, k + 1);
	}
	static T Interp(V3<T> r, const T v[8]) {                                              // voxel_grid.h:285-303
		T x = Clamp(r.x, T(0), T(1)), y = Clamp(r.y, T(0), T(1)), z = Clamp(r.z, T(0), T(1));
		T c00 = Lerp(x, v[0], v[4]);
		T c10 = Lerp(x, v[2], v[6]);
		T c01 = Lerp(x, v[1], v[5]);
		T c11 = Lerp(x, v[3], v[7]);
		T c0 = Lerp(y, c00, c10);
		T c1 = Lerp(y, c01, c11);
		return Lerp(z, c0, c1);
	}
};

template <class T> bool Body<T>::Collide(const WorldFixture<T>& w, V3<T> Q, T sradius, Manifold<T>* m) const {
	const T eps = Eps<T>::value();
	switch (w.type) {
	case B3S_SHAPE_SPHERE: {  // sphere_shape.cpp:50-77
		V3<T> center = w.a;
		T radius = w.radius + sradius;
		T rr = radius * radius;
		V3<T> d = Q - center;
		T dd = Dot(d, d);
		if (dd <= rr) {
			m->point = center;
			if (dd > eps * eps) { T distance = Sqrt(dd); m->normal = d / distance; }
			else m->normal = MakeV3(T(0), T(1), T(0));
			return true;
		}
		return false;
	}
	case B3S_SHAPE_CAPSULE: {  // capsule_shape.cpp:52-138
		V3<T> A = w.a, B = w.b, AB = B - A;
		T u = Dot(B - Q, AB), v = Dot(Q - A, AB);
		T radius = w.radius + sradius;
		if (v <= T(0) || u <= T(0)) {
			V3<T> P = v <= T(0) ? A : B;
			V3<T> d = Q - P;
			T dd = Dot(d, d);
			if (dd > radius * radius) return false;
			V3<T> n = MakeV3(T(0), T(1), T(0));
			T len = Sqrt(dd);
			if (len > eps) n = d / len;
			m->point = P;
			m->normal = n;
			return true;
		}
		T s = Dot(AB, AB);
		V3<T> P = (u * A + v * B) / s;
		V3<T> d = Q - P;
		T dd = Dot(d, d);
		if (dd > radius * radius) return false;
		V3<T> AQ = Q - A;
		V3<T> AB_x_AQ = Cross(AB, AQ);
		V3<T> n = Cross(AB_x_AQ, AB);
		if (Dot(n, AQ) < T(0)) n = -n;
		n.Normalize();
		m->point = P;
		m->normal = n;
		return true;
	}
	case B3S_SHAPE_TRIANGLE: {  // triangle_shape.cpp:65-312
		V3<T> A = w.a, B = w.b, C = w.c;
		T radius = w.radius + sradius;
		V3<T> triN = Cross(B - A, C - A);
		triN.Normalize();
		T wAB[3], wBC[3], wCA[3];
		Barycentric2(wAB, A, B, Q);
		Barycentric2(wBC, B, C, Q);
		Barycentric2(wCA, C, A, Q);
		auto finish = [&](V3<T> P) -> bool {
			V3<T> d = Q - P;
			T dd = Dot(d, d);
			if (dd > radius * radius) return false;
			V3<T> n = triN;
			T len = Length(d);
			if (len > eps) n = d / len;
			m->point = P;
			m->normal = n;
			return true;
		};
		if (wAB[1] <= T(0) && wCA[0] <= T(0)) return finish(A);
		if (wAB[0] <= T(0) && wBC[1] <= T(0)) return finish(B);
		if (wBC[0] <= T(0) && wCA[1] <= T(0)) return finish(C);
		T wABC[4];
		Barycentric3(wABC, A, B, C, Q);
		// Edge regions: a neighbouring face can veto (triangle_shape.cpp:148-164)
		auto edge = [&](V3<T> P, int k, V3<T> B1, V3<T> C1) -> bool {
			V3<T> d = Q - P;
			T dd = Dot(d, d);
			if (dd > radius * radius) return false;
			if (w.hasE[k]) {
				T w1[4];
				Barycentric3(w1, w.e[k], B1, C1, Q);
				if (w1[0] > T(0) && w1[1] > T(0) && w1[2] > T(0)) return false;
			}
			V3<T> n = triN;
			T len = Length(d);
			if (len > eps) n = d / len;
			m->point = P;
			m->normal = n;
			return true;
		};
		if (wAB[0] > T(0) && wAB[1] > T(0) && wABC[3] * wABC[2] <= T(0))
			return edge((wAB[0] * A + wAB[1] * B) / wAB[2], 0, B, A);
		if (wBC[0] > T(0) && wBC[1] > T(0) && wABC[3] * wABC[0] <= T(0))
			return edge((wBC[0] * B + wBC[1] * C) / wBC[2], 1, C, B);
		if (wCA[0] > T(0) && wCA[1] > T(0) && wABC[3] * wABC[1] <= T(0))
			return edge((wCA[0] * C + wCA[1] * A) / wCA[2], 2, A, C);
		return finish((wABC[0] * A + wABC[1] * B + wABC[2] * C) / wABC[3]);
	}
	case B3S_SHAPE_MESH: {  // mesh_shape.cpp:87-121
		// The sphere centre in the unscaled mesh frame; the query box keeps the
		// world radius there (the tree is built over unscaled vertices).
		V3<T> invScale = MakeV3(T(1) / w.scale.x, T(1) / w.scale.y, T(1) / w.scale.z);
		V3<T> c0 = XMulT(w.xf, Q);
		V3<T> center = MakeV3(invScale.x * c0.x, invScale.y * c0.y, invScale.z * c0.z);
		AABB<T> box = AABB<T>::FromCenterRadius(center, sradius + w.radius);
		const uint32_t nt = (uint32_t)(w.meshT.size() / 6);
		T dd0 = std::numeric_limits<T>::max();
		bool found = false;
		auto world = [&](uint32_t vi) {   // GetChildTriangle (mesh_shape.cpp:123-155): xf * (scale . v)
			V3<T> v = w.meshV[vi];
			return XMul(w.xf, MakeV3(w.scale.x * v.x, w.scale.y * v.y, w.scale.z * v.z));
		};
		// Brute force over the triangles whose tree box (mesh.h:107-121: the
		// triangle's box grown by B3_LINEAR_SLOP) overlaps the query box: the
		// same candidates the static tree reports (static_tree.h:100-139), in
		// index order, so among equal distances the lowest index wins.
		for (uint32_t ti = 0; ti < nt; ++ti) {
			const uint32_t* t = &w.meshT[6 * (size_t)ti];
			V3<T> v1 = w.meshV[t[0]], v2 = w.meshV[t[1]], v3 = w.meshV[t[2]];
			AABB<T> leaf = {Min(v1, Min(v2, v3)), Max(v1, Max(v2, v3))};
			leaf.Extend(T(0.005));
			if (leaf.hi.x < box.lo.x || leaf.lo.x > box.hi.x) continue;   // aabb.h:371-391
			if (leaf.hi.y < box.lo.y || leaf.lo.y > box.hi.y) continue;
			if (leaf.hi.z < box.lo.z || leaf.lo.z > box.hi.z) continue;
			WorldFixture<T> tri;
			tri.type = B3S_SHAPE_TRIANGLE;
			tri.radius = w.radius;
			tri.a = world(t[0]); tri.b = world(t[1]); tri.c = world(t[2]);
			for (int k = 0; k < 3; ++k) {
				tri.hasE[k] = t[3 + k] != 0xFFFFFFFFu;
				tri.e[k] = tri.hasE[k] ? world(t[3 + k]) : MakeV3(T(0), T(0), T(0));
			}
			Manifold<T> mm;
			if (!Collide(tri, Q, sradius, &mm)) continue;
			T dd = LengthSquared(Q - mm.point);
			if (dd < dd0) {
				dd0 = dd;
				*m = mm;
				found = true;
			}
		}
		return found;
	}
	case B3S_SHAPE_BOX: {  // box_shape.cpp:63-142
		V3<T> e = w.a;
		T radius = w.radius + sradius;
		V3<T> cLocal = XMulT(w.xf, Q);
		V3<T> pn[6] = {MakeV3(T(1), T(0), T(0)), MakeV3(T(-1), T(0), T(0)), MakeV3(T(0), T(1), T(0)),
		               MakeV3(T(0), T(-1), T(0)), MakeV3(T(0), T(0), T(1)), MakeV3(T(0), T(0), T(-1))};
		T po[6] = {e.x, e.x, e.y, e.y, e.z, e.z};
		uint32_t face = 0;
		T separation = -std::numeric_limits<T>::max();
		for (uint32_t i = 0; i < 6; ++i) {
			T s = Dot(pn[i], cLocal) - po[i];
			if (s > radius) return false;
			if (s > separation) { face = i; separation = s; }
		}
		if (separation < T(0)) {
			T dist = Dot(pn[face], cLocal) - po[face];
			V3<T> cBox = cLocal - dist * pn[face];
			m->point = XMul(w.xf, cBox);
			m->normal = w.xf.q.Rotate(pn[face]);
			return true;
		}
		V3<T> cBox = Clamp(cLocal, -e, e);
		V3<T> d = cLocal - cBox;
		T dd = LengthSquared(d);
		if (dd > radius * radius) return false;
		if (dd > eps * eps) {
			T distance = Sqrt(dd);
			V3<T> normal = d / distance;
			m->point = XMul(w.xf, cBox);
			m->normal = w.xf.q.Rotate(normal);
			return true;
		}
		return false;
	}
	case B3S_SHAPE_SDF: {  // sdf_shape.cpp:50-85
		V3<T> point = XMulT(w.xf, Q);
		T radius = sradius + w.radius;
		Grid<T> g{w};
		int64_t c[3];
		g.CellIndex(point, c);
		if (!g.ContainsCell(c)) return false;
		T vox[8];
		g.CellVoxels(c, vox);
		V3<T> rel = g.CellAABB(c).Relative(point);
		T distance = Grid<T>::Interp(rel, vox);
		if (w.invert) distance *= T(-1);
		if (distance > radius) return false;
		V3<T> grad;   // voxel_grid.h:183-199 (Sample* recompute the same cell)
		grad.x = Grid<T>::Interp(MakeV3(T(1), rel.y, rel.z), vox) - Grid<T>::Interp(MakeV3(T(0), rel.y, rel.z), vox);
		grad.y = Grid<T>::Interp(MakeV3(rel.x, T(1), rel.z), vox) - Grid<T>::Interp(MakeV3(rel.x, T(0), rel.z), vox);
		grad.z = Grid<T>::Interp(MakeV3(rel.x, rel.y, T(1)), vox) - Grid<T>::Interp(MakeV3(rel.x, rel.y, T(0)), vox);
		V3<T> normal = Normalize(grad);
		if (w.invert) normal *= T(-1);
		V3<T> sn = w.xf.q.Rotate(normal);
		m->point = Q - distance * sn;
		m->normal = sn;
		return true;
	}
	default: return false;
	}
}

template <class T> AABB<T> Body<T>::WorldAABB(const WorldFixture<T>& w) const {
	switch (w.type) {
	case B3S_SHAPE_SPHERE: return AABB<T>::FromCenterRadius(w.a, w.radius);   // sphere_shape.cpp:34-43
	case B3S_SHAPE_CAPSULE: {                                                 // capsule_shape.cpp:34-44
		V3<T> r = MakeV3(w.radius, w.radius, w.radius);
		return {Min(w.a, w.b) - r, Max(w.a, w.b) + r};
	}
	case B3S_SHAPE_TRIANGLE: {                                                // triangle_shape.cpp:52-63
		V3<T> r = MakeV3(w.radius, w.radius, w.radius);
		return {Min(w.a, Min(w.b, w.c)) - r, Max(w.a, Max(w.b, w.c)) + r};
	}
	case B3S_SHAPE_BOX: {                                                     // box_shape.cpp:38-61
		V3<T> e = w.a;
		V3<T> vs[8] = {MakeV3(-e.x, -e.y, -e.z), MakeV3(-e.x, -e.y, e.z), MakeV3(-e.x, e.y, -e.z), MakeV3(-e.x, e.y, e.z),
		               MakeV3(e.x, -e.y, -e.z),  MakeV3(e.x, -e.y, e.z),  MakeV3(e.x, e.y, -e.z),  MakeV3(e.x, e.y, e.z)};
		AABB<T> b;
		b.lo = b.hi = XMul(w.xf, vs[0]);
		for (int i = 1; i < 8; ++i) {
			V3<T> v = XMul(w.xf, vs[i]);
			b.lo = Min(b.lo, v);
			b.hi = Max(b.hi, v);
		}
		b.Extend(w.radius);
		return b;
	}
	case B3S_SHAPE_SDF: {                                                     // sdf_shape.cpp:42-48
		AABB<T> b = w.gridBox;
		b.Transform(w.xf);
		b.Extend(w.radius);
		return b;
	}
	case B3S_SHAPE_MESH: {                                                    // mesh_shape.cpp:44-50, aabb.h:90-99
		auto world = [&](V3<T> v) { return XMul(w.xf, MakeV3(w.scale.x * v.x, w.scale.y * v.y, w.scale.z * v.z)); };
		AABB<T> b;
		b.lo = b.hi = world(w.meshV[0]);
		for (size_t i = 1; i < w.meshV.size(); ++i) {
			V3<T> v = world(w.meshV[i]);
			b.lo = Min(b.lo, v);
			b.hi = Max(b.hi, v);
		}
		b.Extend(w.radius);
		return b;
	}
	default: return {MakeV3(T(1), T(1), T(1)), MakeV3(T(0), T(0), T(0))};
	}
}

// sphere_shape_contact.cpp:53-143
template <class T> void Body<T>::ApplyContact(Contact<T>& c, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx) {
	(void)v;
	const SphereFixture<T>& s = spheres[c.sphere];
	const WorldFixture<T>& w = worlds[c.world];
	uint32_t i1 = s.p;
	V3<T> x1 = x[i1];
	T r1 = s.radius, r2 = w.radius;
	Manifold<T> m2;
	if (!Collide(w, x1, r1, &m2)) return;
	if (!c.applyFriction) { c.manifold = m2; c.applyFriction = true; }
	V3<T> x2 = m2.point, n2 = m2.normal, n1 = -n2;
	V3<T> c1 = x1 + r1 * n1;
	V3<T> c2 = x2 + r2 * n2;
	T C = Length(c2 - c1);
	C = Min(T(0.2) * C, T(0.2));                       // B3_BAUMGARTE, B3_MAX_CONTACT_LINEAR_CORRECTION
	V3<T> f1 = -T(1000.0) * C * n1;                    // B3_CONTACT_STIFFNESS (settings.h:83)
	M33<T> I = M33<T>::Identity();
	M33<T> K11 = -T(1000.0) * (Outer(n1, n1) + C * (I - Outer(n1, n1)));
	f[i1] += f1;
	dfdx.at(i1, i1) += K11;
	c.normalForce += Length(f1);
	// B3_CONTACT_DAMPING_STIFFNESS is 0 (settings.h:86): the damping branch never runs.
}

// sphere_shape_contact.cpp:145-189
template <class T> void Body<T>::ApplyFriction(Contact<T>& c, T dt) {
	if (!c.applyFriction) return;
	V3<T> normal = c.manifold.normal;
	T normalForce = c.normalForce;
	V3<T> t1 = Perp(normal);
	V3<T> t2 = Cross(t1, normal);
	Particle<T>& p = particles[spheres[c.sphere].p];
	V3<T> v1 = p.v;
	T im1 = p.invMass;
	T friction = Sqrt(spheres[c.sphere].friction * worlds[c.world].friction);   // contact.h:28-31
	T tangentMass = im1 > T(0) ? T(1) / im1 : T(0);
	V2<T> Cdot;
	Cdot.x = Dot(v1, t1);
	Cdot.y = Dot(v1, t2);
	V2<T> impulse = tangentMass * -Cdot;
	T normalImpulse = dt * normalForce;
	T maxImpulse = friction * normalImpulse;
	if (Dot(impulse, impulse) > maxImpulse * maxImpulse) {
		impulse.Normalize();
		impulse.x *= maxImpulse;
		impulse.y *= maxImpulse;
	}
	V3<T> P1 = impulse.x * t1;
	V3<T> P2 = impulse.y * t2;
	V3<T> P = P1 + P2;
	v1 += im1 * P;
	p.v = v1;
}

// contact_manager.cpp:108-146
template <class T> void Body<T>::UpdateContacts() {
	for (auto it = contacts.begin(); it != contacts.end();) {
		const SphereFixture<T>& s = spheres[it->sphere];
		bool keep = particles[s.p].type == kDynamic && TestOverlap(SphereAABB(s), WorldAABB(worlds[it->world]));
		if (!keep) {
			contactKeys.erase({it->sphere, it->world});
			it = contacts.erase(it);
			continue;
		}
		it->normalForce = T(0);        // sphere_shape_contact.cpp:191-195
		it->applyFriction = false;
		++it;
	}
}

// contact_manager.cpp:33-82
template <class T> void Body<T>::FindNewContacts() {
	for (uint32_t si = 0; si < (uint32_t)spheres.size(); ++si) {
		AABB<T> a1 = SphereAABB(spheres[si]);
		for (uint32_t wi = 0; wi < (uint32_t)worlds.size(); ++wi) {
			if (!TestOverlap(a1, WorldAABB(worlds[wi]))) continue;
			if (contactKeys.count({si, wi})) continue;
			if (particles[spheres[si].p].type != kDynamic) continue;
			Contact<T> c;
			c.sphere = si;
			c.world = wi;
			c.normalForce = T(0);
			c.applyFriction = false;
			c.manifold.point.SetZero();
			c.manifold.normal.SetZero();
			contacts.push_front(c);
			contactKeys.insert({si, wi});
		}
	}
}

// force_solver.cpp:70-86
template <class T>
void Body<T>::ApplyForces(T h, const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
	(void)h;
	for (uint32_t i = 0; i < (uint32_t)particles.size(); ++i) {   // particle.cpp:182-199
		const Particle<T>& p = particles[i];
		if (p.damping > T(0)) {
			V3<T> fd = -p.damping * p.mass * v[i];
			f[i] += fd;
			dfdv.at(i, i) += M33<T>::Diagonal(-p.damping * p.mass);
		}
	}
	for (auto& fr : forces) {
		switch (fr.first) {
		case B3S_FORCE_TET_ELEMENT: ApplyTet(tets[fr.second], x, v, f, dfdx, dfdv); break;
		case B3S_FORCE_TRI_ELEMENT: ApplyTri(tris[fr.second], x, v, f, dfdx, dfdv); break;
		case B3S_FORCE_STRETCH: ApplyStretch(stretches[fr.second], x, v, f, dfdx, dfdv); break;
		case B3S_FORCE_SHEAR: ApplyShear(shears[fr.second], x, v, f, dfdx, dfdv); break;
		case B3S_FORCE_SPRING: ApplySpring(springs[fr.second], x, v, f, dfdx, dfdv); break;
		case B3S_FORCE_MOUSE: ApplyMouse(mice[fr.second], x, v, f, dfdx, dfdv); break;
		default: break;
		}
	}
	for (auto& c : contacts) ApplyContact(c, x, v, f, dfdx);
}

// ======================================================================
// Solvers
// ======================================================================

// sparse_solver.cpp:28-109
template <class T>
uint32_t SolveCGFree(const SpMat<T>& A, const DVec<T>& b, uint32_t maxIt, T eps, DVec<T>& x, bool wide = false) {
	auto dot = [wide](const DVec<T>& u, const DVec<T>& w) { return wide ? VDotWide(u, w) : VDot(u, w); };
	const uint32_t n = A.n();
	DMat<T> invM(n);
	for (uint32_t i = 0; i < n; ++i) {
		M33<T> a = A.get(i, i);
		invM[i] = M33<T>::Diagonal(T(1) / a.x.x, T(1) / a.y.y, T(1) / a.z.z);
	}
	DVec<T> r = VSub(b, SpMV(A, x));
	DVec<T> d = DiagMul(invM, r);
	T delta_new = dot(r, d);
	T delta_0 = delta_new;
	uint32_t it = 0;
	for (;;) {
		if (it == maxIt) break;
		if (delta_new <= eps * eps * delta_0) break;
		DVec<T> q = SpMV(A, d);
		T alpha = delta_new / dot(d, q);
		x = VAdd(x, VScale(alpha, d));
		if (it % 50 == 0) r = VSub(b, SpMV(A, x));
		else r = VSub(r, VScale(alpha, q));
		DVec<T> s = DiagMul(invM, r);
		T delta_old = delta_new;
		delta_new = dot(r, s);
		T beta = delta_new / delta_old;
		d = VAdd(s, VScale(beta, d));
		++it;
	}
	return it;
}

template <class T> uint32_t Body<T>::SolveCG(const SpMat<T>& A, const DVec<T>& b, uint32_t maxIt, T eps, DVec<T>& x) {
	return SolveCGFree(A, b, maxIt, eps, x, wideDots);
}

// sparse_force_solver.cpp:104-232, with the force model as a callable
// model(x, v, f, dfdx, dfdv) (b3SparseForceModel::ApplyForces,
// sparse_force_solver.h:44-49): f arrives zeroed and the matrices empty.
// wide: b3Dot / b3LengthSquared accumulated in double (VDotWide). stats
// keeps the caller's running min / max of the PCG counts
// (sparse_force_solver.cpp:186-188); returns the Newton iteration count.
// The final error goes to *errorOut when given.
template <class T, class Model>
uint32_t SolveBEGeneric(T h, uint32_t maxIt, T eps, uint32_t maxSub, T subEps, const DVec<T>& x0, const DVec<T>& v0,
                        const DVec<T>& fe, const DMat<T>& M, const DVec<T>& y, const DMat<T>& S, const DVec<T>& z,
                        DVec<T>& xo, DVec<T>& vo, Model&& model, bool wide, StepStats& stats,
                        Capture<T>* capture = nullptr, T* errorOut = nullptr) {
	const uint32_t n = (uint32_t)x0.size();
	DMat<T> ST(n), I(n);
	for (uint32_t i = 0; i < n; ++i) { ST[i] = Transpose(S[i]); I[i] = M33<T>::Identity(); }
	DVec<T> py(n);
	for (auto& e : py) e.SetZero();
	DVec<T> x = x0, v = v0;
	T error0 = T(0), error = T(0);
	uint32_t it = 0;
	while (it < maxIt) {
		DVec<T> fi(n);
		for (auto& e : fi) e.SetZero();
		SpMat<T> dfdx(n), dfdv(n);
		model(x, v, fi, dfdx, dfdv);
		// A = M - h * dfdv - (h * h) * dfdx
		SpMat<T> A = Minus(DiagMinus(M, Scale(h, dfdv)), Scale(h * h, dfdx));
		// b = M * (v0 - v) + h * (fe + fi) + h * dfdx * (x0 - x + h * v + y)
		DVec<T> b = VAdd(VAdd(DiagMul(M, VSub(v0, v)), VScale(h, VAdd(fe, fi))),
		                 SpMV(Scale(h, dfdx), VAdd(VAdd(VSub(x0, x), VScale(h, v)), y)));
		// pA = S * A * ST + I - S ; pb = S * (b - A * z)
		SpMat<T> pA = MinusDiag(PlusDiag(TimesDiag(DiagTimes(S, A), ST), I), S);
		DVec<T> pb = DiagMul(S, VSub(b, SpMV(A, z)));
		if (capture && capture->want && it == 0) {
			if (capture->pA) *capture->pA = pA;
			if (capture->pb) *capture->pb = pb;
			if (capture->fi) *capture->fi = fi;
		}
		uint32_t subIt = SolveCGFree(pA, pb, maxSub, subEps, py, wide);
		DVec<T> dv = VAdd(py, z);
		stats.minSub = Min(stats.minSub, subIt);
		stats.maxSub = Max(stats.maxSub, subIt);
		v = VAdd(v, dv);
		x = VAdd(VAdd(x0, VScale(h, v)), y);
		error = wide ? VLengthSquaredWide(dv) : VLengthSquared(dv);
		if (it == 0) {
			error0 = error;
			if (error <= eps * eps) break;
		} else {
			if (error <= eps * eps * error0) break;
		}
		++it;
	}
	xo = x;
	vo = v;
	if (errorOut) *errorOut = error;
	return it;
}

template <class T>
void Body<T>::SolveBE(T h, uint32_t maxIt, T eps, uint32_t maxSub, T subEps, const DVec<T>& x0, const DVec<T>& v0,
                      const DVec<T>& fe, const DMat<T>& M, const DVec<T>& y, const DMat<T>& S, const DVec<T>& z,
                      DVec<T>& xo, DVec<T>& vo) {
	auto model = [this, h](const DVec<T>& x, const DVec<T>& v, DVec<T>& f, SpMat<T>& dfdx, SpMat<T>& dfdv) {
		ApplyForces(h, x, v, f, dfdx, dfdv);
	};
	stats.newtonIterations = SolveBEGeneric(h, maxIt, eps, maxSub, subEps, x0, v0, fe, M, y, S, z, xo, vo, model,
	                                        wideDots, stats, &capture);
}

// force_solver.cpp:88-180 (+ body_solver.cpp:67-93 friction)
template <class T> void Body<T>::Solve(T dt, uint32_t iters, uint32_t subIters, T tol, T subTol) {
	const uint32_t n = (uint32_t)particles.size();
	DVec<T> x0(n), v0(n), fe(n), y(n), z(n), x(n), v(n);
	DMat<T> M(n), S(n);
	for (uint32_t i = 0; i < n; ++i) {
		const Particle<T>& p = particles[i];
		x0[i] = p.x;
		v0[i] = p.v;
		fe[i] = p.force;
		y[i] = p.translation;
		z[i].SetZero();
		if (p.type == kDynamic) {
			M[i] = M33<T>::Diagonal(p.mass);
			fe[i] += p.mass * gravity;
			S[i] = M33<T>::Identity();
		} else {
			M[i] = M33<T>::Diagonal(T(1));
			S[i] = M33<T>::Zero();
		}
	}
	SolveBE(dt, iters, tol, subIters, subTol, x0, v0, fe, M, y, S, z, x, v);
	for (uint32_t i = 0; i < n; ++i) {
		particles[i].x = x[i];
		particles[i].v = v[i];
	}
	for (auto& c : contacts) ApplyFriction(c, dt);
}

// b3TriangleShape::RayCast (triangle_shape.cpp:314-382); the shape's radius
// does not enter.
template <class T> bool RayCastTriangle(V3<T> v1, V3<T> v2, V3<T> v3, V3<T> p1, V3<T> p2, T maxFraction, T* fraction,
                                        V3<T>* normal) {
	V3<T> d = p2 - p1;
	if (LengthSquared(d) < Eps<T>::value() * Eps<T>::value()) return false;
	V3<T> n = Cross(v2 - v1, v3 - v1);
	n.Normalize();
	T num = Dot(n, v1 - p1);
	T den = Dot(n, d);
	if (den == T(0)) return false;
	T t = num / den;
	if (t < T(0) || maxFraction < t) return false;
	V3<T> Q = p1 + t * d;
	V3<T> A = v1, B = v2, C = v3;
	V3<T> AB = B - A, AC = C - A;
	V3<T> QA = A - Q, QB = B - Q, QC = C - Q;
	V3<T> QB_x_QC = Cross(QB, QC), QC_x_QA = Cross(QC, QA), QA_x_QB = Cross(QA, QB);
	V3<T> AB_x_AC = Cross(AB, AC);
	T u = Dot(QB_x_QC, AB_x_AC), v = Dot(QC_x_QA, AB_x_AC), w = Dot(QA_x_QB, AB_x_AC);
	if (u >= T(0) && v >= T(0) && w >= T(0)) {
		*fraction = t;
		*normal = num > T(0) ? -n : n;
		return true;
	}
	return false;
}

// body.cpp:497-521: the body ray cast reports the triangle fixture whose
// RayCast (triangle_fixture.cpp:58-66) gives the smallest fraction in [0, 1].
template <class T> int64_t Body<T>::RayCastSingle(V3<T> p1, V3<T> p2, T* fraction, V3<T>* normal) const {
	int64_t best = -1;
	T bestT = T(0);
	V3<T> bestN = MakeV3(T(0), T(0), T(0));
	for (const auto& t : triFixtures) {
		T f;
		V3<T> n;
		if (!RayCastTriangle(particles[t.p[0]].x, particles[t.p[1]].x, particles[t.p[2]].x, p1, p2, T(1), &f, &n)) continue;
		if (best < 0 || f < bestT || (f == bestT && (int64_t)t.index < best)) {
			best = t.index;
			bestT = f;
			bestN = n;
		}
	}
	if (best >= 0) {
		*fraction = bestT;
		*normal = bestN;
	}
	return best;
}

// b3ClosestPointOnTriangle (collision/geometry/geometry.cpp:45-104).
template <class T> V3<T> ClosestPointOnTriangle(V3<T> A, V3<T> B, V3<T> C, V3<T> Q) {
	T wAB[3], wBC[3], wCA[3];
	Barycentric2(wAB, A, B, Q);
	Barycentric2(wBC, B, C, Q);
	Barycentric2(wCA, C, A, Q);
	if (wAB[1] <= T(0) && wCA[0] <= T(0)) return A;
	if (wAB[0] <= T(0) && wBC[1] <= T(0)) return B;
	if (wBC[0] <= T(0) && wCA[1] <= T(0)) return C;
	T wABC[4];
	Barycentric3(wABC, A, B, C, Q);
	if (wAB[0] > T(0) && wAB[1] > T(0) && wABC[3] * wABC[2] <= T(0)) return (wAB[0] * A + wAB[1] * B) / wAB[2];
	if (wBC[0] > T(0) && wBC[1] > T(0) && wABC[3] * wABC[0] <= T(0)) return (wBC[0] * B + wBC[1] * C) / wBC[2];
	if (wCA[0] > T(0) && wCA[1] > T(0) && wABC[3] * wABC[1] <= T(0)) return (wCA[0] * C + wCA[1] * A) / wCA[2];
	return (wABC[0] * A + wABC[1] * B + wABC[2] * C) / wABC[3];
}

// The voxel grid b3BuildSDF creates (sdf_builder.cpp:155-174): the mesh's box
// (mesh.h:123-128, aabb.h:67-75) grown by `extension`, ceil(size / cell) + 1
// voxels per axis.
inline void SdfGrid(const float* v, uint32_t nv, const float cell[3], float extension, AABB<float>* box, uint32_t dims[3]) {
	AABB<float> b;
	b.lo = b.hi = LoadV3<float>(v);
	for (uint32_t i = 1; i < nv; ++i) {
		V3<float> p = LoadV3<float>(v + 3 * i);
		b.lo = Min(b.lo, p);
		b.hi = Max(b.hi, p);
	}
	b.Extend(extension);
	V3<float> size = b.Dimensions();
	dims[0] = uint32_t(std::ceil(size.x / cell[0])) + 1;
	dims[1] = uint32_t(std::ceil(size.y / cell[1])) + 1;
	dims[2] = uint32_t(std::ceil(size.z / cell[2])) + 1;
	*box = b;
}

// b3BuildSDF's distances (sdf_builder.cpp:26-148): per voxel (the lower
// corner of its cell, regular_grid_indexer.h:86-104) the distance to the
// closest triangle, negative when at least 3 of 5 rays to fixed far points
// (|coordinate| = the grid box's volume) cross the mesh an odd number of
// times. Voxel (i, j, k) is stored at i + j W + k W H.
inline void BuildSdf(const float* v, const uint32_t* tri, uint32_t nt, uint32_t stride, const AABB<float>& box,
                     const uint32_t dims[3], float* out) {
	const uint32_t W = dims[0], H = dims[1], D = dims[2];
	V3<float> gd = box.Dimensions();
	V3<float> cs = MakeV3(gd.x / float(W - 1), gd.y / float(H - 1), gd.z / float(D - 1));
	V3<float> d = box.hi - box.lo;
	const float far = d.x * d.y * d.z;
	const V3<float> farPoints[5] = {MakeV3(far, 0.0f, 0.0f), MakeV3(0.0f, far, 0.0f), MakeV3(0.0f, 0.0f, far),
	                                MakeV3(far, 0.0f, far), MakeV3(-far, far, 0.0f)};
	auto vert = [&](uint32_t t, int k) { return LoadV3<float>(v + 3 * (size_t)tri[(size_t)stride * t + k]); };
	for (uint32_t x = 0; x < W; ++x)
		for (uint32_t y = 0; y < H; ++y)
			for (uint32_t z = 0; z < D; ++z) {
				V3<float> local = MakeV3(float(x) * cs.x, float(y) * cs.y, float(z) * cs.z);
				V3<float> center = box.lo + local + cs / 2.0f;
				V3<float> P = center - cs / 2.0f;
				float best = std::numeric_limits<float>::max();
				for (uint32_t t = 0; t < nt; ++t) {
					V3<float> q = ClosestPointOnTriangle(vert(t, 0), vert(t, 1), vert(t, 2), P);
					float dd = LengthSquared(q - P);
					if (dd < best) best = dd;
				}
				float dist = Sqrt(best);
				uint32_t inCount = 0;
				for (int r = 0; r < 5 && inCount < 3; ++r) {
					uint32_t hits = 0;
					for (uint32_t t = 0; t < nt; ++t) {
						float f;
						V3<float> n;
						if (RayCastTriangle(vert(t, 0), vert(t, 1), vert(t, 2), P, farPoints[r], 1.0f, &f, &n)) ++hits;
					}
					if (hits % 2 != 0) ++inCount;
				}
				if (inCount >= 3) dist *= -1.0f;
				out[x + (size_t)y * W + (size_t)z * W * H] = dist;
			}
}

// body.cpp:552-604
template <class T> void Body<T>::Step(T dt, uint32_t iters, uint32_t subIters, T tol, T subTol) {
	if (!(tol > T(0))) tol = Eps<T>::value();
	if (!(subTol > T(0))) subTol = Eps<T>::value();
	UpdateContacts();
	// b3Force::ClearForces (body.cpp:566-570): the action forces restart every step
	for (auto& e : stretches) for (auto& a : e.act) a.SetZero();
	for (auto& e : shears) for (auto& a : e.act) a.SetZero();
	for (auto& e : springs) for (auto& a : e.act) a.SetZero();
	for (auto& e : mice) for (auto& a : e.act) a.SetZero();
	if (dt > T(0)) Solve(dt, iters, subIters, tol, subTol);
	for (auto& p : particles) { p.force.SetZero(); p.translation.SetZero(); }
	// Triangle-fixture proxy synchronisation (body.cpp:584-600) only moves
	// broad-phase tree proxies used for ray casts; it does not feed the step.
	FindNewContacts();
}

} // namespace b3o

#endif
