// Narrow phase of a particle sphere against a world shape, and contact
// force / friction of b3SphereAndShapeContact, host + device.
// Operation order follows collision/shapes/*_shape.cpp,
// collision/geometry/{voxel_grid,regular_grid_indexer,aabb,geometry}.h and
// dynamics/contacts/sphere_shape_contact.cpp.
#pragma once

#include "b3g_bvh.h"
#include "b3g_internal.h"

namespace b3g {

struct Manifold { V3 point, normal; };

// geometry.h:27-40
B3G_HD void bary2(float out[3], V3 A, V3 B, V3 Q) {
	V3 AB = B - A, QA = A - Q, QB = B - Q;
	float divisor = dot(AB, AB);
	out[0] = dot(QB, AB);
	out[1] = -dot(QA, AB);
	out[2] = divisor;
}
// geometry.h:45-68
B3G_HD void bary3(float out[4], V3 A, V3 B, V3 C, V3 Q) {
	V3 AB = B - A, AC = C - A;
	V3 QA = A - Q, QB = B - Q, QC = C - Q;
	V3 QB_x_QC = cross(QB, QC), QC_x_QA = cross(QC, QA), QA_x_QB = cross(QA, QB);
	V3 AB_x_AC = cross(AB, AC);
	float divisor = dot(AB_x_AC, AB_x_AC);
	out[0] = dot(QB_x_QC, AB_x_AC);
	out[1] = dot(QC_x_QA, AB_x_AC);
	out[2] = dot(QA_x_QB, AB_x_AC);
	out[3] = divisor;
}

// voxel_grid.h:285-303
B3G_HD float grid_interp(V3 r, const float v[8]) {
	float x = clamp_(r.x, 0.0f, 1.0f), y = clamp_(r.y, 0.0f, 1.0f), z = clamp_(r.z, 0.0f, 1.0f);
	float c00 = lerp_(x, v[0], v[4]);
	float c10 = lerp_(x, v[2], v[6]);
	float c01 = lerp_(x, v[1], v[5]);
	float c11 = lerp_(x, v[3], v[7]);
	float c0 = lerp_(y, c00, c10);
	float c1 = lerp_(y, c01, c11);
	return lerp_(z, c0, c1);
}

B3G_HD bool edge_or_vertex(V3 Q, V3 P, float radius, V3 triN, Manifold* m) {
	V3 d = Q - P;
	float dd = dot(d, d);
	if (dd > radius * radius) return false;
	V3 n = triN;
	float len = length(d);
	if (len > kEps) n = d / len;
	m->point = P;
	m->normal = n;
	return true;
}

// b3TriangleShape::Collide (triangle_shape.cpp:65-312) for triangle A B C of
// radius `tradius` with optional edge-neighbour (wing) vertices.
B3G_HD bool collide_triangle(V3 A, V3 B, V3 C, float tradius, const uint32_t hasE[3], const V3 e[3], V3 Q, float sradius,
                             Manifold* m) {
	float radius = tradius + sradius;
	V3 triN = cross(B - A, C - A);
	normalize_inplace(triN);
	float wAB[3], wBC[3], wCA[3];
	bary2(wAB, A, B, Q);
	bary2(wBC, B, C, Q);
	bary2(wCA, C, A, Q);
	if (wAB[1] <= 0.0f && wCA[0] <= 0.0f) return edge_or_vertex(Q, A, radius, triN, m);
	if (wAB[0] <= 0.0f && wBC[1] <= 0.0f) return edge_or_vertex(Q, B, radius, triN, m);
	if (wBC[0] <= 0.0f && wCA[1] <= 0.0f) return edge_or_vertex(Q, C, radius, triN, m);
	float wABC[4];
	bary3(wABC, A, B, C, Q);
	int edge = -1;
	V3 P, B1, C1;
	if (wAB[0] > 0.0f && wAB[1] > 0.0f && wABC[3] * wABC[2] <= 0.0f) {
		edge = 0; P = (wAB[0] * A + wAB[1] * B) / wAB[2]; B1 = B; C1 = A;
	} else if (wBC[0] > 0.0f && wBC[1] > 0.0f && wABC[3] * wABC[0] <= 0.0f) {
		edge = 1; P = (wBC[0] * B + wBC[1] * C) / wBC[2]; B1 = C; C1 = B;
	} else if (wCA[0] > 0.0f && wCA[1] > 0.0f && wABC[3] * wABC[1] <= 0.0f) {
		edge = 2; P = (wCA[0] * C + wCA[1] * A) / wCA[2]; B1 = A; C1 = C;
	}
	if (edge >= 0) {
		V3 d = Q - P;
		float dd = dot(d, d);
		if (dd > radius * radius) return false;
		if (hasE[edge]) {   // adjacent face veto (triangle_shape.cpp:148-164)
			float w1[4];
			bary3(w1, e[edge], B1, C1, Q);
			if (w1[0] > 0.0f && w1[1] > 0.0f && w1[2] > 0.0f) return false;
		}
		return edge_or_vertex(Q, P, radius, triN, m);
	}
	return edge_or_vertex(Q, (wABC[0] * A + wABC[1] * B + wABC[2] * C) / wABC[3], radius, triN, m);
}

// Static data the shapes refer to: SDF voxels and the mesh arrays.
struct ShapeData {
	const float* vox;
	const V3* meshV;
	const uint32_t* meshT;       // 6 per triangle: v1 v2 v3 u1 u2 u3
	const uint32_t* meshLeaf;    // tree position -> local triangle index
	const RtBox* meshLeafBox;
	const uint2* meshChild;
	const RtBox* meshNodeBox;
};

B3G_HD bool rt_overlap(const RtBox& a, const AABB& b) {   // b3TestOverlap (aabb.h:371-391)
	if (a.hi.x < b.lo.x || a.lo.x > b.hi.x) return false;
	if (a.hi.y < b.lo.y || a.lo.y > b.hi.y) return false;
	if (a.hi.z < b.lo.z || a.lo.z > b.hi.z) return false;
	return true;
}

// b3MeshShape::Collide (mesh_shape.cpp:87-121): the candidates are the
// triangles whose tree box overlaps the query box around the sphere centre in
// the unscaled mesh frame (radius in world units, as the reference has it);
// the one whose manifold point is closest wins, ties to the lowest index.
B3G_HD bool collide_mesh(const WorldShape& w, const ShapeData& sd, V3 Q, float sradius, Manifold* m) {
	const V3 scale = w.a;
	const V3 invScale = v3(1.0f / scale.x, 1.0f / scale.y, 1.0f / scale.z);
	const V3 c0 = xmulT(w.xf, Q);
	const V3 center = v3(invScale.x * c0.x, invScale.y * c0.y, invScale.z * c0.z);
	const AABB query = aabb_cr(center, sradius + w.radius);
	auto world = [&](uint32_t vi) {   // GetChildTriangle (mesh_shape.cpp:123-155)
		const V3 v = sd.meshV[w.meshVert + vi];
		return xmul(w.xf, v3(scale.x * v.x, scale.y * v.y, scale.z * v.z));
	};
	float dd0 = 3.40282347e+38f;   // B3_MAX_SCALAR
	uint32_t best = 0xFFFFFFFFu;
	const uint32_t s0 = w.meshNode, e0 = w.meshNode + w.meshCount;
	uint32_t stack[kRtStack];
	int sp = 0;
	stack[sp++] = w.meshCount == 1 ? (kRtLeaf | s0) : s0;
	while (sp > 0) {
		const uint32_t c = stack[--sp];
		if (c & kRtLeaf) {
			const uint32_t k = c & ~kRtLeaf;
			if (k < s0 || k >= e0 || !rt_overlap(sd.meshLeafBox[k], query)) continue;
			const uint32_t ti = sd.meshLeaf[k];
			const uint32_t* t = sd.meshT + 6 * (size_t)(w.meshTri + ti);
			uint32_t hasE[3];
			V3 e[3];
			for (int j = 0; j < 3; ++j) {
				hasE[j] = t[3 + j] != 0xFFFFFFFFu ? 1u : 0u;
				e[j] = hasE[j] ? world(t[3 + j]) : zero3();
			}
			Manifold mm;
			if (!collide_triangle(world(t[0]), world(t[1]), world(t[2]), w.radius, hasE, e, Q, sradius, &mm)) continue;
			const float dd = length_sq(Q - mm.point);
			if (dd < dd0 || (best != 0xFFFFFFFFu && dd == dd0 && ti < best)) {
				dd0 = dd;
				best = ti;
				*m = mm;
			}
			continue;
		}
		if (c < s0 || c + 1 >= e0 || !rt_overlap(sd.meshNodeBox[c], query)) continue;
		if (sp + 2 > kRtStack) break;   // unreachable: the tree depth is at most 62
		const uint2 ch = sd.meshChild[c];
		stack[sp++] = ch.y;
		stack[sp++] = ch.x;
	}
	return best != 0xFFFFFFFFu;
}

// b3Shape::Collide for each supported shape (shape.h:67).
B3G_HD bool collide(const WorldShape& w, const ShapeData& sd, V3 Q, float sradius, Manifold* m) {
	const float* vox = sd.vox;
	switch (w.type) {
	case B3G_SHAPE_SPHERE: {                       // sphere_shape.cpp:50-77
		V3 center = w.a;
		float radius = w.radius + sradius;
		float rr = radius * radius;
		V3 d = Q - center;
		float dd = dot(d, d);
		if (dd <= rr) {
			m->point = center;
			if (dd > kEps * kEps) { float distance = sqrt_(dd); m->normal = d / distance; }
			else m->normal = v3(0.0f, 1.0f, 0.0f);
			return true;
		}
		return false;
	}
	case B3G_SHAPE_CAPSULE: {                      // capsule_shape.cpp:52-138
		V3 A = w.a, B = w.b, AB = B - A;
		float u = dot(B - Q, AB), v = dot(Q - A, AB);
		float radius = w.radius + sradius;
		if (v <= 0.0f || u <= 0.0f) {
			V3 P = v <= 0.0f ? A : B;
			V3 d = Q - P;
			float dd = dot(d, d);
			if (dd > radius * radius) return false;
			V3 n = v3(0.0f, 1.0f, 0.0f);
			float len = sqrt_(dd);
			if (len > kEps) n = d / len;
			m->point = P;
			m->normal = n;
			return true;
		}
		float s = dot(AB, AB);
		V3 P = (u * A + v * B) / s;
		V3 d = Q - P;
		float dd = dot(d, d);
		if (dd > radius * radius) return false;
		V3 AQ = Q - A;
		V3 n = cross(cross(AB, AQ), AB);
		if (dot(n, AQ) < 0.0f) n = -n;
		normalize_inplace(n);
		m->point = P;
		m->normal = n;
		return true;
	}
	case B3G_SHAPE_TRIANGLE:                       // triangle_shape.cpp:65-312
		return collide_triangle(w.a, w.b, w.c, w.radius, w.hasE, w.e, Q, sradius, m);
	case B3G_SHAPE_MESH:                           // mesh_shape.cpp:87-121
		return collide_mesh(w, sd, Q, sradius, m);
	case B3G_SHAPE_BOX: {                          // box_shape.cpp:63-142
		V3 e = w.a;
		float radius = w.radius + sradius;
		V3 cLocal = xmulT(w.xf, Q);
		const V3 pn[6] = {v3(1.0f, 0.0f, 0.0f), v3(-1.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f),
		                  v3(0.0f, -1.0f, 0.0f), v3(0.0f, 0.0f, 1.0f), v3(0.0f, 0.0f, -1.0f)};
		const float po[6] = {e.x, e.x, e.y, e.y, e.z, e.z};
		int face = 0;
		float separation = -3.40282347e+38f;   // -FLT_MAX
		for (int i = 0; i < 6; ++i) {
			float s = dot(pn[i], cLocal) - po[i];
			if (s > radius) return false;
			if (s > separation) { face = i; separation = s; }
		}
		if (separation < 0.0f) {
			float dist = dot(pn[face], cLocal) - po[face];
			V3 cBox = cLocal - dist * pn[face];
			m->point = xmul(w.xf, cBox);
			m->normal = qrotate(w.xf.q, pn[face]);
			return true;
		}
		V3 cBox = vclamp(cLocal, -e, e);
		V3 d = cLocal - cBox;
		float dd = length_sq(d);
		if (dd > radius * radius) return false;
		if (dd > kEps * kEps) {
			float distance = sqrt_(dd);
			V3 normal = d / distance;
			m->point = xmul(w.xf, cBox);
			m->normal = qrotate(w.xf.q, normal);
			return true;
		}
		return false;
	}
	case B3G_SHAPE_SDF: {                          // sdf_shape.cpp:50-85
		V3 point = xmulT(w.xf, Q);
		float radius = sradius + w.radius;
		V3 dims = v3(float(w.W - 1), float(w.H - 1), float(w.D - 1));   // regular_grid_indexer.h:66-69
		V3 cs = aabb_dims(w.grid) / dims;                                 // :78-81
		V3 cp = (point - w.grid.lo) / cs;                                 // :104-110
		int64_t ci = (int64_t)floorf(cp.x), cj = (int64_t)floorf(cp.y), ck = (int64_t)floorf(cp.z);
		if (!(ci >= 0 && ci < int64_t(w.W - 1) && cj >= 0 && cj < int64_t(w.H - 1) && ck >= 0 && ck < int64_t(w.D - 1)))
			return false;
		// cell aabb (:89-101) and relative position (aabb.h GetRelativePosition)
		V3 lp = v3(float(ci) * cs.x, float(cj) * cs.y, float(ck) * cs.z);
		V3 center = w.grid.lo + lp + cs / 2.0f;
		AABB cell = aabb_cr(center, (aabb_dims(w.grid) / dims) / 2.0f);
		V3 rel = (point - cell.lo) / aabb_dims(cell);
		const float* g = vox + w.voxOffset;
		const int64_t W = w.W, WH = int64_t(w.W) * int64_t(w.H);
		auto at = [&](int64_t i, int64_t j, int64_t k) { return g[(uint32_t)(i + j * W + k * WH)]; };
		float vv[8] = {at(ci, cj, ck),         at(ci, cj, ck + 1),         at(ci, cj + 1, ck),     at(ci, cj + 1, ck + 1),
		               at(ci + 1, cj, ck),     at(ci + 1, cj, ck + 1),     at(ci + 1, cj + 1, ck), at(ci + 1, cj + 1, ck + 1)};
		float distance = grid_interp(rel, vv);
		if (w.invert) distance *= -1.0f;
		if (distance > radius) return false;
		V3 grad;                                   // voxel_grid.h:183-199
		grad.x = grid_interp(v3(1.0f, rel.y, rel.z), vv) - grid_interp(v3(0.0f, rel.y, rel.z), vv);
		grad.y = grid_interp(v3(rel.x, 1.0f, rel.z), vv) - grid_interp(v3(rel.x, 0.0f, rel.z), vv);
		grad.z = grid_interp(v3(rel.x, rel.y, 1.0f), vv) - grid_interp(v3(rel.x, rel.y, 0.0f), vv);
		V3 normal = normalized(grad);
		if (w.invert) { normal.x *= -1.0f; normal.y *= -1.0f; normal.z *= -1.0f; }
		V3 sn = qrotate(w.xf.q, normal);
		m->point = Q - distance * sn;
		m->normal = sn;
		return true;
	}
	default: return false;
	}
}

} // namespace b3g
