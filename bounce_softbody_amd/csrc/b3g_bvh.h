// Linear BVH over triangle fixtures for b3Body::RayCastSingle
// (dynamics/body.cpp:497-521).
//
// The reference keeps a dynamic AABB tree of the triangle fixtures and moves
// every proxy after each step (body.cpp:584-600) so that a ray query stays
// cheap. Here nothing happens per step: a query builds a linear BVH (Karras,
// "Maximizing parallelism in the construction of BVHs, octrees and k-d trees",
// HPG 2012) over the current positions when they moved since the last query,
// one sub-tree per instance of a batch, and traverses it one ray per thread.
//
// The same header holds the static trees of mesh world fixtures
// (b3MeshShape), built once on the host. The functions are __host__ __device__
// so that a CPU harness can run the same construction and traversal.
#pragma once

#include "b3g_math.h"

#include <algorithm>
#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

namespace b3g {

constexpr uint32_t kRtLeaf = 0x80000000u;   // child code: leaf (sorted position) when set, else internal node
constexpr int kRtStack = 64;                // > the tree depth: 30 Morton bits + 32 index bits
constexpr uint32_t kRtChunk = 32;           // range-union chunk (leaves per level-1 box, boxes per higher level)

struct RtBox {
	float4 lo, hi;   // w unused
};

B3G_HD RtBox rt_empty() {
	RtBox b;
	b.lo = make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.0f);
	b.hi = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
	return b;
}

B3G_HD void rt_unite(RtBox& a, const RtBox& b) {
	a.lo.x = fminf(a.lo.x, b.lo.x); a.lo.y = fminf(a.lo.y, b.lo.y); a.lo.z = fminf(a.lo.z, b.lo.z);
	a.hi.x = fmaxf(a.hi.x, b.hi.x); a.hi.y = fmaxf(a.hi.y, b.hi.y); a.hi.z = fmaxf(a.hi.z, b.hi.z);
}

// 10 bits per axis interleaved (x highest); u in [0, 1] (NaN maps to 0).
B3G_HD uint32_t rt_spread10(uint32_t v) {
	v = (v * 0x00010001u) & 0xFF0000FFu;
	v = (v * 0x00000101u) & 0x0F00F00Fu;
	v = (v * 0x00000011u) & 0xC30C30C3u;
	v = (v * 0x00000005u) & 0x49249249u;
	return v;
}
B3G_HD uint32_t rt_quant(float u) {
	const float c = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);   // fmaxf / fminf drop a NaN
	return (uint32_t)c;
}
B3G_HD uint32_t rt_morton(float x, float y, float z) {
	return (rt_spread10(rt_quant(x)) << 2) | (rt_spread10(rt_quant(y)) << 1) | rt_spread10(rt_quant(z));
}

// Common-prefix length of sorted positions i and j of one instance's segment
// [s, e): -1 outside it; equal codes are told apart by position.
B3G_HD int rt_delta(const uint32_t* code, int64_t i, int64_t j, int64_t s, int64_t e) {
	if (j < s || j >= e) return -1;
	const uint32_t a = code[i], b = code[j];
	if (a != b) return __builtin_clz(a ^ b);
	return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i (s <= i < e - 1) of the segment's radix tree: its children
// and the range [first, last] of sorted leaves it covers. Node s is the root.
B3G_HD void rt_karras(const uint32_t* code, int64_t i, int64_t s, int64_t e, uint32_t& left, uint32_t& right,
                      uint32_t& first, uint32_t& last) {
	const int d = rt_delta(code, i, i + 1, s, e) - rt_delta(code, i, i - 1, s, e) >= 0 ? 1 : -1;
	const int dmin = rt_delta(code, i, i - d, s, e);
	int64_t lmax = 2;
	while (rt_delta(code, i, i + lmax * d, s, e) > dmin) lmax *= 2;
	int64_t l = 0;
	for (int64_t t = lmax / 2; t >= 1; t /= 2)
		if (rt_delta(code, i, i + (l + t) * d, s, e) > dmin) l += t;
	const int64_t j = i + l * d;
	const int dnode = rt_delta(code, i, j, s, e);
	int64_t split = 0, step = l;
	do {
		step = (step + 1) / 2;
		const int64_t ns = split + step;
		if (ns < l && rt_delta(code, i, i + ns * d, s, e) > dnode) split = ns;
	} while (step > 1);
	const int64_t gamma = i + split * d + (d < 0 ? -1 : 0);
	const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
	left = lo == gamma ? kRtLeaf | (uint32_t)gamma : (uint32_t)gamma;
	right = hi == gamma + 1 ? kRtLeaf | (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
	first = (uint32_t)lo;
	last = (uint32_t)hi;
}

// Union of the sorted leaf boxes [a, b] (inclusive) from the leaf boxes and
// three levels of chunk boxes (chunk k of level L covers 32^L leaves).
B3G_HD RtBox rt_range_box(const RtBox* const* lv, const uint64_t* count, uint64_t a, uint64_t b) {
	RtBox r = rt_empty();
	uint64_t lo = a, hi = b + 1;   // [lo, hi) in units of the current level
	for (int L = 0; L < 4 && lo < hi; ++L) {
		if (L == 3 || hi - lo < 2 * kRtChunk) {
			for (uint64_t k = lo; k < hi; ++k) rt_unite(r, lv[L][k]);
			return r;
		}
		const uint64_t lo2 = (lo + kRtChunk - 1) / kRtChunk, hi2 = hi / kRtChunk;   // whole chunks of the next level
		for (uint64_t k = lo; k < lo2 * kRtChunk; ++k) rt_unite(r, lv[L][k]);
		for (uint64_t k = hi2 * kRtChunk; k < hi; ++k) rt_unite(r, lv[L][k]);
		lo = lo2;
		hi = hi2 < count[L + 1] ? hi2 : count[L + 1];
	}
	return r;
}

// b3TriangleShape::RayCast (collision/shapes/triangle_shape.cpp:314-382), the
// reference's operations in its order; the shape radius does not enter.
B3G_HD bool rt_triangle(V3 v1, V3 v2, V3 v3, V3 p1, V3 p2, float maxFraction, float& fraction, V3& normal) {
	const V3 d = p2 - p1;
	if (length_sq(d) < kEps * kEps) return false;
	V3 n = cross(v2 - v1, v3 - v1);
	normalize_inplace(n);
	const float num = dot(n, v1 - p1);
	const float den = dot(n, d);
	if (den == 0.0f) return false;
	const float t = num / den;
	if (t < 0.0f || maxFraction < t) return false;
	const V3 Q = p1 + t * d;
	const V3 A = v1, B = v2, C = v3;
	const V3 AB = B - A, AC = C - A;
	const V3 QA = A - Q, QB = B - Q, QC = C - Q;
	const V3 QB_x_QC = cross(QB, QC), QC_x_QA = cross(QC, QA), QA_x_QB = cross(QA, QB);
	const V3 AB_x_AC = cross(AB, AC);
	const float u = dot(QB_x_QC, AB_x_AC), v = dot(QC_x_QA, AB_x_AC), w = dot(QA_x_QB, AB_x_AC);
	if (u >= 0.0f && v >= 0.0f && w >= 0.0f) {
		fraction = t;
		normal = num > 0.0f ? -n : n;
		return true;
	}
	return false;
}

// Whether the segment p1 + t d, t in [0, tmax], meets the box grown by a
// margin (the triangle test's rounding can accept points a few ulps outside
// the exact box; the margin keeps the pruning conservative).
B3G_HD bool rt_slab(const RtBox& b, V3 p1, V3 d, float tmax) {
	float t0 = 0.0f, t1 = tmax;
	const float o[3] = {p1.x, p1.y, p1.z}, dd[3] = {d.x, d.y, d.z};
	const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
	for (int k = 0; k < 3; ++k) {
		const float m = (fabsf(lo[k]) + fabsf(hi[k]) + fabsf(o[k])) * 1.0e-5f + 1.0e-30f;
		const float l = lo[k] - m, h = hi[k] + m;
		if (dd[k] == 0.0f) {
			if (!(o[k] >= l && o[k] <= h)) return false;
			continue;
		}
		const float inv = 1.0f / dd[k];
		float a = (l - o[k]) * inv, c = (h - o[k]) * inv;
		if (a > c) { const float tmp = a; a = c; c = tmp; }
		t0 = fmaxf(t0, a - 1.0e-6f * fabsf(a));
		t1 = fminf(t1, c + 1.0e-6f * fabsf(c));
		if (!(t0 <= t1)) return false;
	}
	return true;
}

// Host only from here.

// Static tree of a mesh world fixture (the device's counterpart of
// b3Mesh::BuildTree, mesh.cpp:26-39): leaf boxes as b3Mesh::GetTriangleAABB
// (mesh.h:107-121: the triangle's box in mesh space grown by B3_LINEAR_SLOP),
// leaves in Morton order of their centres, internal nodes by Karras'
// construction, node boxes the unions of their leaves. The candidate set of a
// query (leaves whose box overlaps the query box) is the reference tree's.
// Appends one segment at position `base` of the tree arrays.
struct MeshTrees {
	std::vector<uint32_t> leaf;   // local triangle index per position
	std::vector<RtBox> leafBox, nodeBox;
	std::vector<uint2> child;
};

inline void build_mesh_tree(const float* verts, const uint32_t* tris, uint32_t nt, MeshTrees& out, uint32_t stride = 6,
                            float pad = kLinearSlop) {
	const uint32_t base = (uint32_t)out.leaf.size();
	std::vector<RtBox> box(nt);
	RtBox cb = rt_empty();
	for (uint32_t t = 0; t < nt; ++t) {
		auto vert = [&](int j) {
			const float* p = verts + 3 * (size_t)tris[stride * (size_t)t + j];
			return v3(p[0], p[1], p[2]);
		};
		const V3 v1 = vert(0), v2 = vert(1), v3_ = vert(2);
		const V3 lo = vmin(v1, vmin(v2, v3_)) - v3(pad, pad, pad);
		const V3 hi = vmax(v1, vmax(v2, v3_)) + v3(pad, pad, pad);
		box[t].lo = make_float4(lo.x, lo.y, lo.z, 0.0f);
		box[t].hi = make_float4(hi.x, hi.y, hi.z, 0.0f);
		RtBox c;
		c.lo = c.hi = make_float4(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z), 0.0f);
		rt_unite(cb, c);
	}
	std::vector<std::pair<uint32_t, uint32_t>> order(nt);
	for (uint32_t t = 0; t < nt; ++t) {
		const float ex = cb.hi.x - cb.lo.x, ey = cb.hi.y - cb.lo.y, ez = cb.hi.z - cb.lo.z;
		const float cx = 0.5f * (box[t].lo.x + box[t].hi.x), cy = 0.5f * (box[t].lo.y + box[t].hi.y),
		            cz = 0.5f * (box[t].lo.z + box[t].hi.z);
		order[t] = {rt_morton(ex > 0.0f ? (cx - cb.lo.x) / ex : 0.0f, ey > 0.0f ? (cy - cb.lo.y) / ey : 0.0f,
		                      ez > 0.0f ? (cz - cb.lo.z) / ez : 0.0f), t};
	}
	std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b2) { return a.first < b2.first; });
	std::vector<uint32_t> code(nt);
	for (uint32_t i = 0; i < nt; ++i) {
		code[i] = order[i].first;
		out.leaf.push_back(order[i].second);
		out.leafBox.push_back(box[order[i].second]);
	}
	std::vector<uint2> child(nt, make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu));
	for (uint32_t i = 0; i + 1 < nt; ++i) {
		uint32_t l, r, first, last;
		rt_karras(code.data(), i, 0, nt, l, r, first, last);
		child[i] = make_uint2(l, r);
	}
	// node boxes: unions of the children, from the root down (depth <= 62)
	std::vector<RtBox> nodeBox(nt, rt_empty());
	std::function<RtBox(uint32_t)> boxOf = [&](uint32_t c) -> RtBox {
		if (c & kRtLeaf) return box[order[c & ~kRtLeaf].second];
		RtBox r = boxOf(child[c].x);
		rt_unite(r, boxOf(child[c].y));
		nodeBox[c] = r;
		return r;
	};
	if (nt > 1) boxOf(0u);
	for (uint32_t i = 0; i < nt; ++i) {
		uint2 c = child[i];
		if (c.x != 0xFFFFFFFFu) {   // positions are global: offset by the segment's base
			c.x = (c.x & kRtLeaf) | ((c.x & ~kRtLeaf) + base);
			c.y = (c.y & kRtLeaf) | ((c.y & ~kRtLeaf) + base);
		}
		out.child.push_back(c);
		out.nodeBox.push_back(nodeBox[i]);
	}
}

} // namespace b3g
