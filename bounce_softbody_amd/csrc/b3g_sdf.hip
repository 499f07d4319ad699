// b3BuildSDF (collision/geometry/sdf_builder.cpp:155-174) on the device.
//
// The reference fills the voxel grid one voxel at a time on one core: for
// each voxel the closest point on every triangle (O(voxels x triangles)) and
// five parity rays to fixed far points, each against every triangle. Each
// voxel's value is a minimum and five hit counts, so the order in which
// triangles are visited does not change it. Here one thread owns one voxel.
// k_sdf_tree walks a static tree of the triangles and tests only those whose
// padded box is near enough or crossed by a ray (the default); k_sdf visits
// every triangle, staged through LDS in tiles of 256 (the fallback when the
// tree is too deep for the walk's stack). Per triangle the closest-point test
// and the rays share one pass. Quantities that do not depend on the voxel
// (the edge cross product and the unit normal of b3TriangleShape::RayCast)
// are computed once per triangle with the same operations, and the ray terms
// that do not depend on the triangle once per voxel, so every value is the
// reference's bit for bit.

#include "b3g_bvh.h"
#include "b3g_collide.h"
#include "b3g_internal.h"

#include <cmath>
#include <vector>

namespace b3g {
namespace {

constexpr int kSdfBlock = 256;

// b3ClosestPointOnTriangle (collision/geometry/geometry.cpp:45-104).
B3G_HD V3 closest_point(V3 A, V3 B, V3 C, V3 Q) {
	float wAB[3], wBC[3], wCA[3];
	bary2(wAB, A, B, Q);
	bary2(wBC, B, C, Q);
	bary2(wCA, C, A, Q);
	if (wAB[1] <= 0.0f && wCA[0] <= 0.0f) return A;
	if (wAB[0] <= 0.0f && wBC[1] <= 0.0f) return B;
	if (wBC[0] <= 0.0f && wCA[1] <= 0.0f) return C;
	float wABC[4];
	bary3(wABC, A, B, C, Q);
	if (wAB[0] > 0.0f && wAB[1] > 0.0f && wABC[3] * wABC[2] <= 0.0f) return (wAB[0] * A + wAB[1] * B) / wAB[2];
	if (wBC[0] > 0.0f && wBC[1] > 0.0f && wABC[3] * wABC[0] <= 0.0f) return (wBC[0] * B + wBC[1] * C) / wBC[2];
	if (wCA[0] > 0.0f && wCA[1] > 0.0f && wABC[3] * wABC[1] <= 0.0f) return (wCA[0] * C + wCA[1] * A) / wCA[2];
	return (wABC[0] * A + wABC[1] * B + wABC[2] * C) / wABC[3];
}

struct SdfArgs {
	uint32_t W, H, D, nt;
	V3 lo, cs;                // grid box lower bound, cell size
	float far;                // the grid box's volume (sdf_builder.cpp:118)
	const float4* tri;        // per triangle: (A, B.x) (B.yz, C.xy) (C.z, cross) (unit normal, 0)
	float* out;
};

// One voxel: its position (b3VoxelGrid::GetVoxelPosition, the lower corner
// of its cell, regular_grid_indexer.h:86-104) and the five parity rays.
struct SdfVoxel {
	V3 P;
	V3 d[5];
	bool live[5];
	float best;
	uint32_t hits[5];
};

B3G_HD void sdf_begin(const SdfArgs& a, uint32_t x, uint32_t y, uint32_t z, SdfVoxel& s) {
	const V3 local = v3(float(x) * a.cs.x, float(y) * a.cs.y, float(z) * a.cs.z);
	const V3 center = a.lo + local + a.cs / 2.0f;
	s.P = center - a.cs / 2.0f;
	const V3 farPoint[5] = {v3(a.far, 0.0f, 0.0f), v3(0.0f, a.far, 0.0f), v3(0.0f, 0.0f, a.far), v3(a.far, 0.0f, a.far),
	                        v3(-a.far, a.far, 0.0f)};
	for (int r = 0; r < 5; ++r) {
		s.d[r] = farPoint[r] - s.P;
		s.live[r] = !(length_sq(s.d[r]) < kEps * kEps);   // triangle_shape.cpp:320-323
		s.hits[r] = 0;
	}
	s.best = 3.40282347e+38f;   // B3_MAX_SCALAR
}

// One triangle record (4 float4: A, B, C, cross(B - A, C - A), unit normal):
// its distance when `dist`, and the parity rays of `rays` (bit r: ray r).
B3G_HD void sdf_triangle(const float4* rec, SdfVoxel& s, bool dist = true, uint32_t rays = 0x1Fu) {
	const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
	const V3 A = v3(r0.x, r0.y, r0.z), B = v3(r0.w, r1.x, r1.y), C = v3(r1.z, r1.w, r2.x);
	const V3 ABxAC = v3(r2.y, r2.z, r2.w), nrm = v3(r3.x, r3.y, r3.z);
	// distance (sdf_builder.cpp:82-104)
	if (dist) {
		const float dd = length_sq(closest_point(A, B, C, s.P) - s.P);
		if (dd < s.best) s.best = dd;
	}
	// parity rays (sdf_builder.cpp:26-78, triangle_shape.cpp:314-382)
	const float num = dot(nrm, A - s.P);
	for (int r = 0; r < 5; ++r) {
		if (!s.live[r] || !(rays >> r & 1u)) continue;
		const float den = dot(nrm, s.d[r]);
		if (den == 0.0f) continue;
		const float tt = num / den;
		if (tt < 0.0f || 1.0f < tt) continue;
		const V3 Q = s.P + tt * s.d[r];
		const V3 QA = A - Q, QB = B - Q, QC = C - Q;
		const float u = dot(cross(QB, QC), ABxAC), v = dot(cross(QC, QA), ABxAC), w = dot(cross(QA, QB), ABxAC);
		if (u >= 0.0f && v >= 0.0f && w >= 0.0f) ++s.hits[r];
	}
}

// Inside when at least 3 of the 5 rays cross an odd number of times; inside distances are negative.
B3G_HD float sdf_end(const SdfVoxel& s) {
	uint32_t inCount = 0;
	for (int r = 0; r < 5; ++r) inCount += s.hits[r] & 1u;
	float dist = sqrt_(s.best);
	if (inCount >= 3) dist *= -1.0f;
	return dist;
}

__global__ void __launch_bounds__(kSdfBlock) k_sdf(SdfArgs a) {
	const uint64_t v = (uint64_t)blockIdx.x * kSdfBlock + threadIdx.x;
	const uint64_t count = (uint64_t)a.W * a.H * a.D;
	const bool own = v < count;
	const uint32_t x = own ? (uint32_t)(v % a.W) : 0, y = own ? (uint32_t)((v / a.W) % a.H) : 0,
	               z = own ? (uint32_t)(v / ((uint64_t)a.W * a.H)) : 0;
	SdfVoxel s;
	sdf_begin(a, x, y, z, s);
	__shared__ float4 tile[4 * kSdfBlock];
	for (uint32_t base = 0; base < a.nt; base += kSdfBlock) {
		const uint32_t t = base + threadIdx.x;
		if (t < a.nt)
			for (int k = 0; k < 4; ++k) tile[4 * threadIdx.x + k] = a.tri[4 * (size_t)t + k];
		__syncthreads();
		const uint32_t n = min((uint32_t)kSdfBlock, a.nt - base);
		for (uint32_t k = 0; k < n; ++k) sdf_triangle(&tile[4 * k], s);
		__syncthreads();
	}
	if (own) a.out[x + (size_t)y * a.W + (size_t)z * a.W * a.H] = sdf_end(s);
}

// The same voxels over a static tree of the triangles. A node is entered when
// the voxel's squared distance to its box is at most the best so far, or
// when a live parity ray's segment passes its box; a leaf then runs only
// those tests. The leaf boxes are padded by far more than the rounding of a
// computed distance or ray hit point (the padding grows with the far points'
// distance), so a skipped triangle can neither lower the minimum nor add a
// counted crossing: every value equals the brute-force kernel's.
constexpr int kSdfStack = 64;

struct SdfTree {
	const uint2* child;       // internal nodes: children (kRtLeaf marks a leaf position)
	const RtBox* nodeBox;
	const RtBox* leafBox;     // per leaf position
	const uint32_t* leaf;     // per leaf position: triangle index
};

// rt_slab's conservative test with the ray's reciprocal direction computed
// once per voxel (the 1e-6 relative margin on t covers its rounding).
B3G_HD bool sdf_slab(const RtBox& b, V3 p1, V3 d, V3 inv) {
	float t0 = 0.0f, t1 = 1.0f;
	const float o[3] = {p1.x, p1.y, p1.z}, dd[3] = {d.x, d.y, d.z}, iv[3] = {inv.x, inv.y, inv.z};
	const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
	for (int k = 0; k < 3; ++k) {
		const float m = (fabsf(lo[k]) + fabsf(hi[k]) + fabsf(o[k])) * 1.0e-5f + 1.0e-30f;
		const float l = lo[k] - m, h = hi[k] + m;
		if (dd[k] == 0.0f) {
			if (!(o[k] >= l && o[k] <= h)) return false;
			continue;
		}
		float a = (l - o[k]) * iv[k], c = (h - o[k]) * iv[k];
		if (a > c) { const float tmp = a; a = c; c = tmp; }
		t0 = fmaxf(t0, a - 1.0e-6f * fabsf(a));
		t1 = fminf(t1, c + 1.0e-6f * fabsf(c));
		if (!(t0 <= t1)) return false;
	}
	return true;
}

B3G_HD float box_dist2(const RtBox& b, V3 P) {
	const float dx = fmaxf(fmaxf(b.lo.x - P.x, 0.0f), P.x - b.hi.x);
	const float dy = fmaxf(fmaxf(b.lo.y - P.y, 0.0f), P.y - b.hi.y);
	const float dz = fmaxf(fmaxf(b.lo.z - P.z, 0.0f), P.z - b.hi.z);
	return dx * dx + dy * dy + dz * dz;
}

__global__ void __launch_bounds__(kSdfBlock) k_sdf_tree(SdfArgs a, SdfTree tr) {
	const uint64_t v = (uint64_t)blockIdx.x * kSdfBlock + threadIdx.x;
	const uint64_t count = (uint64_t)a.W * a.H * a.D;
	if (v >= count) return;
	const uint32_t x = (uint32_t)(v % a.W), y = (uint32_t)((v / a.W) % a.H), z = (uint32_t)(v / ((uint64_t)a.W * a.H));
	SdfVoxel s;
	sdf_begin(a, x, y, z, s);
	V3 inv[5];
	for (int r = 0; r < 5; ++r)
		inv[r] = v3(s.d[r].x != 0.0f ? 1.0f / s.d[r].x : 0.0f, s.d[r].y != 0.0f ? 1.0f / s.d[r].y : 0.0f,
		            s.d[r].z != 0.0f ? 1.0f / s.d[r].z : 0.0f);
	uint32_t stack[kSdfStack];
	int sp = 0;
	stack[sp++] = a.nt == 1 ? kRtLeaf : 0u;   // the host checked that the tree's depth fits the stack
	while (sp > 0) {
		const uint32_t c = stack[--sp];
		const bool leaf = (c & kRtLeaf) != 0;
		const RtBox bx = leaf ? tr.leafBox[c & ~kRtLeaf] : tr.nodeBox[c];
		const bool dist = box_dist2(bx, s.P) <= s.best;
		uint32_t rays = 0;
		for (int r = 0; r < 5; ++r)
			if (s.live[r] && sdf_slab(bx, s.P, s.d[r], inv[r])) rays |= 1u << r;
		if (!dist && !rays) continue;
		if (leaf) {
			sdf_triangle(a.tri + 4 * (size_t)tr.leaf[c & ~kRtLeaf], s, dist, rays);
		} else {
			const uint2 ch = tr.child[c];
			stack[sp++] = ch.y;
			stack[sp++] = ch.x;
		}
	}
	a.out[x + (size_t)y * a.W + (size_t)z * a.W * a.H] = sdf_end(s);
}

// Depth of a tree (root 0, kRtLeaf leaves): the stack of k_sdf_tree holds at
// most depth + 1 entries.
inline uint32_t tree_depth(const std::vector<uint2>& child, uint32_t nt) {
	if (nt < 2) return 0;
	uint32_t deepest = 0;
	std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, 0u}};
	while (!todo.empty()) {
		const auto [c, d] = todo.back();
		todo.pop_back();
		deepest = std::max(deepest, d);
		if (c & kRtLeaf) continue;
		todo.push_back({child[c].x, d + 1});
		todo.push_back({child[c].y, d + 1});
	}
	return deepest;
}

// per triangle: the vertices, cross(B - A, C - A) and its unit vector as
// b3TriangleShape::RayCast forms them (triangle_shape.cpp:327-328, 347-357)
inline std::vector<float4> sdf_records(const float* vertices, const uint32_t* triangles, uint32_t nt, uint32_t stride) {
	std::vector<float4> rec(4 * (size_t)nt);
	for (uint64_t t = 0; t < nt; ++t) {
		auto vert = [&](int k) {
			const float* p = vertices + 3 * (size_t)triangles[stride * t + k];
			return v3(p[0], p[1], p[2]);
		};
		const V3 A = vert(0), B = vert(1), C = vert(2);
		const V3 c = cross(B - A, C - A);
		V3 n = c;
		normalize_inplace(n);
		rec[4 * t] = make_float4(A.x, A.y, A.z, B.x);
		rec[4 * t + 1] = make_float4(B.y, B.z, C.x, C.y);
		rec[4 * t + 2] = make_float4(C.z, c.x, c.y, c.z);
		rec[4 * t + 3] = make_float4(n.x, n.y, n.z, 0.0f);
	}
	return rec;
}

inline SdfArgs sdf_args(const AABB& box, const uint32_t dims[3], uint32_t nt) {
	SdfArgs a;
	a.W = dims[0]; a.H = dims[1]; a.D = dims[2]; a.nt = nt;
	a.lo = box.lo;
	const V3 gd = box.hi - box.lo;   // aabb.GetDimensions() / indexer.GetDimensions() (regular_grid_indexer.h:78-81)
	a.cs = v3(gd.x / float(dims[0] - 1), gd.y / float(dims[1] - 1), gd.z / float(dims[2] - 1));
	const V3 e = box.hi - box.lo;    // aabb.h:174-178
	a.far = e.x * e.y * e.z;
	a.tri = nullptr;
	a.out = nullptr;
	return a;
}

} // namespace

// The grid b3BuildSDF creates: the mesh's box (mesh.h:123-128, aabb.h:67-75)
// grown by `extension` (aabb.h:109-114), ceil(size / cell) + 1 voxels per axis.
b3g_status sdf_grid(const float* vertices, uint32_t nv, const float cell[3], float extension, AABB& box, uint32_t dims[3]) {
	if (!vertices || nv == 0 || !cell) { set_error("a mesh needs vertices and a cell size"); return B3G_INVALID_ARGUMENT; }
	if (!(cell[0] > 0.0f && cell[1] > 0.0f && cell[2] > 0.0f)) { set_error("cell sizes must be positive"); return B3G_INVALID_ARGUMENT; }
	box.lo = box.hi = v3(vertices[0], vertices[1], vertices[2]);
	for (uint32_t i = 1; i < nv; ++i) {
		const V3 p = v3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
		box.lo = vmin(box.lo, p);
		box.hi = vmax(box.hi, p);
	}
	aabb_extend(box, extension);
	const V3 size = box.hi - box.lo;
	const double cells[3] = {std::ceil(size.x / cell[0]), std::ceil(size.y / cell[1]), std::ceil(size.z / cell[2])};
	for (int k = 0; k < 3; ++k) {
		if (!(cells[k] >= 1.0 && cells[k] < 65535.0)) { set_error("voxel grid dimensions out of range"); return B3G_INVALID_ARGUMENT; }
		dims[k] = (uint32_t)cells[k] + 1;
	}
	return B3G_OK;
}

b3g_status build_sdf(int device, const float* vertices, uint32_t nv, const uint32_t* triangles, uint32_t nt,
                     uint32_t stride, const float cell[3], float extension, float* voxels, uint64_t capacity) {
	AABB box;
	uint32_t dims[3];
	b3g_status s = sdf_grid(vertices, nv, cell, extension, box, dims);
	if (s != B3G_OK) return s;
	if (!triangles || nt == 0 || stride < 3) { set_error("a mesh needs triangles (stride >= 3)"); return B3G_INVALID_ARGUMENT; }
	for (uint64_t t = 0; t < nt; ++t)
		for (int k = 0; k < 3; ++k)
			if (triangles[stride * t + k] >= nv) { set_error("triangle vertex index out of range"); return B3G_INVALID_ARGUMENT; }
	const uint64_t count = (uint64_t)dims[0] * dims[1] * dims[2];
	if (!voxels || capacity < count) {
		set_error("voxel buffer holds " + std::to_string(capacity) + " values; the grid has " + std::to_string(count));
		return B3G_INVALID_ARGUMENT;
	}
	if (count > (1ull << 32)) { set_error("more than 2^32 voxels"); return B3G_UNSUPPORTED; }
	std::vector<float4> rec = sdf_records(vertices, triangles, nt, stride);
	// the triangle tree, padded beyond the rounding of distances and ray hits
	SdfArgs args = sdf_args(box, dims, nt);
	const V3 ext = box.hi - box.lo;
	const float pad = kLinearSlop + 1.0e-5f * (fabsf(args.far) + ext.x + ext.y + ext.z);
	MeshTrees tree;
	build_mesh_tree(vertices, triangles, nt, tree, stride, pad);
	const bool useTree = tree_depth(tree.child, nt) + 2 <= (uint32_t)kSdfStack;
	B3G_HIP(hipSetDevice(device));
	hipStream_t stream;
	B3G_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
	float4* dTri = nullptr;
	float* dOut = nullptr;
	void* dTree = nullptr;
	b3g_status st = B3G_OK;
	do {
		if (hipMalloc(&dTri, rec.size() * sizeof(float4)) != hipSuccess || hipMalloc(&dOut, count * 4) != hipSuccess) {
			set_error("out of device memory for the voxel grid");
			st = B3G_OUT_OF_MEMORY;
			break;
		}
		if (hipMemcpyAsync(dTri, rec.data(), rec.size() * sizeof(float4), hipMemcpyHostToDevice, stream) != hipSuccess) {
			st = B3G_HIP_ERROR;
			break;
		}
		SdfArgs a = args;
		a.tri = dTri;
		a.out = dOut;
		const unsigned grid = (unsigned)((count + kSdfBlock - 1) / kSdfBlock);
		if (useTree) {
			// one allocation: child, node boxes, leaf boxes, leaf triangles
			const size_t nc = std::max<size_t>(1, tree.child.size());
			const size_t bChild = nc * sizeof(uint2), bNode = nc * sizeof(RtBox), bLeaf = tree.leafBox.size() * sizeof(RtBox),
			             bIdx = tree.leaf.size() * 4;
			if (hipMalloc(&dTree, bChild + bNode + bLeaf + bIdx) != hipSuccess) {
				set_error("out of device memory for the triangle tree");
				st = B3G_OUT_OF_MEMORY;
				break;
			}
			char* p = static_cast<char*>(dTree);
			SdfTree tr;
			tr.child = reinterpret_cast<const uint2*>(p);
			tr.nodeBox = reinterpret_cast<const RtBox*>(p + bChild);
			tr.leafBox = reinterpret_cast<const RtBox*>(p + bChild + bNode);
			tr.leaf = reinterpret_cast<const uint32_t*>(p + bChild + bNode + bLeaf);
			if ((!tree.child.empty() &&
			     (hipMemcpyAsync(p, tree.child.data(), tree.child.size() * sizeof(uint2), hipMemcpyHostToDevice, stream) != hipSuccess ||
			      hipMemcpyAsync(p + bChild, tree.nodeBox.data(), tree.nodeBox.size() * sizeof(RtBox), hipMemcpyHostToDevice,
			                     stream) != hipSuccess)) ||
			    hipMemcpyAsync(p + bChild + bNode, tree.leafBox.data(), bLeaf, hipMemcpyHostToDevice, stream) != hipSuccess ||
			    hipMemcpyAsync(p + bChild + bNode + bLeaf, tree.leaf.data(), bIdx, hipMemcpyHostToDevice, stream) != hipSuccess) {
				st = B3G_HIP_ERROR;
				break;
			}
			hipLaunchKernelGGL(k_sdf_tree, dim3(grid), dim3(kSdfBlock), 0, stream, a, tr);
		} else {
			hipLaunchKernelGGL(k_sdf, dim3(grid), dim3(kSdfBlock), 0, stream, a);
		}
		if (hipGetLastError() != hipSuccess) { st = B3G_HIP_ERROR; break; }
		if (hipMemcpyAsync(voxels, dOut, count * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
		    hipStreamSynchronize(stream) != hipSuccess)
			st = B3G_HIP_ERROR;
	} while (0);
	if (st == B3G_HIP_ERROR) set_error("the SDF build failed on the device");
	(void)hipStreamSynchronize(stream);
	if (dTri) (void)hipFree(dTri);
	if (dOut) (void)hipFree(dOut);
	if (dTree) (void)hipFree(dTree);
	(void)hipStreamDestroy(stream);
	return st;
}

} // namespace b3g
