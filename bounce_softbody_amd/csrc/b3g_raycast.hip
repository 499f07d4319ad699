// b3Body::RayCastSingle (dynamics/body.cpp:497-521) on the device, for many
// rays at once, against one instance's triangle fixtures.
//
// The reference moves one dynamic-tree proxy per triangle fixture after every
// step (body.cpp:584-600) only so that this query can prune. Here the step
// does no tree work. A query rebuilds a linear BVH when particles moved since
// the previous query:
//   1. leaf boxes from the current positions, per-instance centroid bounds;
//   2. keys (instance << 30 | 30-bit Morton code of the box centre), radix
//      sorted, so that every instance is one contiguous segment;
//   3. internal nodes by Karras' construction, each segment its own tree whose
//      root is the segment's first internal node;
//   4. node boxes as unions over each node's contiguous leaf range, from the
//      leaf boxes and three levels of 32-wide chunk boxes: every node is
//      independent, so no bottom-up pass with atomics and device-scope fences
//      is needed (such fences write back the L2 of each XCD on MI355X).
// Traversal is one ray per thread with a stack; a node is visited when the
// ray segment up to the best fraction so far meets its (slightly grown) box,
// and the exact triangle test decides. The result is the fixture with the
// smallest fraction, ties going to the lowest creation index.

#include "b3g_bvh.h"
#include "b3g_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

namespace b3g {
namespace {

constexpr int kRtBlock = 256;

inline uint32_t rt_blocks(uint64_t n) { return (uint32_t)((n + kRtBlock - 1) / kRtBlock); }

// Instance whose segment [begin[k], begin[k + 1]) holds position i.
__device__ __forceinline__ uint32_t rt_instance(const uint32_t* begin, uint32_t nInst, uint32_t i) {
	uint32_t lo = 0, hi = nInst;   // begin[lo] <= i < begin[hi]
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) / 2;
		if (begin[mid] <= i) lo = mid;
		else hi = mid;
	}
	return lo;
}

__global__ void __launch_bounds__(kRtBlock) k_rt_leaf(uint32_t T, const uint4* tri, const V3* x, RtBox* box) {
	const uint32_t t = blockIdx.x * kRtBlock + threadIdx.x;
	if (t >= T) return;
	const uint4 r = tri[t];
	const V3 a = x[r.x], b = x[r.y], c = x[r.z];
	RtBox o;
	o.lo = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)), 0.0f);
	o.hi = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)), 0.0f);
	box[t] = o;
}

// Bounds of the leaf-box centres of each instance: `parts` workgroups per
// instance each cover a strided share of its leaves (partial[k parts + part]),
// then k_rt_inst_join unites an instance's partials (min / max are exact in
// any order).
__global__ void __launch_bounds__(kRtBlock) k_rt_inst_box(const uint32_t* begin, uint32_t parts, const RtBox* box,
                                                          RtBox* partial) {
	const uint32_t k = blockIdx.x / parts, part = blockIdx.x % parts;
	const uint32_t s = begin[k], e = begin[k + 1];
	RtBox r = rt_empty();
	for (uint32_t t = s + part * kRtBlock + threadIdx.x; t < e; t += parts * kRtBlock) {
		const RtBox b = box[t];
		const float4 c = make_float4(0.5f * (b.lo.x + b.hi.x), 0.5f * (b.lo.y + b.hi.y), 0.5f * (b.lo.z + b.hi.z), 0.0f);
		RtBox p;
		p.lo = c;
		p.hi = c;
		rt_unite(r, p);
	}
	__shared__ RtBox lds[kRtBlock];
	lds[threadIdx.x] = r;
	__syncthreads();
	for (uint32_t w = kRtBlock / 2; w > 0; w /= 2) {
		if (threadIdx.x < w) rt_unite(lds[threadIdx.x], lds[threadIdx.x + w]);
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void __launch_bounds__(kRtBlock) k_rt_inst_join(uint32_t nInst, uint32_t parts, const RtBox* partial,
                                                           RtBox* out) {
	const uint32_t k = blockIdx.x * kRtBlock + threadIdx.x;
	if (k >= nInst) return;
	RtBox r = rt_empty();
	for (uint32_t p = 0; p < parts; ++p) rt_unite(r, partial[(size_t)k * parts + p]);
	out[k] = r;
}

__global__ void __launch_bounds__(kRtBlock) k_rt_morton(uint32_t T, uint32_t nInst, const uint32_t* begin,
                                                        const RtBox* box, const RtBox* instBox, uint64_t* key,
                                                        uint32_t* val) {
	const uint32_t t = blockIdx.x * kRtBlock + threadIdx.x;
	if (t >= T) return;
	const uint32_t k = rt_instance(begin, nInst, t);
	const RtBox b = box[t], ib = instBox[k];
	const float cx = 0.5f * (b.lo.x + b.hi.x), cy = 0.5f * (b.lo.y + b.hi.y), cz = 0.5f * (b.lo.z + b.hi.z);
	const float ex = ib.hi.x - ib.lo.x, ey = ib.hi.y - ib.lo.y, ez = ib.hi.z - ib.lo.z;
	const float ux = ex > 0.0f ? (cx - ib.lo.x) / ex : 0.0f;
	const float uy = ey > 0.0f ? (cy - ib.lo.y) / ey : 0.0f;
	const float uz = ez > 0.0f ? (cz - ib.lo.z) / ez : 0.0f;
	key[t] = ((uint64_t)k << 30) | rt_morton(ux, uy, uz);
	val[t] = t;
}

// Sorted order: 30-bit codes, leaf records and leaf boxes.
__global__ void __launch_bounds__(kRtBlock) k_rt_sorted(uint32_t T, const uint64_t* key, const uint32_t* val,
                                                        const uint4* tri, const RtBox* box, uint32_t* code,
                                                        uint4* leafTri, RtBox* leafBox) {
	const uint32_t i = blockIdx.x * kRtBlock + threadIdx.x;
	if (i >= T) return;
	const uint32_t t = val[i];
	code[i] = (uint32_t)(key[i] & 0x3FFFFFFFull);
	leafTri[i] = tri[t];
	leafBox[i] = box[t];
}

__global__ void __launch_bounds__(kRtBlock) k_rt_karras(uint32_t T, uint32_t nInst, const uint32_t* begin,
                                                        const uint32_t* code, uint2* child, uint2* range) {
	const uint32_t i = blockIdx.x * kRtBlock + threadIdx.x;
	if (i >= T) return;
	const uint32_t k = rt_instance(begin, nInst, i);
	const uint32_t s = begin[k], e = begin[k + 1];
	if (i + 1 >= e) {   // the last position of a segment is no internal node
		child[i] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
		range[i] = make_uint2(i, i);
		return;
	}
	uint32_t l, r, first, last;
	rt_karras(code, i, s, e, l, r, first, last);
	child[i] = make_uint2(l, r);
	range[i] = make_uint2(first, last);
}

// Level L + 1 chunk boxes from level L: box k = union of items [32 k, 32 k + 32).
__global__ void __launch_bounds__(kRtBlock) k_rt_chunk(uint64_t nIn, uint64_t nOut, const RtBox* in, RtBox* out) {
	const uint64_t k = (uint64_t)blockIdx.x * kRtBlock + threadIdx.x;
	if (k >= nOut) return;
	RtBox r = rt_empty();
	const uint64_t e = std::min<uint64_t>(nIn, (k + 1) * kRtChunk);
	for (uint64_t j = k * kRtChunk; j < e; ++j) rt_unite(r, in[j]);
	out[k] = r;
}

struct RtLevels {
	const RtBox* lv[4];
	uint64_t count[4];
};

__global__ void __launch_bounds__(kRtBlock) k_rt_node_box(uint32_t T, RtLevels L, const uint2* child, const uint2* range,
                                                          RtBox* nodeBox) {
	const uint32_t i = blockIdx.x * kRtBlock + threadIdx.x;
	if (i >= T) return;
	if (child[i].x == 0xFFFFFFFFu) return;
	const uint2 r = range[i];
	nodeBox[i] = rt_range_box(L.lv, L.count, r.x, r.y);
}

struct RtCast {
	uint32_t count, s, e;       // rays; the instance's segment of sorted positions
	const float* p1;            // 3 floats per ray
	const float* p2;
	const uint2* child;
	const RtBox* nodeBox;
	const RtBox* leafBox;
	const uint4* leafTri;       // list-order particles, creation index
	const V3* x;
	int64_t* index;             // per ray: creation index of the triangle fixture hit, or -1
	float* fraction;
	float* normal;              // 3 floats per ray, or null
	uint32_t* fault;            // set when a stack or index bound would be exceeded
};

__global__ void __launch_bounds__(kRtBlock) k_rt_cast(RtCast a) {
	const uint32_t r = blockIdx.x * kRtBlock + threadIdx.x;
	if (r >= a.count) return;
	const V3 p1 = v3(a.p1[3 * (size_t)r], a.p1[3 * (size_t)r + 1], a.p1[3 * (size_t)r + 2]);
	const V3 p2 = v3(a.p2[3 * (size_t)r], a.p2[3 * (size_t)r + 1], a.p2[3 * (size_t)r + 2]);
	const V3 d = p2 - p1;
	uint32_t bestIdx = 0xFFFFFFFFu;
	float bestT = 1.0f;
	V3 bestN = zero3();
	if (a.e > a.s && !(length_sq(d) < kEps * kEps)) {
		uint32_t stack[kRtStack];
		int sp = 0;
		stack[sp++] = a.e - a.s == 1 ? (kRtLeaf | a.s) : a.s;
		while (sp > 0) {
			const uint32_t c = stack[--sp];
			const float tmax = bestIdx != 0xFFFFFFFFu ? bestT : 1.0f;
			if (c & kRtLeaf) {
				const uint32_t k = c & ~kRtLeaf;
				if (k < a.s || k >= a.e) { atomicOr(a.fault, 1u); break; }
				if (!rt_slab(a.leafBox[k], p1, d, tmax)) continue;
				const uint4 t = a.leafTri[k];
				float f;
				V3 n;
				if (!rt_triangle(a.x[t.x], a.x[t.y], a.x[t.z], p1, p2, 1.0f, f, n)) continue;
				if (bestIdx == 0xFFFFFFFFu || f < bestT || (f == bestT && t.w < bestIdx)) {
					bestIdx = t.w;
					bestT = f;
					bestN = n;
				}
				continue;
			}
			if (c < a.s || c + 1 >= a.e) { atomicOr(a.fault, 1u); break; }
			if (!rt_slab(a.nodeBox[c], p1, d, tmax)) continue;
			if (sp + 2 > kRtStack) { atomicOr(a.fault, 2u); break; }
			const uint2 ch = a.child[c];
			stack[sp++] = ch.y;
			stack[sp++] = ch.x;
		}
	}
	const bool hit = bestIdx != 0xFFFFFFFFu;
	a.index[r] = hit ? (int64_t)bestIdx : -1;
	a.fraction[r] = hit ? bestT : 0.0f;
	if (a.normal) {
		a.normal[3 * (size_t)r] = bestN.x;
		a.normal[3 * (size_t)r + 1] = bestN.y;
		a.normal[3 * (size_t)r + 2] = bestN.z;
	}
}

b3g_status rt_alloc(Body& b, DBuf& buf, size_t bytes) {
	if (buf.p && buf.bytes >= bytes) return B3G_OK;
	return dalloc(b, buf, bytes);
}

// Builds the tree over the current positions (all instances).
b3g_status rt_build(Body& b) {
	const uint32_t T = b.nRayTri, nInst = b.nInst;
	b3g_status s;
	if (!b.rtTri.p) {
		if ((s = dupload(b, b.rtTri, b.hostRayTri.data(), b.hostRayTri.size() * 4)) != B3G_OK) return s;
		if ((s = dupload(b, b.rtInstBegin, b.hostRayInstBegin.data(), b.hostRayInstBegin.size() * 4)) != B3G_OK) return s;
	}
	const uint64_t n1 = (T + kRtChunk - 1) / kRtChunk, n2 = (n1 + kRtChunk - 1) / kRtChunk, n3 = (n2 + kRtChunk - 1) / kRtChunk;
	if ((s = rt_alloc(b, b.rtBox, (size_t)T * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtKey, (size_t)T * 8)) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtKey2, (size_t)T * 8)) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtVal, (size_t)T * 4)) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtVal2, (size_t)T * 4)) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtLeafTri, (size_t)T * sizeof(uint4))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtLeafBox, (size_t)T * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtChild, (size_t)T * sizeof(uint2))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtNodeBox, (size_t)T * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtChunk1, (size_t)std::max<uint64_t>(1, n1) * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtChunk2, (size_t)std::max<uint64_t>(1, n2) * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtChunk3, (size_t)std::max<uint64_t>(1, n3) * sizeof(RtBox))) != B3G_OK) return s;
	if ((s = rt_alloc(b, b.rtInstBox, (size_t)nInst * sizeof(RtBox))) != B3G_OK) return s;
	// workgroups per instance for its bounds: about 32 leaves per thread for
	// the largest instance, at most 65536 workgroups in all
	uint32_t maxLeaves = 0;
	for (uint32_t k = 0; k < nInst; ++k)
		maxLeaves = std::max(maxLeaves, b.hostRayInstBegin[k + 1] - b.hostRayInstBegin[k]);
	uint32_t parts = std::max(1u, (maxLeaves + 32 * kRtBlock - 1) / (32 * kRtBlock));
	parts = std::max(1u, std::min(parts, 65536u / std::max(1u, nInst)));
	if ((s = rt_alloc(b, b.rtInstPart, (size_t)nInst * parts * sizeof(RtBox))) != B3G_OK) return s;
	int instBits = 0;
	while (instBits < 34 && (1ull << instBits) < nInst) ++instBits;
	const int endBit = 30 + instBits;
	uint64_t* key = dptr<uint64_t>(b.rtKey);
	uint64_t* key2 = dptr<uint64_t>(b.rtKey2);
	uint32_t* val = dptr<uint32_t>(b.rtVal);
	uint32_t* val2 = dptr<uint32_t>(b.rtVal2);
	if (!b.rtSortBytes) {
		size_t bytes = 0;
		B3G_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key2, val, val2, (int)T, 0, endBit, b.stream));
		if ((s = rt_alloc(b, b.rtSortTemp, std::max<size_t>(bytes, 16))) != B3G_OK) return s;
		b.rtSortBytes = std::max<size_t>(bytes, 16);
	}
	const uint32_t* begin = dptr<uint32_t>(b.rtInstBegin);
	RtBox* box = dptr<RtBox>(b.rtBox);
	hipLaunchKernelGGL(k_rt_leaf, dim3(rt_blocks(T)), dim3(kRtBlock), 0, b.stream, T, dptr<uint4>(b.rtTri), dptr<V3>(b.x),
	                   box);
	B3G_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_rt_inst_box, dim3(nInst * parts), dim3(kRtBlock), 0, b.stream, begin, parts, box,
	                   dptr<RtBox>(b.rtInstPart));
	B3G_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_rt_inst_join, dim3(rt_blocks(nInst)), dim3(kRtBlock), 0, b.stream, nInst, parts,
	                   dptr<RtBox>(b.rtInstPart), dptr<RtBox>(b.rtInstBox));
	B3G_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_rt_morton, dim3(rt_blocks(T)), dim3(kRtBlock), 0, b.stream, T, nInst, begin, box,
	                   dptr<RtBox>(b.rtInstBox), key, val);
	B3G_HIP(hipGetLastError());
	size_t bytes = b.rtSortBytes;
	B3G_HIP(hipcub::DeviceRadixSort::SortPairs(b.rtSortTemp.p, bytes, key, key2, val, val2, (int)T, 0, endBit, b.stream));
	// codes go to rtVal (free after the sort); node ranges to rtKey (free after the sort)
	uint32_t* code = val;
	uint2* range = reinterpret_cast<uint2*>(key);
	hipLaunchKernelGGL(k_rt_sorted, dim3(rt_blocks(T)), dim3(kRtBlock), 0, b.stream, T, key2, val2, dptr<uint4>(b.rtTri),
	                   box, code, dptr<uint4>(b.rtLeafTri), dptr<RtBox>(b.rtLeafBox));
	B3G_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_rt_karras, dim3(rt_blocks(T)), dim3(kRtBlock), 0, b.stream, T, nInst, begin, code,
	                   dptr<uint2>(b.rtChild), range);
	B3G_HIP(hipGetLastError());
	RtLevels L;
	L.lv[0] = dptr<RtBox>(b.rtLeafBox); L.count[0] = T;
	L.lv[1] = dptr<RtBox>(b.rtChunk1); L.count[1] = n1;
	L.lv[2] = dptr<RtBox>(b.rtChunk2); L.count[2] = n2;
	L.lv[3] = dptr<RtBox>(b.rtChunk3); L.count[3] = n3;
	for (int l = 0; l < 3; ++l) {
		hipLaunchKernelGGL(k_rt_chunk, dim3(rt_blocks(L.count[l + 1])), dim3(kRtBlock), 0, b.stream, L.count[l],
		                   L.count[l + 1], L.lv[l], const_cast<RtBox*>(L.lv[l + 1]));
		B3G_HIP(hipGetLastError());
	}
	hipLaunchKernelGGL(k_rt_node_box, dim3(rt_blocks(T)), dim3(kRtBlock), 0, b.stream, T, L, dptr<uint2>(b.rtChild),
	                   range, dptr<RtBox>(b.rtNodeBox));
	B3G_HIP(hipGetLastError());
	b.rtVersion = b.xVersion;
	return B3G_OK;
}

} // namespace

namespace {

// The instance's tree (rebuilt when particles moved since the last query)
// and the cast of `count` rays whose arrays are all in device memory; the
// fault word is cleared first and checked after (the call blocks).
b3g_status cast_rays(Body& b, uint32_t inst, uint32_t count, const float* p1, const float* p2, int64_t* index,
                     float* fraction, float* normal, uint32_t* fault) {
	const uint32_t s0 = b.hostRayInstBegin.empty() ? 0 : b.hostRayInstBegin[inst];
	const uint32_t e0 = b.hostRayInstBegin.empty() ? 0 : b.hostRayInstBegin[inst + 1];
	b3g_status s;
	if (e0 > s0 && b.rtVersion != b.xVersion && (s = rt_build(b)) != B3G_OK) return s;
	B3G_HIP(hipMemsetAsync(fault, 0, 4, b.stream));
	RtCast a;
	a.count = count;
	a.s = s0;
	a.e = e0;
	a.p1 = p1;
	a.p2 = p2;
	a.child = dptr<uint2>(b.rtChild);
	a.nodeBox = dptr<RtBox>(b.rtNodeBox);
	a.leafBox = dptr<RtBox>(b.rtLeafBox);
	a.leafTri = dptr<uint4>(b.rtLeafTri);
	a.x = dptr<V3>(b.x);
	a.index = index;
	a.fraction = fraction;
	a.normal = normal;
	a.fault = fault;
	hipLaunchKernelGGL(k_rt_cast, dim3(rt_blocks(count)), dim3(kRtBlock), 0, b.stream, a);
	B3G_HIP(hipGetLastError());
	uint32_t f = 0;
	B3G_HIP(hipMemcpyAsync(&f, fault, 4, hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	if (f) {
		set_error("ray cast: tree traversal exceeded its bounds");
		return B3G_HIP_ERROR;
	}
	return B3G_OK;
}

} // namespace

b3g_status ray_cast(Body& b, uint32_t inst, uint32_t count, const float* p1, const float* p2, int64_t* index,
                    float* fraction, float* normal) {
	if (count == 0) return B3G_OK;
	b3g_status s;
	// staging: p1, p2 (3 floats per ray each); index (int64), fraction,
	// normal (3 floats), then the fault word
	if (b.rtRayCap < count) {
		if ((s = dalloc(b, b.rtRays, (size_t)count * 6 * 4)) != B3G_OK) return s;
		if ((s = dalloc(b, b.rtHits, (size_t)count * 24 + 4)) != B3G_OK) return s;
		b.rtRayCap = count;
	}
	float* dp1 = dptr<float>(b.rtRays);
	float* dp2 = dp1 + 3 * (size_t)count;
	int64_t* di = dptr<int64_t>(b.rtHits);
	float* df = reinterpret_cast<float*>(di + count);
	float* dn = df + count;
	uint32_t* fault = reinterpret_cast<uint32_t*>(dn + 3 * (size_t)count);
	B3G_HIP(hipMemcpyAsync(dp1, p1, (size_t)count * 12, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(dp2, p2, (size_t)count * 12, hipMemcpyHostToDevice, b.stream));
	if ((s = cast_rays(b, inst, count, dp1, dp2, di, df, normal ? dn : nullptr, fault)) != B3G_OK) return s;
	B3G_HIP(hipMemcpyAsync(index, di, (size_t)count * 8, hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipMemcpyAsync(fraction, df, (size_t)count * 4, hipMemcpyDeviceToHost, b.stream));
	if (normal) B3G_HIP(hipMemcpyAsync(normal, dn, (size_t)count * 12, hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	return B3G_OK;
}

b3g_status ray_cast_device(Body& b, uint32_t inst, uint32_t count, const float* p1, const float* p2, int64_t* index,
                           float* fraction, float* normal) {
	if (count == 0) return B3G_OK;
	b3g_status s;
	if (!b.rtFault.p && (s = dalloc(b, b.rtFault, 4)) != B3G_OK) return s;
	return cast_rays(b, inst, count, p1, p2, index, fraction, normal, dptr<uint32_t>(b.rtFault));
}

} // namespace b3g
