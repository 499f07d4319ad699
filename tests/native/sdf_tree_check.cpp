// Checks the tree walk of the device SDF builder (k_sdf_tree in
// bounce_softbody_amd/csrc/b3g_sdf.hip, its host/device functions compiled
// for the host) against the brute-force loop over every triangle (k_sdf's),
// voxel by voxel, bit for bit. Usage: sdf_tree_check CELLS mesh.bin...
// (mesh.bin: uint32 nv, nt, stride; float vertices[3 nv]; uint32 triangles[stride nt]).
#include "b3g_sdf.hip"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace b3g {
void set_error(const std::string&) {}
}
using namespace b3g;

int main(int argc, char** argv) {
	if (argc < 3) return 2;
	const float cells = (float)atof(argv[1]);
	long total = 0;
	for (int m = 2; m < argc; ++m) {
		FILE* f = fopen(argv[m], "rb");
		if (!f) return 2;
		uint32_t hdr[3];
		if (fread(hdr, 4, 3, f) != 3) return 2;
		std::vector<float> v(3 * (size_t)hdr[0]);
		std::vector<uint32_t> t((size_t)hdr[1] * hdr[2]);
		if (fread(v.data(), 4, v.size(), f) != v.size() || fread(t.data(), 4, t.size(), f) != t.size()) return 2;
		fclose(f);
		const uint32_t nv = hdr[0], nt = hdr[1], stride = hdr[2];
		float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f}, cell[3];
		for (uint32_t i = 0; i < nv; ++i)
			for (int k = 0; k < 3; ++k) {
				lo[k] = std::min(lo[k], v[3 * i + k]);
				hi[k] = std::max(hi[k], v[3 * i + k]);
			}
		for (int k = 0; k < 3; ++k) cell[k] = (hi[k] - lo[k] + 0.4f) / cells;
		AABB box;
		uint32_t dims[3];
		if (sdf_grid(v.data(), nv, cell, 0.2f, box, dims) != B3G_OK) return 2;
		std::vector<float4> rec = sdf_records(v.data(), t.data(), nt, stride);
		SdfArgs a = sdf_args(box, dims, nt);
		a.tri = rec.data();
		const V3 ext = box.hi - box.lo;
		const float pad = kLinearSlop + 1.0e-5f * (fabsf(a.far) + ext.x + ext.y + ext.z);   // as build_sdf
		MeshTrees tree;
		build_mesh_tree(v.data(), t.data(), nt, tree, stride, pad);
		if (tree_depth(tree.child, nt) + 2 > (uint32_t)kSdfStack) return 3;
		long bad = 0, visits = 0, count = 0;
		for (uint32_t z = 0; z < dims[2]; ++z)
			for (uint32_t y = 0; y < dims[1]; ++y)
				for (uint32_t x = 0; x < dims[0]; ++x) {
					SdfVoxel s0;
					sdf_begin(a, x, y, z, s0);
					for (uint32_t k = 0; k < nt; ++k) sdf_triangle(&rec[4 * (size_t)k], s0);
					const float ref = sdf_end(s0);
					// the walk of k_sdf_tree
					SdfVoxel s;
					sdf_begin(a, x, y, z, s);
					V3 inv[5];
					for (int r = 0; r < 5; ++r)
						inv[r] = v3(s.d[r].x != 0.0f ? 1.0f / s.d[r].x : 0.0f, s.d[r].y != 0.0f ? 1.0f / s.d[r].y : 0.0f,
						            s.d[r].z != 0.0f ? 1.0f / s.d[r].z : 0.0f);
					uint32_t stack[kSdfStack];
					int sp = 0;
					stack[sp++] = nt == 1 ? kRtLeaf : 0u;
					while (sp > 0) {
						const uint32_t c = stack[--sp];
						const bool leaf = (c & kRtLeaf) != 0;
						const RtBox bx = leaf ? tree.leafBox[c & ~kRtLeaf] : tree.nodeBox[c];
						const bool dist = box_dist2(bx, s.P) <= s.best;
						uint32_t rays = 0;
						for (int r = 0; r < 5; ++r)
							if (s.live[r] && sdf_slab(bx, s.P, s.d[r], inv[r])) rays |= 1u << r;
						if (!dist && !rays) continue;
						if (leaf) {
							++visits;
							sdf_triangle(&rec[4 * (size_t)tree.leaf[c & ~kRtLeaf]], s, dist, rays);
						} else {
							stack[sp++] = tree.child[c].y;
							stack[sp++] = tree.child[c].x;
						}
					}
					const float got = sdf_end(s);
					++count;
					if (std::memcmp(&ref, &got, 4) != 0) ++bad;
				}
		printf("%s: %u triangles, %u x %u x %u voxels, mismatches %ld, leaf visits per voxel %.1f\n", argv[m], nt, dims[0],
		       dims[1], dims[2], bad, double(visits) / count);
		total += bad;
	}
	printf("total mismatches %ld\n", total);
	return total ? 1 : 0;
}
