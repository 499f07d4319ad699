// Checks the device's recomputed tetrahedron stiffness blocks (b3g_tet.h,
// compiled here for the host) against the oracle's dense restatement of
// ResetElementData (tetrahedron_element_force.cpp:219-265) bit for bit, on
// seeded random elements: jittered cells, axis-aligned corners (exact and
// signed zero gradients), flat slivers and wide scale / material ranges.
// Prints "checked N mismatches M" and exits non-zero on a mismatch.
#include "b3g_tet.h"
#include "cpu_ref.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static bool same(float a, float b) { return std::memcmp(&a, &b, sizeof(float)) == 0; }

int main() {
	std::mt19937 rng(1234);
	std::uniform_real_distribution<float> u(-1.0f, 1.0f);
	long checked = 0, bad = 0;
	for (int n = 0; n < 200000; ++n) {
		b3o::TetForce<float> t;
		const int kind = n % 4;
		const float s = std::ldexp(1.0f, (int)(rng() % 12) - 6);
		const b3o::V3<float> base = {u(rng), u(rng), u(rng)};
		const float corner[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
		for (int k = 0; k < 4; ++k) {
			b3o::V3<float> x;
			if (kind == 0) {        // unit corner plus jitter
				x = {corner[k][0] + 0.05f * u(rng), corner[k][1] + 0.05f * u(rng), corner[k][2] + 0.05f * u(rng)};
			} else if (kind == 1) { // exact axis-aligned corner: zero gradient entries
				x = {corner[k][0], corner[k][1], corner[k][2]};
			} else if (kind == 2) { // random
				x = {u(rng), u(rng), u(rng)};
			} else {                // sliver: nearly flat in z
				x = {u(rng), u(rng), k == 3 ? 1e-3f * u(rng) : 0.0f};
			}
			t.x[k] = {s * x.x + base.x, s * x.y + base.y, s * x.z + base.z};
		}
		const b3o::V3<float> e1 = t.x[1] - t.x[0], e2 = t.x[2] - t.x[0], e3 = t.x[3] - t.x[0];
		const float det = b3o::Det(e1, e2, e3);
		if (det == 0.0f || !std::isfinite(det)) continue;
		t.E = std::pow(10.0f, 3.0f + 5.0f * (0.5f + 0.5f * u(rng)));
		t.nu = 0.45f * (0.5f + 0.5f * u(rng));
		b3o::TetReset(t);
		const float V = b3o::Abs(det) / 6.0f;
		b3g::M33 iE;
		iE.x = b3g::v3(t.invE.x.x, t.invE.x.y, t.invE.x.z);
		iE.y = b3g::v3(t.invE.y.x, t.invE.y.y, t.invE.y.z);
		iE.z = b3g::v3(t.invE.z.x, t.invE.z.y, t.invE.z.z);
		float rec[b3g::kTetRecFloats];
		b3g::tet_record(iE, t.E, t.nu, V, rec);
		const b3g::M33 back = b3g::tet_inv_e(rec);
		const float* pb = &back.x.x;
		const float* pi = &iE.x.x;
		for (int c = 0; c < 9; ++c) bad += !same(pb[c], pi[c]);
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) {
				const b3g::M33 k = b3g::tet_k0(rec, i, j);
				const b3o::M33<float>& r = t.K[i + 4 * j];
				const float a[9] = {k.x.x, k.x.y, k.x.z, k.y.x, k.y.y, k.y.z, k.z.x, k.z.y, k.z.z};
				const float b[9] = {r.x.x, r.x.y, r.x.z, r.y.x, r.y.y, r.y.z, r.z.x, r.z.y, r.z.z};
				for (int c = 0; c < 9; ++c) {
					++checked;
					if (!same(a[c], b[c])) {
						if (bad < 5) std::printf("element %d block (%d,%d) entry %d: %a vs %a\n", n, i, j, c, a[c], b[c]);
						++bad;
					}
				}
			}
	}
	std::printf("checked %ld mismatches %ld\n", checked, bad);
	return bad ? 1 : 0;
}
