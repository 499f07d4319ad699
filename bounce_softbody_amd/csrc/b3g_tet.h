// Tetrahedron stiffness blocks K0_ij = V (B^T D B)_ij recomputed from a
// 16-float rest record instead of stored (144 floats per element).
//
// The reference forms K0 once per element with dense products
// (tetrahedron_element_force.cpp:219-265): BT_D = B^T D (b3Mul, math.h:135-151),
// then K = BT_D B, then K *= V. b3Mul starts every entry at +0 and adds the
// products in index order. Such an accumulator is never -0 (+0 + -0 = +0 and
// an exact cancellation gives +0), so adding a product that is +-0 leaves it
// unchanged. The structural zeros of B (b3ComputeB, :121-159) and D
// (b3ComputeD, :37-57) therefore drop out, and every entry of K0_ij reduces
// to two or three products of the shape-function gradients g_i, g_j with
// lambda + 2 mu, lambda and mu, summed from +0 in the reference's order. The
// results equal the reference's dense products bit for bit.
#pragma once

#include "b3g_math.h"

namespace b3g {

// Record layout (16 floats per element): g_0, g_1, g_2, g_3 (12 floats: the
// shape-function gradients, g_0 = -g_1 - g_2 - g_3 as the reference forms
// it), then V, lambda + 2 mu, lambda, mu.
constexpr int kTetRecFloats = 16;
constexpr int kTetRecV = 12, kTetRecL2 = 13, kTetRecL = 14, kTetRecMu = 15;

// The record from invE (columns as in M33), Young's modulus, Poisson's ratio
// and the volume, with the reference's operations.
B3G_HD void tet_record(const M33& iE, float E, float nu, float V, float* rec) {
	const float lambda = (nu * E) / ((1.0f + nu) * (1.0f - 2.0f * nu));
	const float mu = E / (2.0f * (1.0f + nu));
	// dN2/dx = invE.x.x, dN3/dx = invE.x.y, dN4/dx = invE.x.z, ...
	const V3 g1 = v3(iE.x.x, iE.y.x, iE.z.x);
	const V3 g2 = v3(iE.x.y, iE.y.y, iE.z.y);
	const V3 g3 = v3(iE.x.z, iE.y.z, iE.z.z);
	const V3 g0 = v3(-g1.x - g2.x - g3.x, -g1.y - g2.y - g3.y, -g1.z - g2.z - g3.z);
	const V3 g[4] = {g0, g1, g2, g3};
	for (int v = 0; v < 4; ++v) {
		rec[3 * v] = g[v].x;
		rec[3 * v + 1] = g[v].y;
		rec[3 * v + 2] = g[v].z;
	}
	rec[kTetRecV] = V;
	rec[kTetRecL2] = lambda + 2.0f * mu;
	rec[kTetRecL] = lambda;
	rec[kTetRecMu] = mu;
}

// invE back from the record (the gradients are its rows; no arithmetic).
B3G_HD M33 tet_inv_e(const float* rec) {
	return m33(v3(rec[3], rec[6], rec[9]), v3(rec[4], rec[7], rec[10]), v3(rec[5], rec[8], rec[11]));
}

// K0 block (i, j) (m_K[i + 4 j], rows of vertex i, columns of vertex j) from
// p = g_i, q = g_j. BT_D rows of vertex i (BT_D(3i + a, k), each a single
// product added to +0):
//   a = 0: [px L2, px L, px L, py mu, 0, pz mu]
//   a = 1: [py L, py L2, py L, px mu, pz mu, 0]
//   a = 2: [pz L, pz L, pz L2, 0, py mu, px mu]
// B columns of vertex j: [qx 0 0 qy 0 qz], [0 qy 0 qx qz 0], [0 0 qz 0 qy qx].
//
// Each BT_D entry is its product added to +0, which changes the product only
// when it is -0. Every BT_D entry is then only multiplied by a component of q
// and added into a K sum that starts from +0 and so is never -0; adding a
// zero of either sign to such a sum gives the same result, so the sign of a
// zero BT_D entry never reaches K and its +0 addition is left out.
B3G_HD M33 tet_k0(V3 p, V3 q, float V, float L2, float L, float mu) {
	const float z = 0.0f;
	// BT_D entries (see above: the +0 additions drop out)
	const float a0 = p.x * L2, a1 = p.x * L, a3 = p.y * mu, a5 = p.z * mu;
	const float b0 = p.y * L, b1 = p.y * L2, b3 = p.x * mu, b4 = p.z * mu;
	const float c0 = p.z * L, c2 = p.z * L2, c4 = p.y * mu, c5 = p.x * mu;
	// K(a, b), summed in k order
	const float k00 = ((z + a0 * q.x) + a3 * q.y) + a5 * q.z;
	const float k01 = (z + a1 * q.y) + a3 * q.x;
	const float k02 = (z + a1 * q.z) + a5 * q.x;
	const float k10 = (z + b0 * q.x) + b3 * q.y;
	const float k11 = ((z + b1 * q.y) + b3 * q.x) + b4 * q.z;
	const float k12 = (z + b0 * q.z) + b4 * q.y;
	const float k20 = (z + c0 * q.x) + c5 * q.z;
	const float k21 = (z + c0 * q.y) + c4 * q.z;
	const float k22 = ((z + c2 * q.z) + c4 * q.y) + c5 * q.x;
	// columns of M33: column b holds K(0, b), K(1, b), K(2, b)
	return m33(v3(k00 * V, k10 * V, k20 * V), v3(k01 * V, k11 * V, k21 * V), v3(k02 * V, k12 * V, k22 * V));
}

B3G_HD V3 tet_grad(const float* rec, int v) { return v3(rec[3 * v], rec[3 * v + 1], rec[3 * v + 2]); }

B3G_HD M33 tet_k0(const float* rec, int i, int j) {
	return tet_k0(tet_grad(rec, i), tet_grad(rec, j), rec[kTetRecV], rec[kTetRecL2], rec[kTetRecL], rec[kTetRecMu]);
}

} // namespace b3g
