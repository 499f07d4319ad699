// Host-side build of a body: list ordering, mass lumping, element rest data,
// world shapes and the symbolic pattern of the linear system.
//
// The symbolic pattern reproduces the order in which the reference's linked
// lists hold the blocks of pA, so that the device sums every row in the
// reference's order:
//   * dfdx / dfdv rows are filled by ApplyForces (force_solver.cpp:70-86):
//     particles, then forces in list order, then contacts; an entry is
//     inserted at the list head on first touch (sparse_mat33.h:268-288).
//     Write D_i / V_i for the head-to-tail lists (reverse first touch).
//   * A = M - h dfdv - (h h) dfdx and pA = S A S^T + I - S
//     (sparse_force_solver.cpp:165-173) copy, reverse and extend those lists
//     (sparse_mat33.h:210-230, 452-492); the resulting row of pA is
//         (D_i \ (V_i u {i})) ++ [i] ++ reverse(V_i \ {i}).
//   * b's product h dfdx (x0 - x + h v + y) walks reverse(D_i), i.e. the
//     first-touch order of dfdx.
// Each pA block keeps the list of staged element contributions that the
// reference adds into it, in the reference's order.
#include "b3g_bvh.h"
#include "b3g_internal.h"
#include "b3g_tet.h"

#include <algorithm>
#include <functional>
#include <cstring>
#include <numeric>

namespace b3g {

b3g_status dalloc(Body& b, DBuf& buf, size_t bytes) {
	if (buf.p) {
		(void)hipFree(buf.p);
		buf.p = nullptr;
		buf.bytes = 0;
	}
	if (bytes == 0) return B3G_OK;
	hipError_t e = hipMalloc(&buf.p, bytes);
	if (e != hipSuccess) {
		buf.p = nullptr;
		set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
		return e == hipErrorOutOfMemory ? B3G_OUT_OF_MEMORY : B3G_HIP_ERROR;
	}
	buf.bytes = bytes;
	if (std::find(b.owned.begin(), b.owned.end(), &buf) == b.owned.end()) b.owned.push_back(&buf);
	return B3G_OK;
}

b3g_status dupload(Body& b, DBuf& buf, const void* host, size_t bytes) {
	b3g_status s = dalloc(b, buf, bytes);
	if (s != B3G_OK) return s;
	// Synchronous: set-up sources are temporaries that die right after this call.
	if (bytes) {
		B3G_HIP(hipStreamSynchronize(b.stream));
		B3G_HIP(hipMemcpy(buf.p, host, bytes, hipMemcpyHostToDevice));
	}
	return B3G_OK;
}

void dfree_all(Body& b) {
	for (DBuf* d : b.owned) {
		if (d->p) (void)hipFree(d->p);
		d->p = nullptr;
		d->bytes = 0;
	}
	b.owned.clear();
}

namespace {

V3 load3(const float* p) { return v3(p[0], p[1], p[2]); }

// Planar upload of an (m x k) row-major float table as k planes of m values.
template <class T>
std::vector<T> planes(const std::vector<T>& rows, uint32_t m, uint32_t k) {
	std::vector<T> out((size_t)m * k);
	for (uint32_t e = 0; e < m; ++e)
		for (uint32_t c = 0; c < k; ++c) out[(size_t)c * m + e] = rows[(size_t)e * k + c];
	return out;
}

// triangle_element_force.cpp:30-50
M33 compute_c(float Ex, float Ey, float Es, float nu_xy, float nu_yx) {
	float s = 1.0f - nu_xy * nu_yx;
	M33 C;
	C.x.x = Ex / s; C.x.y = Ey * nu_xy / s; C.x.z = 0.0f;
	C.y.x = Ex * nu_yx / s; C.y.y = Ey / s; C.y.z = 0.0f;
	C.z.x = 0.0f; C.z.y = 0.0f; C.z.z = Es;
	return C;
}

// triangle_element_force.cpp:208-279 (ResetElementData); returns false when det == 0.
bool tri_reset(V3 p1, V3 p2, V3 p3, const float prm[6], V2 x[3], M22& invS, float K0[36]) {
	V3 n = cross(p2 - p1, p3 - p1);
	normalize_inplace(n);
	V3 px, py;
	compute_basis(n, px, py);
	float P[6] = {px.x, py.x, px.y, py.y, px.z, py.z};
	gmul<2, 3, 1>(&x[0].x, P, &p1.x);
	gmul<2, 3, 1>(&x[1].x, P, &p2.x);
	gmul<2, 3, 1>(&x[2].x, P, &p3.x);
	V2 e1 = x[1] - x[0], e2 = x[2] - x[0];
	M22 S;
	S.x = e1;
	S.y = e2;
	float d = det(e1, e2);
	if (d == 0.0f) return false;
	float A = 0.5f * fabs_(d);
	invS = d * adjugate(S);   // the reference scales by det, not 1/det (triangle_element_force.cpp:252)
	M33 C = compute_c(prm[0], prm[1], prm[2], prm[3], prm[4]);
	float dN2dx = invS.x.x, dN3dx = invS.x.y, dN1dx = -dN2dx - dN3dx;
	float dN2dy = invS.y.x, dN3dy = invS.y.y, dN1dy = -dN2dy - dN3dy;
	float B[18] = {dN1dx, 0, dN1dy, 0, dN1dy, dN1dx, dN2dx, 0, dN2dy, 0, dN2dy, dN2dx, dN3dx, 0, dN3dy, 0, dN3dy, dN3dx};
	float BT[18], BT_C[18], K[36];
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 6; ++j) BT[j + 6 * i] = B[i + 3 * j];
	gmul<6, 3, 3>(BT_C, BT, &C.x.x);
	gmul<6, 3, 6>(K, BT_C, B);
	for (int i = 0; i < 36; ++i) K[i] *= A;
	// b3SetK (triangle_element_force.cpp:152-181): block (i0 + 3 j0), element (ii + 2 jj)
	for (int i = 0; i < 6; ++i)
		for (int j = 0; j < 6; ++j) K0[((i / 2) + 3 * (j / 2)) * 4 + (i % 2) + 2 * (j % 2)] = K[i + 6 * j];
	return true;
}

// stretch_force.cpp:62-107, shear_force.cpp:51-92: returns 11 coefficients
// alpha du1 dv1 du2 dv2 inv_det dwudx(3) dwvdx(3) -> 12 with padding
void uv_reset(const float uv[6], float out[12]) {
	float u1 = uv[0], v1 = uv[1], u2 = uv[2], v2 = uv[3], u3 = uv[4], v3_ = uv[5];
	float du1 = u2 - u1, dv1 = v2 - v1, du2 = u3 - u1, dv2 = v3_ - v1;
	float d = du1 * dv2 - du2 * dv1;
	float inv_det = 1.0f / d;
	out[0] = 0.5f * fabs_(d);
	out[1] = du1; out[2] = dv1; out[3] = du2; out[4] = dv2;
	out[5] = inv_det;
	out[6] = inv_det * (dv1 - dv2);
	out[7] = inv_det * dv2;
	out[8] = -inv_det * dv1;
	out[9] = inv_det * (du2 - du1);
	out[10] = -inv_det * du2;
	out[11] = inv_det * du1;
}

// ----------------------------------------------------------------------------
// Symbolic pattern builder
// ----------------------------------------------------------------------------
struct Touch {
	uint32_t row;
	uint32_t stage;   // staged contribution index, or kNoStage
	uint16_t local;   // first-touch index of the column within the row
};
constexpr uint32_t kNoStage = 0xFFFFFFFFu;

struct RowLists {
	std::vector<std::vector<uint32_t>> cols;   // per row: columns in first-touch order
	std::vector<Touch> touches;
	explicit RowLists(uint32_t n) : cols(n) {}
	void touch(uint32_t i, uint32_t j, uint32_t stage) {
		std::vector<uint32_t>& c = cols[i];
		uint16_t k = 0;
		for (; k < c.size(); ++k)
			if (c[k] == j) break;
		if (k == c.size()) c.push_back(j);
		touches.push_back({i, stage, k});
	}
};

struct Pattern {
	std::vector<uint32_t> rowptr, cols, slotRow, diagSlot;
	std::vector<uint32_t> jxPtr, jxIdx, jvPtr, jvIdx, bPtr, bSlot, fPtr, fIdx;
};

// Walks the touch sequence of ApplyForces for the body and derives pA's layout.
struct SymbolicInput {
	uint32_t n;
	const std::vector<uint8_t>* type;
	const std::vector<float>* damping;
	const std::vector<int32_t>* sphereOf;
	const std::vector<char>* contactRow;   // per row: the row's particle can touch a world fixture (or null)
	// forces in list order: (type, creation index within the type)
	std::vector<std::pair<uint32_t, uint32_t>> forces;
	// particle list indices of each element, by type (creation order)
	const std::vector<uint32_t>* tetP; uint32_t tetM; bool tetDamp; const std::vector<float>* tetKd;
	const std::vector<uint32_t>* triP; uint32_t triM; const std::vector<float>* triKd;
	const std::vector<uint32_t>* stP; uint32_t stM; const std::vector<float>* stCoef;   // 20 coefs / element
	const std::vector<uint32_t>* shP; uint32_t shM; const std::vector<float>* shCoef;   // 16
	const std::vector<uint32_t>* spP; uint32_t spM; const std::vector<float>* spCoef;   // 3
	const std::vector<uint32_t>* moP; uint32_t moM; const std::vector<float>* moCoef;   // 6
	// staging bases
	uint32_t tetF, tetFd, triJx, triJv, triF, triFd;
	uint32_t stF, shF, spJx, spJv, spF, moJx, moJv, moF;
};

// A stretch / shear block rebuilt in the assembly (variant: see kUVCode).
inline uint32_t uv_code(uint32_t variant, uint32_t e) { return kUVCode | (variant << kUVShift) | e; }

// A batch is merged into one description with 32-bit counts: every count,
// summed over the instances, must fit.
bool batch_limits(const b3g_body_desc* ds, uint32_t count) {
	uint64_t t[12] = {};
	for (uint32_t k = 0; k < count; ++k) {
		const b3g_body_desc& d = ds[k];
		const uint32_t c[12] = {d.n_particles, d.n_spheres, d.n_trifix, d.n_tetfix, d.n_forces, d.n_tet,
		                        d.n_tri, d.n_stretch, d.n_shear, d.n_spring, d.n_mouse, d.n_world};
		for (int i = 0; i < 12; ++i) t[i] += c[i];
	}
	for (int i = 0; i < 12; ++i)
		if (t[i] > 0xFFFFFFFFull) {
			set_error("a count summed over the batch's instances exceeds 2^32 - 1");
			return false;
		}
	return true;
}

// Element counts the 32-bit contribution codes can address: 16 blocks per
// tetrahedron in 31 bits, stretch / shear indices in kUVShift bits, staged
// blocks (triangle elements, springs, mouse) below kUVCode.
bool code_limits(const b3g_body_desc& d) {
	if (d.n_tet >= kTetMaxElements) {
		set_error("more than 2^27 - 1 tetrahedron elements");
		return false;
	}
	if (d.n_stretch >= kUVMaxElements || d.n_shear >= kUVMaxElements) {
		set_error("more than 2^25 - 1 stretch or shear forces");
		return false;
	}
	const uint64_t staged = 18ull * d.n_tri + 8ull * d.n_spring + 32ull * d.n_mouse;
	if (staged >= kUVCode) {
		set_error("more than 2^30 staged Jacobian blocks (triangle elements, springs, mouse forces)");
		return false;
	}
	return true;
}

bool build_pattern(const SymbolicInput& in, Pattern& P) {
	const uint32_t n = in.n;
	RowLists D(n), V(n);
	std::vector<std::pair<uint32_t, uint32_t>> fTouch;   // (row, stage) in order

	// particles (particle.cpp:182-199): mass damping touches dfdv(i,i)
	for (uint32_t i = 0; i < n; ++i)
		if ((*in.damping)[i] > 0.0f) V.touch(i, i, kNoStage);

	for (const auto& fr : in.forces) {
		const uint32_t t = fr.first, e = fr.second;
		switch (t) {
		case B3G_FORCE_TET_ELEMENT: {   // tetrahedron_element_force.cpp:378-463
			const uint32_t* p = in.tetP->data() + 4 * (size_t)e;
			const uint32_t m = in.tetM;
			for (int i = 0; i < 4; ++i)
				for (int j = 0; j < 4; ++j) D.touch(p[i], p[j], kTetCode | ((uint32_t)(i + 4 * j) * m + e));
			for (int i = 0; i < 4; ++i) fTouch.push_back({p[i], in.tetF + (uint32_t)i * m + e});
			if ((*in.tetKd)[e] > 0.0f) {
				for (int i = 0; i < 4; ++i) fTouch.push_back({p[i], in.tetFd + (uint32_t)i * m + e});
				for (int i = 0; i < 4; ++i)
					for (int j = 0; j < 4; ++j) V.touch(p[i], p[j], kTetCode | ((uint32_t)(i + 4 * j) * m + e));
			}
			break;
		}
		case B3G_FORCE_TRI_ELEMENT: {   // triangle_element_force.cpp:379-467
			const uint32_t* p = in.triP->data() + 3 * (size_t)e;
			const uint32_t m = in.triM;
			for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.triF + (uint32_t)i * m + e});
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j) D.touch(p[i], p[j], in.triJx + (uint32_t)(i + 3 * j) * m + e);
			if ((*in.triKd)[e] > 0.0f) {
				for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.triFd + (uint32_t)i * m + e});
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j) V.touch(p[i], p[j], in.triJv + (uint32_t)(i + 3 * j) * m + e);
			}
			break;
		}
		case B3G_FORCE_STRETCH: {       // stretch_force.cpp:161-388, parts u then v
			const uint32_t* p = in.stP->data() + 3 * (size_t)e;
			const float* c = in.stCoef->data() + 20 * (size_t)e;
			const uint32_t m = in.stM;
			for (int part = 0; part < 2; ++part) {
				float ks = c[12 + 3 * part], kd = c[13 + 3 * part];
				if (ks > 0.0f) {
					for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.stF + (uint32_t)(6 * part + i) * m + e});
					for (int i = 0; i < 3; ++i)
						for (int j = 0; j < 3; ++j) D.touch(p[i], p[j], uv_code(9 * part + 3 * i + j, e));
				}
				if (kd > 0.0f) {
					for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.stF + (uint32_t)(6 * part + 3 + i) * m + e});
					for (int i = 0; i < 3; ++i)
						for (int j = 0; j < 3; ++j) V.touch(p[i], p[j], uv_code(9 * part + 3 * i + j, e));
				}
			}
			break;
		}
		case B3G_FORCE_SHEAR: {         // shear_force.cpp:150-244
			const uint32_t* p = in.shP->data() + 3 * (size_t)e;
			const float* c = in.shCoef->data() + 16 * (size_t)e;
			const uint32_t m = in.shM;
			if (c[12] > 0.0f) {
				for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.shF + (uint32_t)i * m + e});
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j) D.touch(p[i], p[j], uv_code(kShearVariant + 3 * i + j, e));
			}
			if (c[13] > 0.0f) {
				for (int i = 0; i < 3; ++i) fTouch.push_back({p[i], in.shF + (uint32_t)(3 + i) * m + e});
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j) V.touch(p[i], p[j], uv_code(kShearVariant + 3 * i + j, e));
			}
			break;
		}
		case B3G_FORCE_SPRING: {        // spring_force.cpp:84-142 (tension part assumed active)
			const uint32_t* p = in.spP->data() + 2 * (size_t)e;
			const float* c = in.spCoef->data() + 3 * (size_t)e;
			const uint32_t m = in.spM;
			const uint32_t pr[4][2] = {{p[0], p[0]}, {p[0], p[1]}, {p[1], p[0]}, {p[1], p[1]}};
			if (c[1] > 0.0f) {
				for (int i = 0; i < 2; ++i) fTouch.push_back({p[i], in.spF + (uint32_t)i * m + e});
				for (int k = 0; k < 4; ++k) D.touch(pr[k][0], pr[k][1], in.spJx + (uint32_t)k * m + e);
			}
			if (c[2] > 0.0f) {
				for (int i = 0; i < 2; ++i) fTouch.push_back({p[i], in.spF + (uint32_t)(2 + i) * m + e});
				for (int k = 0; k < 4; ++k) V.touch(pr[k][0], pr[k][1], in.spJv + (uint32_t)k * m + e);
			}
			break;
		}
		case B3G_FORCE_MOUSE: {         // mouse_force.cpp:92-245
			const uint32_t* p = in.moP->data() + 4 * (size_t)e;
			const float* c = in.moCoef->data() + 6 * (size_t)e;
			const uint32_t m = in.moM;
			if (c[4] > 0.0f) {
				for (int i = 0; i < 4; ++i) fTouch.push_back({p[i], in.moF + (uint32_t)i * m + e});
				for (int i = 0; i < 4; ++i)
					for (int j = 0; j < 4; ++j) D.touch(p[i], p[j], in.moJx + (uint32_t)(4 * i + j) * m + e);
			}
			if (c[5] > 0.0f) {
				for (int i = 0; i < 4; ++i) fTouch.push_back({p[i], in.moF + (uint32_t)(4 + i) * m + e});
				for (int i = 0; i < 4; ++i)
					for (int j = 0; j < 4; ++j) V.touch(p[i], p[j], in.moJv + (uint32_t)(4 * i + j) * m + e);
			}
			break;
		}
		default: break;
		}
	}
	// contacts (sphere_shape_contact.cpp:121-122): dfdx(i,i) of particles that can touch a world fixture
	if (in.contactRow)
		for (uint32_t i = 0; i < n; ++i)
			if ((*in.contactRow)[i] && (*in.sphereOf)[i] >= 0 && (*in.type)[i] == kDynamic) D.touch(i, i, kNoStage);

	// Layout of pA rows
	P.rowptr.assign(n + 1, 0);
	std::vector<std::vector<uint32_t>> rowCols(n);
	for (uint32_t i = 0; i < n; ++i) {
		const std::vector<uint32_t>& dft = D.cols[i];   // first-touch order
		const std::vector<uint32_t>& vft = V.cols[i];
		std::vector<uint32_t>& rc = rowCols[i];
		// (D_i \ (V_i u {i})) in head-to-tail order = reverse first touch
		for (size_t k = dft.size(); k-- > 0;) {
			uint32_t c = dft[k];
			if (c == i) continue;
			if (std::find(vft.begin(), vft.end(), c) != vft.end()) continue;
			rc.push_back(c);
		}
		rc.push_back(i);
		// reverse(V_i \ {i}) with V_i head-to-tail = first-touch order of dfdv
		for (uint32_t c : vft)
			if (c != i) rc.push_back(c);
		// block indices are 32-bit and the planes of A are padded to a multiple of 4
		const uint64_t end = (uint64_t)P.rowptr[i] + rc.size();
		if (end > 0xFFFFFFFCull) {
			set_error("more than 2^32 - 4 blocks in pA");
			return false;
		}
		P.rowptr[i + 1] = (uint32_t)end;
	}
	const uint32_t nnz = P.rowptr[n];
	P.cols.resize(nnz);
	P.slotRow.resize(nnz);
	P.diagSlot.resize(n);
	for (uint32_t i = 0; i < n; ++i) {
		for (size_t k = 0; k < rowCols[i].size(); ++k) {
			uint32_t s = P.rowptr[i] + (uint32_t)k;
			P.cols[s] = rowCols[i][k];
			P.slotRow[s] = i;
			if (rowCols[i][k] == i) P.diagSlot[i] = s;
		}
	}
	// Map a (row, first-touch local index) of D / V to its pA slot.
	auto slotOf = [&](const RowLists& L, uint32_t i, uint16_t local) -> uint32_t {
		uint32_t c = L.cols[i][local];
		const std::vector<uint32_t>& rc = rowCols[i];
		for (size_t k = 0; k < rc.size(); ++k)
			if (rc[k] == c) return P.rowptr[i] + (uint32_t)k;
		return 0xFFFFFFFFu;
	};
	auto lists = [&](const RowLists& L, std::vector<uint32_t>& ptr, std::vector<uint32_t>& idx) {
		std::vector<uint32_t> slotOfTouch(L.touches.size());
		ptr.assign(nnz + 1, 0);
		// per-row cache of local -> slot
		std::vector<std::vector<uint32_t>> cache(n);
		for (uint32_t i = 0; i < n; ++i) {
			cache[i].resize(L.cols[i].size());
			for (uint16_t k = 0; k < L.cols[i].size(); ++k) cache[i][k] = slotOf(L, i, k);
		}
		for (size_t t = 0; t < L.touches.size(); ++t) {
			const Touch& tc = L.touches[t];
			slotOfTouch[t] = cache[tc.row][tc.local];
			if (tc.stage != kNoStage) ptr[slotOfTouch[t] + 1]++;
		}
		for (uint32_t s = 0; s < nnz; ++s) ptr[s + 1] += ptr[s];
		idx.resize(ptr[nnz]);
		std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
		for (size_t t = 0; t < L.touches.size(); ++t)
			if (L.touches[t].stage != kNoStage) idx[fill[slotOfTouch[t]]++] = L.touches[t].stage;
		return cache;
	};
	std::vector<std::vector<uint32_t>> dcache = lists(D, P.jxPtr, P.jxIdx);
	lists(V, P.jvPtr, P.jvIdx);
	// b's product over reverse(D_i) = first-touch order
	P.bPtr.assign(n + 1, 0);
	for (uint32_t i = 0; i < n; ++i) P.bPtr[i + 1] = P.bPtr[i] + (uint32_t)D.cols[i].size();
	P.bSlot.resize(P.bPtr[n]);
	for (uint32_t i = 0; i < n; ++i)
		for (size_t k = 0; k < D.cols[i].size(); ++k) P.bSlot[P.bPtr[i] + k] = dcache[i][k];
	// force contributions per row, reference order
	P.fPtr.assign(n + 1, 0);
	for (auto& ft : fTouch) P.fPtr[ft.first + 1]++;
	for (uint32_t i = 0; i < n; ++i) P.fPtr[i + 1] += P.fPtr[i];
	P.fIdx.resize(fTouch.size());
	std::vector<uint32_t> fill(P.fPtr.begin(), P.fPtr.end() - 1);
	for (auto& ft : fTouch) P.fIdx[fill[ft.first]++] = ft.second;
	return true;
}

// ComputeAABB of each world shape (shapes/*.cpp), static over the body's life.
AABB shape_aabb(const WorldShape& w) {
	switch (w.type) {
	case B3G_SHAPE_SPHERE: return aabb_cr(w.a, w.radius);                      // sphere_shape.cpp:34-43
	case B3G_SHAPE_CAPSULE: {                                                  // capsule_shape.cpp:34-44
		V3 r = v3(w.radius, w.radius, w.radius);
		AABB a; a.lo = vmin(w.a, w.b) - r; a.hi = vmax(w.a, w.b) + r; return a;
	}
	case B3G_SHAPE_TRIANGLE: {                                                 // triangle_shape.cpp:52-63
		V3 r = v3(w.radius, w.radius, w.radius);
		AABB a; a.lo = vmin(w.a, vmin(w.b, w.c)) - r; a.hi = vmax(w.a, vmax(w.b, w.c)) + r; return a;
	}
	case B3G_SHAPE_BOX: {                                                      // box_shape.cpp:38-61
		V3 e = w.a;
		V3 vs[8] = {v3(-e.x, -e.y, -e.z), v3(-e.x, -e.y, e.z), v3(-e.x, e.y, -e.z), v3(-e.x, e.y, e.z),
		            v3(e.x, -e.y, -e.z),  v3(e.x, -e.y, e.z),  v3(e.x, e.y, -e.z),  v3(e.x, e.y, e.z)};
		AABB a;
		a.lo = a.hi = xmul(w.xf, vs[0]);
		for (int i = 1; i < 8; ++i) { V3 p = xmul(w.xf, vs[i]); a.lo = vmin(a.lo, p); a.hi = vmax(a.hi, p); }
		aabb_extend(a, w.radius);
		return a;
	}
	case B3G_SHAPE_SDF: {                                                      // sdf_shape.cpp:42-48, aabb.h:202-217
		V3 lc = 0.5f * (w.grid.lo + w.grid.hi);
		V3 lr = 0.5f * (w.grid.hi - w.grid.lo);
		V3 c = xmul(w.xf, lc);
		M33 ar = mabs(qmatrix(w.xf.q));
		V3 r = v3(dot(lr, ar.x), dot(lr, ar.y), dot(lr, ar.z));
		AABB a = aabb_cr(c, r);
		aabb_extend(a, w.radius);
		return a;
	}
	default: { AABB a; a.lo = v3(1, 1, 1); a.hi = v3(0, 0, 0); return a; }
	}
}

struct Built {
	std::vector<uint8_t> type;
	std::vector<float> damping, mass, invMass;
	std::vector<V3> x, v;
	std::vector<int32_t> sphereOf;
	std::vector<uint32_t> sphereParticle;
	std::vector<float> sphereRadius, sphereFriction;
	std::vector<std::pair<uint32_t, uint32_t>> forces;   // list order
};

b3g_status validate(const b3g_body_desc& d) {
	auto bad = [](const std::string& m) { set_error(m); return B3G_INVALID_ARGUMENT; };
	if (d.n_particles == 0) return bad("body has no particles");
	if (!d.p_type || !d.p_pos || !d.p_vel || !d.p_damping) return bad("particle arrays must not be NULL");
	for (uint32_t i = 0; i < d.n_particles; ++i)
		if (d.p_type[i] > 2) return bad("particle type must be 0, 1 or 2");
	auto chk = [&](const uint32_t* p, uint32_t m, uint32_t k, const char* what) -> bool {
		for (size_t i = 0; i < (size_t)m * k; ++i)
			if (p[i] >= d.n_particles) { set_error(std::string(what) + " references a particle out of range"); return false; }
		return true;
	};
	if (!chk(d.s_particle, d.n_spheres, 1, "sphere fixture") || !chk(d.tf_p, d.n_trifix, 3, "triangle fixture") ||
	    !chk(d.ttf_p, d.n_tetfix, 4, "tetrahedron fixture") || !chk(d.tet_p, d.n_tet, 4, "tetrahedron element") ||
	    !chk(d.tri_p, d.n_tri, 3, "triangle element") || !chk(d.st_p, d.n_stretch, 3, "stretch force") ||
	    !chk(d.sh_p, d.n_shear, 3, "shear force") || !chk(d.sp_p, d.n_spring, 2, "spring force") ||
	    !chk(d.mo_p, d.n_mouse, 4, "mouse force"))
		return B3G_INVALID_ARGUMENT;
	if (!code_limits(d)) return B3G_UNSUPPORTED;
	const uint32_t counts[7] = {0, d.n_stretch, d.n_shear, d.n_spring, d.n_mouse, d.n_tri, d.n_tet};
	std::vector<std::vector<char>> used(7);
	for (int t = 1; t < 7; ++t) used[t].assign(counts[t], 0);
	for (uint32_t c = 0; c < d.n_forces; ++c) {
		uint32_t t = d.f_type[c], k = d.f_index[c];
		if (t < 1 || t > 6) return bad("unknown force type");
		if (k >= counts[t]) return bad("force index out of range");
		if (used[t][k]) return bad("a force table row is referenced twice");
		used[t][k] = 1;
	}
	for (int t = 1; t < 7; ++t)
		for (char u : used[t])
			if (!u) return bad("every force table row must appear once in the force order");
	for (uint32_t w = 0; w < d.n_world; ++w) {
		uint32_t t = d.w_type[w];
		if (t > B3G_SHAPE_SDF) return bad("unknown world shape type");
		if (t == B3G_SHAPE_MESH) {
			const float* q = d.w_param + (size_t)B3G_WORLD_PARAMS * w;
			if (!d.w_mesh_vertices || !d.w_mesh_triangles || !d.w_mesh_counts || !d.w_mesh_vertices[w] ||
			    !d.w_mesh_triangles[w])
				return bad("mesh world fixture without vertices or triangles");
			const uint32_t nv = d.w_mesh_counts[2 * w], nt = d.w_mesh_counts[2 * w + 1];
			if (nv == 0 || nt == 0) return bad("mesh world fixture with no vertices or no triangles");
			if (q[1] == 0.0f || q[2] == 0.0f || q[3] == 0.0f) return bad("mesh scale has a zero component");   // mesh_shape.cpp:94-96
			const uint32_t* tr = d.w_mesh_triangles[w];
			for (size_t k = 0; k < 6 * (size_t)nt; ++k) {
				const bool wing = k % 6 >= 3;
				if (tr[k] >= nv && !(wing && tr[k] == 0xFFFFFFFFu)) return bad("mesh triangle vertex index out of range");
			}
		}
		if (t == B3G_SHAPE_SDF) {
			const float* q = d.w_param + (size_t)B3G_WORLD_PARAMS * w;
			if (!d.w_voxels || !d.w_voxels[w]) return bad("sdf world fixture without voxels");
			if (q[7] < 2.0f || q[8] < 2.0f || q[9] < 2.0f) return bad("sdf voxel grid needs at least 2 voxels per axis");
		}
	}
	// A contact's list position is its stamp, FindNewContacts call x world fixtures + world (32 bits): at 4096
	// world fixtures it stays exact for 2^20 steps. Every sphere pairs with every world fixture
	// (contact_manager.cpp:65-82), so the pair table holds spheres x world fixtures entries.
	if (d.n_world > 4096) { set_error("at most 4096 world fixtures per body are supported"); return B3G_UNSUPPORTED; }
	if ((uint64_t)d.n_spheres * d.n_world >= (1ull << 31)) {
		set_error("sphere fixtures x world fixtures exceeds 2^31 contact pairs");
		return B3G_UNSUPPORTED;
	}
	return B3G_OK;
}

// Several descriptions as one, in creation order: instance 0's objects first.
// Particle references are offset by the instance's first particle, force
// indices by the instance's first element of the type. List order reverses
// creation order, so each instance stays contiguous and keeps its own list
// orders.
struct Merged {
	b3g_body_desc d;
	std::vector<uint8_t> p_type;
	std::vector<float> p_pos, p_vel, p_damping;
	std::vector<uint32_t> s_particle;
	std::vector<float> s_radius, s_friction;
	std::vector<uint32_t> tf_p;
	std::vector<float> tf_v, tf_density, tf_radius, tf_friction;
	std::vector<uint32_t> ttf_p;
	std::vector<float> ttf_v, ttf_density, ttf_radius, ttf_friction;
	std::vector<uint32_t> f_type, f_index;
	std::vector<uint32_t> tet_p, tri_p, st_p, sh_p, sp_p, mo_p;
	std::vector<float> tet_v, tet_param, tri_v, tri_param, st_param, sh_param, sp_param, mo_param;
	std::vector<uint32_t> w_type;
	std::vector<float> w_param, w_friction;
	std::vector<const float*> w_voxels;
	std::vector<const float*> w_mesh_vertices;
	std::vector<const uint32_t*> w_mesh_triangles;
	std::vector<uint32_t> w_mesh_counts;
	std::vector<uint32_t> partBegin, worldBegin;   // count + 1, creation order
	std::vector<V3> gravity;
};

template <class T>
void app(std::vector<T>& v, const T* p, size_t count) {
	if (count) v.insert(v.end(), p, p + count);
}
void app_p(std::vector<uint32_t>& v, const uint32_t* p, size_t count, uint32_t off) {
	for (size_t i = 0; i < count; ++i) v.push_back(p[i] + off);
}

void merge(const b3g_body_desc* ds, uint32_t count, Merged& M) {
	uint32_t P = 0, W = 0;
	uint32_t typeOff[7] = {0, 0, 0, 0, 0, 0, 0};
	for (uint32_t k = 0; k < count; ++k) {
		const b3g_body_desc& d = ds[k];
		M.partBegin.push_back(P);
		M.worldBegin.push_back(W);
		M.gravity.push_back(v3(d.gravity[0], d.gravity[1], d.gravity[2]));
		app(M.p_type, d.p_type, d.n_particles);
		app(M.p_pos, d.p_pos, 3 * (size_t)d.n_particles);
		app(M.p_vel, d.p_vel, 3 * (size_t)d.n_particles);
		app(M.p_damping, d.p_damping, d.n_particles);
		app_p(M.s_particle, d.s_particle, d.n_spheres, P);
		app(M.s_radius, d.s_radius, d.n_spheres);
		app(M.s_friction, d.s_friction, d.n_spheres);
		app_p(M.tf_p, d.tf_p, 3 * (size_t)d.n_trifix, P);
		app(M.tf_v, d.tf_v, 9 * (size_t)d.n_trifix);
		app(M.tf_density, d.tf_density, d.n_trifix);
		app(M.tf_radius, d.tf_radius, d.n_trifix);
		app(M.tf_friction, d.tf_friction, d.n_trifix);
		app_p(M.ttf_p, d.ttf_p, 4 * (size_t)d.n_tetfix, P);
		app(M.ttf_v, d.ttf_v, 12 * (size_t)d.n_tetfix);
		app(M.ttf_density, d.ttf_density, d.n_tetfix);
		app(M.ttf_radius, d.ttf_radius, d.n_tetfix);
		app(M.ttf_friction, d.ttf_friction, d.n_tetfix);
		for (uint32_t c = 0; c < d.n_forces; ++c) {
			M.f_type.push_back(d.f_type[c]);
			M.f_index.push_back(d.f_index[c] + typeOff[d.f_type[c]]);
		}
		app_p(M.tet_p, d.tet_p, 4 * (size_t)d.n_tet, P);
		app(M.tet_v, d.tet_v, 12 * (size_t)d.n_tet);
		app(M.tet_param, d.tet_param, 3 * (size_t)d.n_tet);
		app_p(M.tri_p, d.tri_p, 3 * (size_t)d.n_tri, P);
		app(M.tri_v, d.tri_v, 9 * (size_t)d.n_tri);
		app(M.tri_param, d.tri_param, 6 * (size_t)d.n_tri);
		app_p(M.st_p, d.st_p, 3 * (size_t)d.n_stretch, P);
		app(M.st_param, d.st_param, 12 * (size_t)d.n_stretch);
		app_p(M.sh_p, d.sh_p, 3 * (size_t)d.n_shear, P);
		app(M.sh_param, d.sh_param, 8 * (size_t)d.n_shear);
		app_p(M.sp_p, d.sp_p, 2 * (size_t)d.n_spring, P);
		app(M.sp_param, d.sp_param, 3 * (size_t)d.n_spring);
		app_p(M.mo_p, d.mo_p, 4 * (size_t)d.n_mouse, P);
		app(M.mo_param, d.mo_param, 6 * (size_t)d.n_mouse);
		app(M.w_type, d.w_type, d.n_world);
		app(M.w_param, d.w_param, (size_t)B3G_WORLD_PARAMS * d.n_world);
		app(M.w_friction, d.w_friction, d.n_world);
		for (uint32_t w = 0; w < d.n_world; ++w) {
			M.w_voxels.push_back(d.w_voxels ? d.w_voxels[w] : nullptr);
			const bool mesh = d.w_type[w] == B3G_SHAPE_MESH;
			M.w_mesh_vertices.push_back(mesh ? d.w_mesh_vertices[w] : nullptr);
			M.w_mesh_triangles.push_back(mesh ? d.w_mesh_triangles[w] : nullptr);
			M.w_mesh_counts.push_back(mesh ? d.w_mesh_counts[2 * w] : 0u);
			M.w_mesh_counts.push_back(mesh ? d.w_mesh_counts[2 * w + 1] : 0u);
		}
		typeOff[B3G_FORCE_STRETCH] += d.n_stretch; typeOff[B3G_FORCE_SHEAR] += d.n_shear;
		typeOff[B3G_FORCE_SPRING] += d.n_spring; typeOff[B3G_FORCE_MOUSE] += d.n_mouse;
		typeOff[B3G_FORCE_TRI_ELEMENT] += d.n_tri; typeOff[B3G_FORCE_TET_ELEMENT] += d.n_tet;
		P += d.n_particles;
		W += d.n_world;
	}
	M.partBegin.push_back(P);
	M.worldBegin.push_back(W);
	b3g_body_desc& o = M.d;
	std::memset(&o, 0, sizeof(o));
	o.n_particles = P;
	o.p_type = M.p_type.data(); o.p_pos = M.p_pos.data(); o.p_vel = M.p_vel.data(); o.p_damping = M.p_damping.data();
	o.n_spheres = (uint32_t)M.s_radius.size();
	o.s_particle = M.s_particle.data(); o.s_radius = M.s_radius.data(); o.s_friction = M.s_friction.data();
	o.n_trifix = (uint32_t)M.tf_density.size();
	o.tf_p = M.tf_p.data(); o.tf_v = M.tf_v.data(); o.tf_density = M.tf_density.data(); o.tf_radius = M.tf_radius.data();
	o.tf_friction = M.tf_friction.data();
	o.n_tetfix = (uint32_t)M.ttf_density.size();
	o.ttf_p = M.ttf_p.data(); o.ttf_v = M.ttf_v.data(); o.ttf_density = M.ttf_density.data();
	o.ttf_radius = M.ttf_radius.data(); o.ttf_friction = M.ttf_friction.data();
	o.n_forces = (uint32_t)M.f_type.size(); o.f_type = M.f_type.data(); o.f_index = M.f_index.data();
	o.n_tet = (uint32_t)(M.tet_param.size() / 3); o.tet_p = M.tet_p.data(); o.tet_v = M.tet_v.data(); o.tet_param = M.tet_param.data();
	o.n_tri = (uint32_t)(M.tri_param.size() / 6); o.tri_p = M.tri_p.data(); o.tri_v = M.tri_v.data(); o.tri_param = M.tri_param.data();
	o.n_stretch = (uint32_t)(M.st_param.size() / 12); o.st_p = M.st_p.data(); o.st_param = M.st_param.data();
	o.n_shear = (uint32_t)(M.sh_param.size() / 8); o.sh_p = M.sh_p.data(); o.sh_param = M.sh_param.data();
	o.n_spring = (uint32_t)(M.sp_param.size() / 3); o.sp_p = M.sp_p.data(); o.sp_param = M.sp_param.data();
	o.n_mouse = (uint32_t)(M.mo_param.size() / 6); o.mo_p = M.mo_p.data(); o.mo_param = M.mo_param.data();
	o.n_world = (uint32_t)M.w_type.size(); o.w_type = M.w_type.data(); o.w_param = M.w_param.data();
	o.w_friction = M.w_friction.data(); o.w_voxels = M.w_voxels.data();
	o.w_mesh_vertices = M.w_mesh_vertices.data(); o.w_mesh_triangles = M.w_mesh_triangles.data();
	o.w_mesh_counts = M.w_mesh_counts.data();
	if (count) for (int c = 0; c < 3; ++c) o.gravity[c] = ds[0].gravity[c];
}

} // namespace

b3g_status host_pattern(const b3g_body_desc& d, std::vector<uint32_t>& rowptr, std::vector<uint32_t>& cols);

b3g_status build_body(Body& b, const b3g_body_desc* descs, uint32_t count) {
	if (count == 0) { set_error("a batch needs at least one body description"); return B3G_INVALID_ARGUMENT; }
	for (uint32_t k = 0; k < count; ++k) {
		b3g_status st = validate(descs[k]);
		if (st != B3G_OK) {
			if (count > 1) set_error("instance " + std::to_string(k) + ": " + b3g_last_error());
			return st;
		}
	}
	if (!batch_limits(descs, count)) return B3G_UNSUPPORTED;
	Merged M;
	merge(descs, count, M);
	const b3g_body_desc& d = M.d;
	if (!code_limits(d)) return B3G_UNSUPPORTED;
	b3g_status st = B3G_OK;
	drop_graph(b);
	dfree_all(b);
	b.built = false;
	b.rtVersion = ~0ull;   // the ray-cast tree is rebuilt at the next query
	b.rtSortBytes = 0;
	b.rtRayCap = 0;
	const uint32_t n = d.n_particles;
	auto L = [n](uint32_t c) { return n - 1 - c; };   // creation -> list order (body.cpp:67-75)
	b.n = n;
	// instances: creation range [partBegin[k], partBegin[k+1]) -> rows [n - partBegin[k+1], n - partBegin[k])
	b.nInst = count;
	b.hostRowInst.assign(n, 0);
	b.hostRowBegin.resize(count);
	b.hostRowEnd.resize(count);
	b.hostGravity = M.gravity;
	for (uint32_t k = 0; k < count; ++k) {
		b.hostRowBegin[k] = n - M.partBegin[k + 1];
		b.hostRowEnd[k] = n - M.partBegin[k];
		for (uint32_t i = b.hostRowBegin[k]; i < b.hostRowEnd[k]; ++i) b.hostRowInst[i] = k;
	}

	// ------------------------------------------------------------ particles
	Built h;
	h.type.resize(n); h.damping.resize(n); h.mass.resize(n); h.invMass.resize(n); h.x.resize(n); h.v.resize(n);
	for (uint32_t c = 0; c < n; ++c) {
		uint32_t i = L(c);
		h.type[i] = d.p_type[c];
		h.x[i] = load3(d.p_pos + 3 * c);
		h.v[i] = load3(d.p_vel + 3 * c);
		h.damping[i] = d.p_damping[c];
		h.mass[i] = h.type[i] == kDynamic ? 1.0f : 0.0f;   // particle.cpp:40-49
	}
	// sphere fixtures: the first fixture of a particle wins (body.cpp:151-176)
	h.sphereOf.assign(n, -1);
	{
		std::vector<uint32_t> created, createdIndex;
		std::vector<float> rad, fr;
		std::vector<char> has(n, 0);
		for (uint32_t c = 0; c < d.n_spheres; ++c) {
			uint32_t p = L(d.s_particle[c]);
			if (has[p]) continue;
			has[p] = 1;
			created.push_back(p); createdIndex.push_back(c); rad.push_back(d.s_radius[c]); fr.push_back(d.s_friction[c]);
		}
		const uint32_t ns = (uint32_t)created.size();
		b.hostSphereCreation.clear();
		for (uint32_t k = 0; k < ns; ++k) {   // list order = reverse
			b.hostSphereCreation.push_back(createdIndex[ns - 1 - k]);
			h.sphereParticle.push_back(created[ns - 1 - k]);
			h.sphereRadius.push_back(rad[ns - 1 - k]);
			h.sphereFriction.push_back(fr[ns - 1 - k]);
			h.sphereOf[created[ns - 1 - k]] = (int32_t)k;
		}
	}
	// mass lumping: body.cpp:394-466 over the deduplicated fixtures in list order
	{
		struct TriF { uint32_t p[3]; float area, density; double area64; };
		struct TetF { uint32_t p[4]; float volume, density; double volume64; };
		std::vector<TriF> tris;
		std::vector<TetF> tets;
		std::vector<std::array<uint32_t, 4>> seen;
		{
			std::vector<std::array<uint32_t, 3>> keys;
			std::vector<char> dup(d.n_trifix, 0);
			std::vector<std::pair<std::array<uint32_t, 3>, uint32_t>> srt;
			for (uint32_t c = 0; c < d.n_trifix; ++c) {
				std::array<uint32_t, 3> k = {d.tf_p[3 * c], d.tf_p[3 * c + 1], d.tf_p[3 * c + 2]};
				std::sort(k.begin(), k.end());
				srt.push_back({k, c});
			}
			std::stable_sort(srt.begin(), srt.end(), [](auto& a, auto& b2) { return a.first < b2.first; });
			for (size_t k = 1; k < srt.size(); ++k)
				if (srt[k].first == srt[k - 1].first) dup[srt[k].second] = 1;
			// ray-cast table: (instance, creation index, particles), ordered by instance then creation index
			std::vector<std::array<uint32_t, 5>> ray;
			for (uint32_t c = d.n_trifix; c-- > 0;) {
				if (dup[c]) continue;
				TriF t;
				for (int k = 0; k < 3; ++k) t.p[k] = L(d.tf_p[3 * c + k]);
				ray.push_back({b.hostRowInst[t.p[0]], c, t.p[0], t.p[1], t.p[2]});
				const float* v = d.tf_v + 9 * (size_t)c;
				V3 A = load3(v), B = load3(v + 3), C = load3(v + 6);
				t.area = 0.5f * length(cross(B - A, C - A));   // triangle_fixture.cpp:35-37
				t.area64 = 0.5 * length(cross(d3(B) - d3(A), d3(C) - d3(A)));
				t.density = d.tf_density[c];
				tris.push_back(t);
			}
			std::sort(ray.begin(), ray.end());
			b.nRayTri = (uint32_t)ray.size();
			b.hostRayTri.resize(4 * ray.size());
			b.hostRayInstBegin.assign(count + 1, 0);
			for (size_t k = 0; k < ray.size(); ++k) {
				for (int j = 0; j < 3; ++j) b.hostRayTri[4 * k + j] = ray[k][2 + j];
				b.hostRayTri[4 * k + 3] = ray[k][1];
				b.hostRayInstBegin[ray[k][0] + 1] += 1;
			}
			for (uint32_t k = 0; k < count; ++k) b.hostRayInstBegin[k + 1] += b.hostRayInstBegin[k];
		}
		{
			std::vector<char> dup(d.n_tetfix, 0);
			std::vector<std::pair<std::array<uint32_t, 4>, uint32_t>> srt;
			for (uint32_t c = 0; c < d.n_tetfix; ++c) {
				std::array<uint32_t, 4> k = {d.ttf_p[4 * c], d.ttf_p[4 * c + 1], d.ttf_p[4 * c + 2], d.ttf_p[4 * c + 3]};
				std::sort(k.begin(), k.end());
				srt.push_back({k, c});
			}
			std::stable_sort(srt.begin(), srt.end(), [](auto& a, auto& b2) { return a.first < b2.first; });
			for (size_t k = 1; k < srt.size(); ++k)
				if (srt[k].first == srt[k - 1].first) dup[srt[k].second] = 1;
			for (uint32_t c = d.n_tetfix; c-- > 0;) {
				if (dup[c]) continue;
				TetF t;
				for (int k = 0; k < 4; ++k) t.p[k] = L(d.ttf_p[4 * c + k]);
				const float* v = d.ttf_v + 12 * (size_t)c;
				V3 A = load3(v), B = load3(v + 3), C = load3(v + 6), D = load3(v + 9);
				const float inv6 = 1.0f / 6.0f;
				float dt = det(B - A, C - A, D - A);
				float sg = sign_(dt);
				t.volume = inv6 * sg * dt;                    // tetrahedron_fixture.cpp:41-45
				t.volume64 = fabs(det(d3(B) - d3(A), d3(C) - d3(A), d3(D) - d3(A))) / 6.0;
				t.density = d.ttf_density[c];
				tets.push_back(t);
			}
		}
		for (auto& t : tris) for (int k = 0; k < 3; ++k) h.mass[t.p[k]] = 0.0f;
		for (auto& t : tets) for (int k = 0; k < 4; ++k) h.mass[t.p[k]] = 0.0f;
		const float inv3 = 1.0f / 3.0f, inv4 = 1.0f / 4.0f;
		for (auto& t : tris) { float m = t.density * t.area; for (int k = 0; k < 3; ++k) h.mass[t.p[k]] += inv3 * m; }
		for (auto& t : tets) { float m = t.density * t.volume; for (int k = 0; k < 4; ++k) h.mass[t.p[k]] += inv4 * m; }
		// the same lumping in double (the precision form's right-hand side)
		std::vector<double> m64(n);
		for (uint32_t i = 0; i < n; ++i) m64[i] = h.type[i] == kDynamic ? 1.0 : 0.0;
		for (auto& t : tris) for (int k = 0; k < 3; ++k) m64[t.p[k]] = 0.0;
		for (auto& t : tets) for (int k = 0; k < 4; ++k) m64[t.p[k]] = 0.0;
		for (auto& t : tris) { double m = (double)t.density * t.area64; for (int k = 0; k < 3; ++k) m64[t.p[k]] += m / 3.0; }
		for (auto& t : tets) { double m = (double)t.density * t.volume64; for (int k = 0; k < 4; ++k) m64[t.p[k]] += m / 4.0; }
		for (uint32_t i = 0; i < n; ++i) {
			if (h.type[i] != kDynamic) { h.mass[i] = 0.0f; h.invMass[i] = 0.0f; m64[i] = 0.0; continue; }
			if (h.mass[i] > 0.0f) h.invMass[i] = 1.0f / h.mass[i];
			else { h.mass[i] = 1.0f; h.invMass[i] = 1.0f; }
			if (!(m64[i] > 0.0)) m64[i] = 1.0;
		}
		b.hostMass64 = std::move(m64);
	}
	b.hostMass = h.mass;
	b.hostType = h.type;
	b.nSpheres = (uint32_t)h.sphereParticle.size();
	b.hostSphereParticle = h.sphereParticle;

	// ------------------------------------------------------------ forces (list order = reverse creation)
	for (uint32_t c = d.n_forces; c-- > 0;) h.forces.push_back({d.f_type[c], d.f_index[c]});

	// element tables (creation order within each type), particle indices in list order
	std::vector<uint32_t> tetP(4 * (size_t)d.n_tet), triP(3 * (size_t)d.n_tri), stP(3 * (size_t)d.n_stretch),
	    shP(3 * (size_t)d.n_shear), spP(2 * (size_t)d.n_spring), moP(4 * (size_t)d.n_mouse);
	for (size_t i = 0; i < tetP.size(); ++i) tetP[i] = L(d.tet_p[i]);
	for (size_t i = 0; i < triP.size(); ++i) triP[i] = L(d.tri_p[i]);
	for (size_t i = 0; i < stP.size(); ++i) stP[i] = L(d.st_p[i]);
	for (size_t i = 0; i < shP.size(); ++i) shP[i] = L(d.sh_p[i]);
	for (size_t i = 0; i < spP.size(); ++i) spP[i] = L(d.sp_p[i]);
	for (size_t i = 0; i < moP.size(); ++i) moP[i] = L(d.mo_p[i]);

	std::vector<float> tetKd(d.n_tet), triKd(d.n_tri);
	for (uint32_t e = 0; e < d.n_tet; ++e) tetKd[e] = d.tet_param[3 * e + 2];
	for (uint32_t e = 0; e < d.n_tri; ++e) triKd[e] = d.tri_param[6 * e + 5];
	// stretch coefficients: 12 from uv_reset + ks_u kd_u b_u ks_v kd_v b_v + 2 pad = 20
	std::vector<float> stCoef(20 * (size_t)d.n_stretch), shCoef(16 * (size_t)d.n_shear);
	for (uint32_t e = 0; e < d.n_stretch; ++e) {
		uv_reset(d.st_param + 12 * (size_t)e, &stCoef[20 * (size_t)e]);
		for (int k = 0; k < 6; ++k) stCoef[20 * (size_t)e + 12 + k] = d.st_param[12 * (size_t)e + 6 + k];
	}
	for (uint32_t e = 0; e < d.n_shear; ++e) {
		uv_reset(d.sh_param + 8 * (size_t)e, &shCoef[16 * (size_t)e]);
		shCoef[16 * (size_t)e + 12] = d.sh_param[8 * (size_t)e + 6];
		shCoef[16 * (size_t)e + 13] = d.sh_param[8 * (size_t)e + 7];
	}
	std::vector<float> spCoef(d.sp_param, d.sp_param + 3 * (size_t)d.n_spring);
	std::vector<float> moCoef(d.mo_param, d.mo_param + 6 * (size_t)d.n_mouse);

	// staging layout: [type][local][element]
	uint32_t J = 0, F = 0;
	auto take = [](uint32_t& cur, uint32_t count) { uint32_t base = cur; cur += count; return base; };
	b.tet.m = d.n_tet;
	b.tet.anyDamping = std::any_of(tetKd.begin(), tetKd.end(), [](float k) { return k > 0.0f; });
	b.tet.f = take(F, 4 * d.n_tet);
	b.tet.fd = b.tet.anyDamping ? take(F, 4 * d.n_tet) : 0;
	b.tri.m = d.n_tri;
	b.tri.anyDamping = std::any_of(triKd.begin(), triKd.end(), [](float k) { return k > 0.0f; });
	b.tri.jx = take(J, 9 * d.n_tri);
	b.tri.jv = b.tri.anyDamping ? take(J, 9 * d.n_tri) : 0;
	b.tri.f = take(F, 3 * d.n_tri);
	b.tri.fd = b.tri.anyDamping ? take(F, 3 * d.n_tri) : 0;
	b.stretch.m = d.n_stretch; b.stretch.ncoef = 20;
	b.stretch.f = take(F, 12 * d.n_stretch);
	b.shear.m = d.n_shear; b.shear.ncoef = 16;
	b.shear.f = take(F, 6 * d.n_shear);
	b.spring.m = d.n_spring; b.spring.np = 2; b.spring.ncoef = 3;
	b.spring.jx = take(J, 4 * d.n_spring); b.spring.jv = take(J, 4 * d.n_spring); b.spring.f = take(F, 4 * d.n_spring);
	b.mouse.m = d.n_mouse; b.mouse.np = 4; b.mouse.ncoef = 6;
	b.mouse.jx = take(J, 16 * d.n_mouse); b.mouse.jv = take(J, 16 * d.n_mouse); b.mouse.f = take(F, 8 * d.n_mouse);
	b.nJStage = J;
	b.nFStage = F;

	// ------------------------------------------------------------ world fixtures (list order)
	b.nWorld = d.n_world;
	b.hostWorlds.clear();
	std::vector<float> vox;
	// A grid passed by several fixtures (e.g. every instance of a batch over
	// one terrain) is uploaded once: fixtures with the same voxel pointer and
	// dimensions share one offset.
	std::vector<std::pair<const float*, std::pair<size_t, uint64_t>>> voxSeen;
	// meshes: one copy (vertices, triangles, tree) per distinct (vertex array, triangle array)
	std::vector<float> meshV;
	std::vector<uint32_t> meshT;
	MeshTrees trees;
	struct MeshSeen { const float* v; const uint32_t* t; uint32_t nv, nt, vOff, tOff, node; };
	std::vector<MeshSeen> meshSeen;
	for (uint32_t c = d.n_world; c-- > 0;) {
		const float* q = d.w_param + (size_t)B3G_WORLD_PARAMS * c;
		WorldShape w;
		std::memset(&w, 0, sizeof(w));
		w.type = d.w_type[c];
		w.radius = q[0];
		w.friction = d.w_friction[c];
		w.xf.q = qidentity();
		switch (w.type) {
		case B3G_SHAPE_SPHERE: w.a = load3(q + 1); break;
		case B3G_SHAPE_CAPSULE: w.a = load3(q + 1); w.b = load3(q + 4); break;
		case B3G_SHAPE_TRIANGLE:
			w.a = load3(q + 1); w.b = load3(q + 4); w.c = load3(q + 7);
			for (int k = 0; k < 3; ++k) { w.hasE[k] = q[10 + k] != 0.0f; w.e[k] = load3(q + 13 + 3 * k); }
			break;
		case B3G_SHAPE_BOX:
			w.a = load3(q + 1); w.xf.p = load3(q + 4); w.xf.q.v = load3(q + 7); w.xf.q.s = q[10];
			break;
		case B3G_SHAPE_SDF: {
			w.grid.lo = load3(q + 1); w.grid.hi = load3(q + 4);
			w.W = (uint32_t)q[7]; w.H = (uint32_t)q[8]; w.D = (uint32_t)q[9];
			w.xf.p = load3(q + 10); w.xf.q.v = load3(q + 13); w.xf.q.s = q[16];
			w.invert = q[17] != 0.0f;
			size_t nv = (size_t)w.W * w.H * w.D;
			auto seen = std::find_if(voxSeen.begin(), voxSeen.end(), [&](const auto& e) {
				return e.first == d.w_voxels[c] && e.second.first == nv;
			});
			if (seen != voxSeen.end()) {
				w.voxOffset = seen->second.second;
			} else {
				w.voxOffset = vox.size();
				voxSeen.push_back({d.w_voxels[c], {nv, w.voxOffset}});
				vox.insert(vox.end(), d.w_voxels[c], d.w_voxels[c] + nv);
			}
			break;
		}
		case B3G_SHAPE_MESH: {
			w.a = load3(q + 1);   // scale
			w.xf.p = load3(q + 4); w.xf.q.v = load3(q + 7); w.xf.q.s = q[10];
			const float* mv = d.w_mesh_vertices[c];
			const uint32_t* mt = d.w_mesh_triangles[c];
			const uint32_t nv = d.w_mesh_counts[2 * c], nt = d.w_mesh_counts[2 * c + 1];
			auto seen = std::find_if(meshSeen.begin(), meshSeen.end(), [&](const MeshSeen& e) {
				return e.v == mv && e.t == mt && e.nv == nv && e.nt == nt;
			});
			if (seen == meshSeen.end()) {
				MeshSeen e{mv, mt, nv, nt, (uint32_t)(meshV.size() / 3), (uint32_t)(meshT.size() / 6), (uint32_t)trees.leaf.size()};
				meshV.insert(meshV.end(), mv, mv + 3 * (size_t)nv);
				meshT.insert(meshT.end(), mt, mt + 6 * (size_t)nt);
				build_mesh_tree(mv, mt, nt, trees);
				meshSeen.push_back(e);
				seen = meshSeen.end() - 1;
			}
			w.meshVert = seen->vOff;
			w.meshTri = seen->tOff;
			w.meshNode = seen->node;
			w.meshCount = nt;
			// b3MeshShape::ComputeAABB (mesh_shape.cpp:44-50, aabb.h:90-99)
			auto world = [&](uint32_t i) {
				const V3 v = load3(mv + 3 * (size_t)i);
				return xmul(w.xf, v3(w.a.x * v.x, w.a.y * v.y, w.a.z * v.z));
			};
			w.aabb.lo = w.aabb.hi = world(0);
			for (uint32_t i = 1; i < nv; ++i) {
				const V3 p = world(i);
				w.aabb.lo = vmin(w.aabb.lo, p);
				w.aabb.hi = vmax(w.aabb.hi, p);
			}
			aabb_extend(w.aabb, w.radius);
			break;
		}
		default: break;
		}
		if (w.type != B3G_SHAPE_MESH) w.aabb = shape_aabb(w);
		b.hostWorlds.push_back(w);
	}

	// ------------------------------------------------------------ symbolic pattern
	SymbolicInput si;
	si.n = n; si.type = &h.type; si.damping = &h.damping; si.sphereOf = &h.sphereOf;
	// contact pairs: each sphere (list order) with the world fixtures of its
	// instance (list order); rows that can touch a world fixture
	std::vector<char> contactRow(n, 0);
	{
		b.hostPairStart.assign(1, 0);
		b.hostPairSphere.clear();
		b.hostPairWorld.clear();
		for (uint32_t sp = 0; sp < b.nSpheres; ++sp) {
			const uint32_t row = h.sphereParticle[sp], k = b.hostRowInst[row];
			const uint32_t w0 = d.n_world - M.worldBegin[k + 1], w1 = d.n_world - M.worldBegin[k];
			for (uint32_t w = w0; w < w1; ++w) {
				b.hostPairSphere.push_back(sp);
				b.hostPairWorld.push_back(w);
			}
			if (w1 > w0) contactRow[row] = 1;
			if (b.hostPairSphere.size() >= (1ull << 31)) {
				set_error("sphere fixtures x world fixtures exceeds 2^31 contact pairs");
				return B3G_UNSUPPORTED;
			}
			b.hostPairStart.push_back((uint32_t)b.hostPairSphere.size());
		}
		b.nPairs = (uint32_t)b.hostPairSphere.size();
	}
	si.contactRow = b.nPairs ? &contactRow : nullptr;
	si.forces = h.forces;
	si.tetP = &tetP; si.tetM = d.n_tet; si.tetKd = &tetKd;
	si.triP = &triP; si.triM = d.n_tri; si.triKd = &triKd;
	si.stP = &stP; si.stM = d.n_stretch; si.stCoef = &stCoef;
	si.shP = &shP; si.shM = d.n_shear; si.shCoef = &shCoef;
	si.spP = &spP; si.spM = d.n_spring; si.spCoef = &spCoef;
	si.moP = &moP; si.moM = d.n_mouse; si.moCoef = &moCoef;
	si.tetF = b.tet.f; si.tetFd = b.tet.fd;
	si.triJx = b.tri.jx; si.triJv = b.tri.jv; si.triF = b.tri.f; si.triFd = b.tri.fd;
	si.stF = b.stretch.f;
	si.shF = b.shear.f;
	si.spJx = b.spring.jx; si.spJv = b.spring.jv; si.spF = b.spring.f;
	si.moJx = b.mouse.jx; si.moJv = b.mouse.jv; si.moF = b.mouse.f;
	Pattern P;
	if (!build_pattern(si, P)) return B3G_UNSUPPORTED;
	b.nnz = P.rowptr[n];
	b.aStride = plane_stride(b.nnz);
	b.hostRowptr = P.rowptr;
	b.hostCols = P.cols;
	// tetrahedron entries: the pattern numbers block (i, j) of element e as
	// (i + 4 j) m + e; the assembly decodes kTetCode | (i + 4 j) << 27 | e
	auto tet_codes = [&](std::vector<uint32_t>& idx) {
		if (!d.n_tet) return;
		for (size_t k = 0; k < idx.size(); ++k)
			if (idx[k] & kTetCode) {
				const uint32_t src = idx[k] & ~kTetCode;
				idx[k] = kTetCode | (src / d.n_tet) << kTetBlockShift | (src % d.n_tet);
			}
	};
	tet_codes(P.jxIdx);
	if (b.tet.anyDamping) tet_codes(P.jvIdx);
	b.anyJv = !P.jvIdx.empty();
	// slot-assembly tiles: <= 256 slots, <= 1024 contributions of each list
	// (one LDS chunk of the slot assembly)
	{
		const uint32_t kSlots = 256, kCap = 1024;
		std::vector<uint32_t> t{0};
		uint32_t slots = 0, cx = 0, cv = 0;
		std::vector<uint32_t> ti;
		uint32_t inst = b.hostRowInst[P.slotRow.empty() ? 0 : P.slotRow[0]];
		for (uint32_t s = 0; s < b.nnz; ++s) {
			const uint32_t lx = P.jxPtr[s + 1] - P.jxPtr[s], lv = P.jvPtr[s + 1] - P.jvPtr[s];
			const uint32_t k = b.hostRowInst[P.slotRow[s]];
			if (slots > 0 && (slots == kSlots || cx + lx > kCap || cv + lv > kCap || k != inst)) {
				t.push_back(s);
				ti.push_back(inst);
				slots = cx = cv = 0;
			}
			inst = k;
			slots += 1;
			cx += lx;
			cv += lv;
		}
		t.push_back(b.nnz);
		ti.push_back(inst);
		b.nSlotTiles = (uint32_t)t.size() - 1;
		st = dupload(b, b.slotTiles, t.data(), t.size() * 4);
		if (st != B3G_OK) return st;
		st = dupload(b, b.slotTileInst, ti.data(), ti.size() * 4);
		if (st != B3G_OK) return st;
	}

	// ------------------------------------------------------------ device uploads
#define UP(buf, vec) do { b3g_status s_ = dupload(b, buf, (vec).data(), (vec).size() * sizeof((vec)[0])); if (s_ != B3G_OK) return s_; } while (0)
#define AL(buf, bytes) do { b3g_status s_ = dalloc(b, buf, (bytes)); if (s_ != B3G_OK) return s_; } while (0)
	UP(b.type, h.type);
	UP(b.mass, h.mass);
	UP(b.invMass, h.invMass);
	UP(b.damping, h.damping);
	UP(b.x, h.x);
	UP(b.v, h.v);
	std::vector<V3> zeros(n, zero3());
	UP(b.force, zeros);
	UP(b.translation, zeros);
	for (DBuf* buf : {&b.x0, &b.v0, &b.fe, &b.y, &b.w, &b.pb, &b.invM, &b.px, &b.r, &b.s, &b.q, &b.d0, &b.d1, &b.fi})
		AL(*buf, (size_t)n * sizeof(V3));
	UP(b.sphereOf, h.sphereOf);
	if (b.nSpheres) {
		UP(b.sphereParticle, h.sphereParticle);
		UP(b.sphereRadius, h.sphereRadius);
		UP(b.sphereFriction, h.sphereFriction);
	}
	// elements
	if (d.n_tet) {
		UP(b.tet.p, planes(tetP, d.n_tet, 4));
		std::vector<float> rest(d.tet_v, d.tet_v + 12 * (size_t)d.n_tet);
		UP(b.tet.rest, planes(rest, d.n_tet, 12));
		std::vector<float> prm(d.tet_param, d.tet_param + 3 * (size_t)d.n_tet);
		UP(b.tet.param, planes(prm, d.n_tet, 3));
		UP(b.tet.kd, tetKd);
		AL(b.tet.rec, (size_t)d.n_tet * kTetRecFloats * 4);
		std::vector<float> q(4 * (size_t)d.n_tet, 0.0f);
		for (uint32_t e = 0; e < d.n_tet; ++e) q[4 * (size_t)e + 3] = 1.0f;   // identity (x y z s per element)
		UP(b.tet.q, q);
	}
	if (d.n_tri) {
		std::vector<float> x2(6 * (size_t)d.n_tri), invS(4 * (size_t)d.n_tri), K0(36 * (size_t)d.n_tri);
		for (uint32_t e = 0; e < d.n_tri; ++e) {
			const float* v = d.tri_v + 9 * (size_t)e;
			V2 xs[3];
			M22 iS;
			if (!tri_reset(load3(v), load3(v + 3), load3(v + 6), d.tri_param + 6 * (size_t)e, xs, iS, &K0[36 * (size_t)e])) {
				set_error("degenerate triangle element " + std::to_string(e));
				return B3G_DEGENERATE_ELEMENT;
			}
			for (int k = 0; k < 3; ++k) { x2[6 * (size_t)e + 2 * k] = xs[k].x; x2[6 * (size_t)e + 2 * k + 1] = xs[k].y; }
			invS[4 * (size_t)e] = iS.x.x; invS[4 * (size_t)e + 1] = iS.x.y; invS[4 * (size_t)e + 2] = iS.y.x; invS[4 * (size_t)e + 3] = iS.y.y;
		}
		UP(b.tri.p, planes(triP, d.n_tri, 3));
		UP(b.tri.x2, planes(x2, d.n_tri, 6));
		UP(b.tri.invS, planes(invS, d.n_tri, 4));
		UP(b.tri.K0, planes(K0, d.n_tri, 36));
		UP(b.tri.kd, triKd);
	}
	if (d.n_stretch) {
		UP(b.stretch.p, planes(stP, d.n_stretch, 3));
		UP(b.stretch.c, planes(stCoef, d.n_stretch, 20));
		// records per part: (n_w, |w|) | (dw/dx, alpha) | (ks, kd, b, 0)
		std::vector<float4> rec(6 * (size_t)d.n_stretch, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
		for (size_t e = 0; e < d.n_stretch; ++e) {
			const float* c = &stCoef[20 * e];
			for (int part = 0; part < 2; ++part) {
				const float* dw = c + 6 + 3 * part;
				rec[6 * e + 3 * part + 1] = make_float4(dw[0], dw[1], dw[2], c[0]);
				rec[6 * e + 3 * part + 2] = make_float4(c[12 + 3 * part], c[13 + 3 * part], c[14 + 3 * part], 0.0f);
			}
		}
		UP(b.stretch.rec, rec);
	}
	if (d.n_shear) {
		UP(b.shear.p, planes(shP, d.n_shear, 3));
		UP(b.shear.c, planes(shCoef, d.n_shear, 16));
		// records: (w_u, alpha) | (w_v, ks) | (dw_u/dx, kd) | (dw_v/dx, 0)
		std::vector<float4> rec(4 * (size_t)d.n_shear, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
		for (size_t e = 0; e < d.n_shear; ++e) {
			const float* c = &shCoef[16 * e];
			rec[4 * e + 2] = make_float4(c[6], c[7], c[8], c[13]);
			rec[4 * e + 3] = make_float4(c[9], c[10], c[11], 0.0f);
		}
		UP(b.shear.rec, rec);
	}
	if (d.n_spring) { UP(b.spring.p, planes(spP, d.n_spring, 2)); UP(b.spring.c, planes(spCoef, d.n_spring, 3)); }
	if (d.n_mouse) { UP(b.mouse.p, planes(moP, d.n_mouse, 4)); UP(b.mouse.c, planes(moCoef, d.n_mouse, 6)); }
	AL(b.jstage, (size_t)std::max(1u, b.nJStage) * sizeof(M33));
	AL(b.fstage, (size_t)std::max(1u, b.nFStage) * sizeof(V3));
	// pattern
	UP(b.rowptr, P.rowptr);
	{
		std::vector<uint32_t> colsPad(P.cols);   // padded to the plane stride for 16-byte loads
		colsPad.resize(b.aStride, 0u);
		UP(b.cols, colsPad);
	}
	UP(b.slotRow, P.slotRow);
	UP(b.diagSlot, P.diagSlot);
	UP(b.jxPtr, P.jxPtr);
	if (!P.jxIdx.empty()) UP(b.jxIdx, P.jxIdx);
	UP(b.jvPtr, P.jvPtr);
	if (!P.jvIdx.empty()) UP(b.jvIdx, P.jvIdx);
	UP(b.bPtr, P.bPtr);
	if (!P.bSlot.empty()) UP(b.bSlot, P.bSlot);
	UP(b.fPtr, P.fPtr);
	if (!P.fIdx.empty()) UP(b.fIdx, P.fIdx);
	AL(b.A, (size_t)b.aStride * 9 * 4);
	AL(b.tB, (size_t)b.nnz * sizeof(V3));
	{
		SolverState s0;
		std::memset(&s0, 0, sizeof(s0));
		s0.statMinSub = 0xFFFFFFFFu;
		std::vector<SolverState> states(count, s0);
		UP(b.state, states);
		const std::vector<uint32_t> noFault(1, 0u);
		UP(b.faultWord, noFault);
	}
	UP(b.gravity, b.hostGravity);
	UP(b.rowInst, b.hostRowInst);
	// contacts
	if (b.nWorld) {
		UP(b.worlds, b.hostWorlds);
		if (!vox.empty()) UP(b.voxels, vox);
		if (!meshT.empty()) {
			UP(b.meshV, meshV);
			UP(b.meshT, meshT);
			UP(b.meshLeaf, trees.leaf);
			UP(b.meshLeafBox, trees.leafBox);
			UP(b.meshChild, trees.child);
			UP(b.meshNodeBox, trees.nodeBox);
		}
	}
	if (b.nPairs) {
		const size_t pairs = b.nPairs;
		UP(b.pairStart, b.hostPairStart);
		UP(b.pairSphere, b.hostPairSphere);
		UP(b.pairWorld, b.hostPairWorld);
		AL(b.cActive, pairs * 4); AL(b.cStamp, pairs * 4); AL(b.cFriction, pairs * 4); AL(b.cHit, pairs * 4);
		AL(b.cNormalForce, pairs * 4); AL(b.cManifold, pairs * 2 * sizeof(V3)); AL(b.cF, pairs * sizeof(V3));
		AL(b.cK, pairs * sizeof(M33));
		std::vector<uint32_t> zeros32(pairs, 0u);
		UP(b.cActive, zeros32);
	}
	{
		const uint32_t zero = 0;
		st = dupload(b, b.stampCounter, &zero, 4);
		if (st != B3G_OK) return st;
	}
#undef UP
#undef AL
	st = build_tiles(b);
	if (st != B3G_OK) return st;
	if (d.n_tet) {
		st = launch_tet_setup(b);
		if (st != B3G_OK) return st;
	}
	st = ensure_actions(b);
	if (st != B3G_OK) return st;
	B3G_HIP(hipStreamSynchronize(b.stream));
	b.built = true;
	return B3G_OK;
}

// Pattern only (no device): for host-side tests of the symbolic phase.
b3g_status host_pattern(const b3g_body_desc& d, std::vector<uint32_t>& rowptr, std::vector<uint32_t>& cols) {
	b3g_status st = validate(d);
	if (st != B3G_OK) return st;
	// Re-use the body build's host half through a throwaway body with no device.
	const uint32_t n = d.n_particles;
	auto L = [n](uint32_t c) { return n - 1 - c; };
	std::vector<uint8_t> type(n);
	std::vector<float> damping(n);
	std::vector<int32_t> sphereOf(n, -1);
	for (uint32_t c = 0; c < n; ++c) { type[L(c)] = d.p_type[c]; damping[L(c)] = d.p_damping[c]; }
	uint32_t nSpheres = 0;
	for (uint32_t c = 0; c < d.n_spheres; ++c)
		if (sphereOf[L(d.s_particle[c])] < 0) sphereOf[L(d.s_particle[c])] = (int32_t)nSpheres++;
	SymbolicInput si;
	si.n = n; si.type = &type; si.damping = &damping; si.sphereOf = &sphereOf;
	std::vector<char> contactRow(n, d.n_world > 0 ? 1 : 0);
	si.contactRow = d.n_world > 0 && nSpheres > 0 ? &contactRow : nullptr;
	for (uint32_t c = d.n_forces; c-- > 0;) si.forces.push_back({d.f_type[c], d.f_index[c]});
	std::vector<uint32_t> tetP(4 * (size_t)d.n_tet), triP(3 * (size_t)d.n_tri), stP(3 * (size_t)d.n_stretch),
	    shP(3 * (size_t)d.n_shear), spP(2 * (size_t)d.n_spring), moP(4 * (size_t)d.n_mouse);
	for (size_t i = 0; i < tetP.size(); ++i) tetP[i] = L(d.tet_p[i]);
	for (size_t i = 0; i < triP.size(); ++i) triP[i] = L(d.tri_p[i]);
	for (size_t i = 0; i < stP.size(); ++i) stP[i] = L(d.st_p[i]);
	for (size_t i = 0; i < shP.size(); ++i) shP[i] = L(d.sh_p[i]);
	for (size_t i = 0; i < spP.size(); ++i) spP[i] = L(d.sp_p[i]);
	for (size_t i = 0; i < moP.size(); ++i) moP[i] = L(d.mo_p[i]);
	std::vector<float> tetKd(d.n_tet), triKd(d.n_tri), stCoef(20 * (size_t)d.n_stretch), shCoef(16 * (size_t)d.n_shear);
	for (uint32_t e = 0; e < d.n_tet; ++e) tetKd[e] = d.tet_param[3 * e + 2];
	for (uint32_t e = 0; e < d.n_tri; ++e) triKd[e] = d.tri_param[6 * e + 5];
	for (uint32_t e = 0; e < d.n_stretch; ++e)
		for (int k = 0; k < 6; ++k) stCoef[20 * (size_t)e + 12 + k] = d.st_param[12 * (size_t)e + 6 + k];
	for (uint32_t e = 0; e < d.n_shear; ++e) { shCoef[16 * (size_t)e + 12] = d.sh_param[8 * (size_t)e + 6]; shCoef[16 * (size_t)e + 13] = d.sh_param[8 * (size_t)e + 7]; }
	std::vector<float> spCoef(d.sp_param, d.sp_param + 3 * (size_t)d.n_spring), moCoef(d.mo_param, d.mo_param + 6 * (size_t)d.n_mouse);
	si.tetP = &tetP; si.tetM = d.n_tet; si.tetKd = &tetKd;
	si.triP = &triP; si.triM = d.n_tri; si.triKd = &triKd;
	si.stP = &stP; si.stM = d.n_stretch; si.stCoef = &stCoef;
	si.shP = &shP; si.shM = d.n_shear; si.shCoef = &shCoef;
	si.spP = &spP; si.spM = d.n_spring; si.spCoef = &spCoef;
	si.moP = &moP; si.moM = d.n_mouse; si.moCoef = &moCoef;
	si.tetF = si.tetFd = si.triJx = si.triJv = si.triF = si.triFd = 0;
	si.stF = si.shF = si.spJx = si.spJv = si.spF = si.moJx = si.moJv = si.moF = 0;
	Pattern P;
	if (!build_pattern(si, P)) return B3G_UNSUPPORTED;
	rowptr = P.rowptr;
	cols = P.cols;
	return B3G_OK;
}

} // namespace b3g
