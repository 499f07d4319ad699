// Operator-level backward Euler: b3SparseSolveBE (sparse/sparse_force_solver.cpp:104-232)
// behind its plugin, b3SparseForceModel (sparse_force_solver.h:44-49).
//
// The force model is the caller's host code, called once per Newton
// iteration as the reference calls it. It fills f and adds blocks to two
// matrices with the reference's list semantics (b3g_matrix). After the call
// the host flattens the two matrices (rows head to tail), works out pA's row
// order once per pattern (every sparse copy in the reference reverses a row,
// sparse_mat33.h:210-230), and the device forms A, pA, b and pb, runs the PCG
// and the Newton update (b3g_kernels.hip, k_be_*).
#include "b3g_internal.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace b3g;

// A b3SparseMat33 on the host: a row is a vector whose back() is the list head.
struct b3g_matrix_s {
	struct Entry {
		uint32_t col;
		M33 val;
	};
	uint32_t n = 0;
	std::vector<std::vector<Entry>> rows;

	void reset(uint32_t rowsCount) {
		if (rows.size() != rowsCount) rows.assign(rowsCount, {});
		else
			for (auto& r : rows) r.clear();
		n = rowsCount;
	}
	// b3SparseMat33::operator()(i, j) (sparse_mat33.h:268-288): search from the
	// head, else insert a zero block at the head.
	M33& at(uint32_t i, uint32_t j) {
		std::vector<Entry>& r = rows[i];
		for (size_t k = r.size(); k-- > 0;)
			if (r[k].col == j) return r[k].val;
		r.push_back({j, mzero()});
		return r.back().val;
	}
	size_t count() const {
		size_t c = 0;
		for (const auto& r : rows) c += r.size();
		return c;
	}
};

namespace {

// Pinned host staging (hipHostMalloc), grown on demand.
struct Pinned {
	void* p = nullptr;
	size_t bytes = 0;
	b3g_status ensure(size_t need) {
		if (bytes >= need && p) return B3G_OK;
		if (p) (void)hipHostFree(p);
		p = nullptr;
		bytes = 0;
		const size_t cap = std::max<size_t>(need, 4096);
		B3G_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
		bytes = cap;
		return B3G_OK;
	}
	void release() {
		if (p) (void)hipHostFree(p);
		p = nullptr;
		bytes = 0;
	}
	template <class T> T* as() { return static_cast<T*>(p); }
};

} // namespace

struct b3g_be_s {
	Body body;
	// force-model output on the device: dfdx / dfdv in list order, f
	DBuf rowptrX, colsX, valX, valV, fi;
	// the system's symbolic structure (per pattern) and constant inputs (per solve)
	DBuf slotRow, slotX, slotV, diagSlot, Araw;
	DBuf M, S, x0, v0, fe, y, z, w;
	uint32_t rowsCap = 0;        // rows the per-row buffers hold
	size_t slotsCap = 0;         // pA slots Araw / slot arrays hold
	size_t xCap = 0, vCap = 0;   // dfdx / dfdv blocks valX / valV hold

	b3g_matrix_s dfdx, dfdv;
	// flattened force-model output
	std::vector<uint32_t> rpX, clX, rpV, clV;
	// the pattern the symbolic structure was built for
	std::vector<uint32_t> pRpX, pClX, pRpV, pClV;
	bool pUseV = false, pUseX = false, havePattern = false;
	std::vector<uint32_t> markX, posX, markV, posV;

	Pinned stage;   // force-model values and f on the way up, x / v on the way down
};

namespace {

b3g_status ensure_rows(b3g_be_s& e, uint32_t n) {
	if (e.rowsCap >= n && e.rowsCap > 0) return B3G_OK;
	Body& b = e.body;
	const size_t v3 = (size_t)n * sizeof(V3), m33 = (size_t)n * sizeof(M33);
	b3g_status s;
	for (DBuf* buf : {&b.x, &b.v, &b.pb, &b.invM, &b.px, &b.r, &b.s, &b.q, &b.d0, &b.d1, &e.x0, &e.v0, &e.fe, &e.y, &e.z,
	                  &e.w, &e.fi})
		if ((s = dalloc(b, *buf, v3)) != B3G_OK) return s;
	for (DBuf* buf : {&e.M, &e.S})
		if ((s = dalloc(b, *buf, m33)) != B3G_OK) return s;
	if ((s = dalloc(b, e.rowptrX, ((size_t)n + 1) * 4)) != B3G_OK) return s;
	if ((s = dalloc(b, e.diagSlot, (size_t)n * 4)) != B3G_OK) return s;
	if (!b.state.p && (s = dalloc(b, b.state, sizeof(SolverState))) != B3G_OK) return s;
	e.rowsCap = n;
	e.havePattern = false;   // rowptr / tiles sized for another system
	return B3G_OK;
}

// pA's slots in the reference's row order (sparse_force_solver.cpp:165-174):
// with D = dfdx's row and V = dfdv's row (head to tail, each empty when the
// scalar that scales it is 0), A's row is (D \ (V + {i})) ++ [i] ++ reverse(V \ {i})
// -- M - h dfdv inserts V's entries at the head after the diagonal, the copy
// in A - (h h) dfdx reverses that row before D's new entries go in at the
// head -- and the four copies of S A S^T + I - S leave that order.
b3g_status build_pattern(b3g_be_s& e, uint32_t n, bool useV, bool useX) {
	Body& b = e.body;
	std::vector<uint32_t> rowptr(1, 0u), cols, slotRow;
	std::vector<int32_t> sX, sV;
	std::vector<uint32_t> diag(n);
	const size_t reserve = (size_t)n + (useX ? e.clX.size() : 0) + (useV ? e.clV.size() : 0);
	cols.reserve(reserve);
	slotRow.reserve(reserve);
	sX.reserve(reserve);
	sV.reserve(reserve);
	e.markX.assign(n, 0u);
	e.posX.resize(n);
	e.markV.assign(n, 0u);
	e.posV.resize(n);
	auto emit = [&](uint32_t i, uint32_t j, int32_t kx, int32_t kv) {
		cols.push_back(j);
		slotRow.push_back(i);
		sX.push_back(kx);
		sV.push_back(kv);
	};
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t stamp = i + 1;
		if (useX)
			for (uint32_t k = e.rpX[i]; k < e.rpX[i + 1]; ++k) { e.markX[e.clX[k]] = stamp; e.posX[e.clX[k]] = k; }
		if (useV)
			for (uint32_t k = e.rpV[i]; k < e.rpV[i + 1]; ++k) { e.markV[e.clV[k]] = stamp; e.posV[e.clV[k]] = k; }
		auto inX = [&](uint32_t j) { return useX && e.markX[j] == stamp ? (int32_t)e.posX[j] : -1; };
		auto inV = [&](uint32_t j) { return useV && e.markV[j] == stamp ? (int32_t)e.posV[j] : -1; };
		if (useX)
			for (uint32_t k = e.rpX[i]; k < e.rpX[i + 1]; ++k) {
				const uint32_t j = e.clX[k];
				if (j != i && inV(j) < 0) emit(i, j, (int32_t)k, -1);
			}
		diag[i] = (uint32_t)cols.size();
		emit(i, i, inX(i), inV(i));
		if (useV)
			for (uint32_t k = e.rpV[i + 1]; k-- > e.rpV[i];) {
				const uint32_t j = e.clV[k];
				if (j != i) emit(i, j, inX(j), (int32_t)k);
			}
		if (cols.size() > 0xFFFFFF00ull) { set_error("system too large: more than 2^32 - 256 blocks"); return B3G_UNSUPPORTED; }
		rowptr.push_back((uint32_t)cols.size());
	}
	const uint32_t nnz = (uint32_t)cols.size();
	b.n = n;
	b.nnz = nnz;
	b.nInst = 1;
	b.aStride = plane_stride(nnz);
	b.hostRowptr = rowptr;
	b.hostCols = cols;
	std::vector<uint32_t> colsPad(b.aStride, 0u);
	std::copy(cols.begin(), cols.end(), colsPad.begin());
	b3g_status s;
	if ((s = dupload(b, b.rowptr, rowptr.data(), rowptr.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, b.cols, colsPad.data(), colsPad.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, e.slotRow, slotRow.data(), slotRow.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, e.slotX, sX.data(), sX.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, e.slotV, sV.data(), sV.size() * 4)) != B3G_OK) return s;
	B3G_HIP(hipMemcpy(e.diagSlot.p, diag.data(), (size_t)n * 4, hipMemcpyHostToDevice));
	if (e.slotsCap < nnz || !e.Araw.p) {
		if ((s = dalloc(b, e.Araw, std::max<size_t>(1, nnz) * sizeof(M33))) != B3G_OK) return s;
		e.slotsCap = nnz;
	}
	if ((s = dalloc(b, b.A, (size_t)9 * b.aStride * 4)) != B3G_OK) return s;
	return build_tiles(b, true, true);   // column-sorted storage, the default schedule's tiles and row blocks
}

// Flattens a matrix's rows head to tail into rowptr / cols and the block
// values into `vals` (M33 each).
void flatten(const b3g_matrix_s& m, std::vector<uint32_t>& rp, std::vector<uint32_t>& cl, M33* vals) {
	rp.assign(1, 0u);
	cl.clear();
	size_t k = 0;
	for (const auto& r : m.rows) {
		for (size_t t = r.size(); t-- > 0;) {
			cl.push_back(r[t].col);
			if (vals) vals[k] = r[t].val;
			++k;
		}
		rp.push_back((uint32_t)cl.size());
	}
}

b3g_status matrix_apply(b3g_matrix m, uint32_t count, const uint32_t* rows, const uint32_t* cols, const float* blocks,
                        bool sub) {
	if (!m) { set_error("null matrix"); return B3G_INVALID_ARGUMENT; }
	if (count == 0) return B3G_OK;
	if (!rows || !cols || !blocks) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	for (uint32_t k = 0; k < count; ++k)
		if (rows[k] >= m->n || cols[k] >= m->n) {
			set_error("block (" + std::to_string(rows[k]) + ", " + std::to_string(cols[k]) + ") outside a " +
			          std::to_string(m->n) + "-row matrix");
			return B3G_INVALID_ARGUMENT;
		}
	for (uint32_t k = 0; k < count; ++k) {
		M33 K;
		std::memcpy(&K, blocks + 9 * (size_t)k, sizeof(M33));
		M33& a = m->at(rows[k], cols[k]);
		a = sub ? a - K : a + K;   // b3Mat33::operator-= / += (mat33.h:58-72)
	}
	return B3G_OK;
}

b3g_status solve(b3g_be_s& e, b3g_be_output* out, const b3g_be_input* in) {
	Body& b = e.body;
	const uint32_t n = in->n;
	B3G_HIP(hipSetDevice(b.device));
	b3g_status s;
	if ((s = ensure_rows(e, n)) != B3G_OK) return s;
	const size_t v3 = (size_t)n * sizeof(V3), m33 = (size_t)n * sizeof(M33);
	// constant inputs of the solve; x = x0, v = v0 (sparse_force_solver.cpp:136-137)
	B3G_HIP(hipMemcpyAsync(e.x0.p, in->x0, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(e.v0.p, in->v0, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(b.x.p, in->x0, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(b.v.p, in->v0, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(e.fe.p, in->fe, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(e.y.p, in->y, v3, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(e.M.p, in->M, m33, hipMemcpyHostToDevice, b.stream));
	B3G_HIP(hipMemcpyAsync(e.S.p, in->S, m33, hipMemcpyHostToDevice, b.stream));
	if (in->z) B3G_HIP(hipMemcpyAsync(e.z.p, in->z, v3, hipMemcpyHostToDevice, b.stream));
	else B3G_HIP(hipMemsetAsync(e.z.p, 0, v3, b.stream));   // +0 in every component
	B3G_HIP(hipStreamSynchronize(b.stream));   // the caller's arrays may change after the call returns

	const float h = in->h, hh = h * h;
	const bool useV = h != 0.0f, useX = hh != 0.0f;
	std::vector<float> xh(in->x0, in->x0 + 3 * (size_t)n), vh(in->v0, in->v0 + 3 * (size_t)n), fh(3 * (size_t)n);
	SolverState st{};
	bool begun = false;
	for (;;) {
		// the force model (sparse_force_solver.cpp:141-162)
		e.dfdx.reset(n);
		e.dfdv.reset(n);
		std::fill(fh.begin(), fh.end(), 0.0f);
		b3g_force_data data;
		data.h = in->h;
		data.inv_h = in->inv_h;
		data.n = n;
		data.x = xh.data();
		data.v = vh.data();
		data.f = fh.data();
		data.dfdx = &e.dfdx;
		data.dfdv = &e.dfdv;
		const b3g_status ms = in->force_model(in->user, &data);
		if (ms != B3G_OK) {
			set_error("the force model returned status " + std::to_string((int)ms));
			return ms;
		}
		// flatten into pinned staging: dfdx values, dfdv values, f
		const size_t nx = e.dfdx.count(), nv = e.dfdv.count();
		if (nx > 0xFFFFFF00ull || nv > 0xFFFFFF00ull) { set_error("force model matrix too large"); return B3G_UNSUPPORTED; }
		// one staging size for the uploads and, after them on the same stream, the downloads
		if ((s = e.stage.ensure(std::max((nx + nv) * sizeof(M33) + v3, 2 * v3))) != B3G_OK) return s;
		M33* sx = e.stage.as<M33>();
		M33* sv = sx + nx;
		float* sf = reinterpret_cast<float*>(sv + nv);
		flatten(e.dfdx, e.rpX, e.clX, sx);
		flatten(e.dfdv, e.rpV, e.clV, sv);
		std::memcpy(sf, fh.data(), v3);
		// symbolic structure, when the pattern differs from the last one
		if (!e.havePattern || useV != e.pUseV || useX != e.pUseX || e.rpX != e.pRpX || e.clX != e.pClX ||
		    e.rpV != e.pRpV || e.clV != e.pClV || b.n != n) {
			e.havePattern = false;
			if ((s = build_pattern(e, n, useV, useX)) != B3G_OK) return s;
			if (e.xCap < nx || !e.valX.p || !e.colsX.p) {
				if ((s = dalloc(b, e.valX, std::max<size_t>(1, nx) * sizeof(M33))) != B3G_OK) return s;
				if ((s = dalloc(b, e.colsX, std::max<size_t>(1, nx) * 4)) != B3G_OK) return s;
				e.xCap = nx;
			}
			if (e.vCap < nv || !e.valV.p) {
				if ((s = dalloc(b, e.valV, std::max<size_t>(1, nv) * sizeof(M33))) != B3G_OK) return s;
				e.vCap = nv;
			}
			B3G_HIP(hipMemcpy(e.rowptrX.p, e.rpX.data(), e.rpX.size() * 4, hipMemcpyHostToDevice));
			if (nx) B3G_HIP(hipMemcpy(e.colsX.p, e.clX.data(), nx * 4, hipMemcpyHostToDevice));
			e.pRpX = e.rpX; e.pClX = e.clX; e.pRpV = e.rpV; e.pClV = e.clV;
			e.pUseV = useV; e.pUseX = useX;
			e.havePattern = true;
		}
		if (nx) B3G_HIP(hipMemcpyAsync(e.valX.p, sx, nx * sizeof(M33), hipMemcpyHostToDevice, b.stream));
		if (nv) B3G_HIP(hipMemcpyAsync(e.valV.p, sv, nv * sizeof(M33), hipMemcpyHostToDevice, b.stream));
		B3G_HIP(hipMemcpyAsync(e.fi.p, sf, v3, hipMemcpyHostToDevice, b.stream));

		BEArgs a;
		a.n = n;
		a.nnz = b.nnz;
		a.h = h;
		a.hh = hh;
		a.useV = useV ? 1u : 0u;
		a.rowptr = dptr<uint32_t>(b.rowptr);
		a.cols = dptr<uint32_t>(b.cols);
		a.slotRow = dptr<uint32_t>(e.slotRow);
		a.slotPos = b.sortedA ? dptr<uint32_t>(b.slotPos) : nullptr;
		a.diagSlot = dptr<uint32_t>(e.diagSlot);
		a.slotX = dptr<int32_t>(e.slotX);
		a.slotV = dptr<int32_t>(e.slotV);
		a.rowptrX = dptr<uint32_t>(e.rowptrX);
		a.colsX = dptr<uint32_t>(e.colsX);
		a.valX = dptr<M33>(e.valX);
		a.valV = dptr<M33>(e.valV);
		a.M = dptr<M33>(e.M);
		a.S = dptr<M33>(e.S);
		a.x0 = dptr<V3>(e.x0);
		a.v0 = dptr<V3>(e.v0);
		a.fe = dptr<V3>(e.fe);
		a.fi = dptr<V3>(e.fi);
		a.y = dptr<V3>(e.y);
		a.z = dptr<V3>(e.z);
		a.x = dptr<V3>(b.x);
		a.v = dptr<V3>(b.v);
		a.w = dptr<V3>(e.w);
		a.px = dptr<V3>(b.px);
		a.pb = dptr<V3>(b.pb);
		a.invM = dptr<V3>(b.invM);
		a.Araw = dptr<M33>(e.Araw);
		a.A = dptr<float>(b.A);
		a.stride = b.aStride;
		a.st = dptr<SolverState>(b.state);
		if (!begun) {
			if ((s = enqueue_be_begin(b, a, in->max_iterations, in->tolerance, in->max_sub_iterations, in->sub_tolerance,
			                          out->min_sub_iterations, out->max_sub_iterations)) != B3G_OK)
				return s;
			begun = true;
		}
		if ((s = enqueue_be_iteration(b, a, in->max_sub_iterations, in->sub_tolerance)) != B3G_OK) return s;
		// the iterate for the next model call (or the result), and the Newton state
		float* sxh = e.stage.as<float>();
		float* svh = sxh + 3 * (size_t)n;
		B3G_HIP(hipMemcpyAsync(sxh, b.x.p, v3, hipMemcpyDeviceToHost, b.stream));
		B3G_HIP(hipMemcpyAsync(svh, b.v.p, v3, hipMemcpyDeviceToHost, b.stream));
		B3G_HIP(hipMemcpyAsync(&st, b.state.p, sizeof(SolverState), hipMemcpyDeviceToHost, b.stream));
		B3G_HIP(hipStreamSynchronize(b.stream));
		std::memcpy(xh.data(), sxh, v3);
		std::memcpy(vh.data(), svh, v3);
		if (!st.newtonActive) break;
	}
	std::memcpy(out->x, xh.data(), v3);
	std::memcpy(out->v, vh.data(), v3);
	out->iterations = st.statNewton;
	out->error = st.error;
	out->min_sub_iterations = st.statMinSub;
	out->max_sub_iterations = st.statMaxSub;
	// the reference's release build goes on past its asserts; the results are
	// written as it would write them, and the status says what happened
	if (st.faults & B3G_FAULT_DIAGONAL) {
		set_error("a Jacobi diagonal entry of pA was not > 0 (sparse_solver.cpp:45-52)");
		return B3G_NON_POSITIVE_DIAGONAL;
	}
	if (st.faults & B3G_FAULT_NON_FINITE) {
		set_error("the solve produced a non-finite position or velocity");
		return B3G_NON_FINITE;
	}
	return B3G_OK;
}

} // namespace

extern "C" {

b3g_status b3g_matrix_add(b3g_matrix m, uint32_t count, const uint32_t* rows, const uint32_t* cols,
                          const float* blocks) {
	return matrix_apply(m, count, rows, cols, blocks, false);
}

b3g_status b3g_matrix_sub(b3g_matrix m, uint32_t count, const uint32_t* rows, const uint32_t* cols,
                          const float* blocks) {
	return matrix_apply(m, count, rows, cols, blocks, true);
}

b3g_status b3g_matrix_set_row(b3g_matrix m, uint32_t row, uint32_t count, const uint32_t* cols, const float* blocks) {
	if (!m) { set_error("null matrix"); return B3G_INVALID_ARGUMENT; }
	if (row >= m->n) { set_error("row " + std::to_string(row) + " outside a " + std::to_string(m->n) + "-row matrix"); return B3G_INVALID_ARGUMENT; }
	if (count && (!cols || !blocks)) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	std::vector<uint32_t> sorted(cols, cols + count);
	std::sort(sorted.begin(), sorted.end());
	for (uint32_t k = 0; k < count; ++k) {
		if (sorted[k] >= m->n) { set_error("column " + std::to_string(sorted[k]) + " out of range"); return B3G_INVALID_ARGUMENT; }
		if (k && sorted[k] == sorted[k - 1]) { set_error("column " + std::to_string(sorted[k]) + " repeated in a row"); return B3G_INVALID_ARGUMENT; }
	}
	auto& r = m->rows[row];
	r.clear();
	for (uint32_t k = count; k-- > 0;) {   // the head is the vector's back
		b3g_matrix_s::Entry e;
		e.col = cols[k];
		std::memcpy(&e.val, blocks + 9 * (size_t)k, sizeof(M33));
		r.push_back(e);
	}
	return B3G_OK;
}

b3g_status b3g_be_create(b3g_be* out, int device) {
	if (!out) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	*out = nullptr;
	int count = 0;
	B3G_HIP(hipGetDeviceCount(&count));
	if (device < 0 || device >= count) { set_error("no such device"); return B3G_INVALID_ARGUMENT; }
	std::unique_ptr<b3g_be_s> e(new b3g_be_s());
	B3G_HIP(hipSetDevice(device));
	e->body.device = device;
	B3G_HIP(hipStreamCreateWithFlags(&e->body.stream, hipStreamNonBlocking));
	b3g_status s = ensure_exit_flags(e->body);
	if (s != B3G_OK) {
		(void)hipStreamDestroy(e->body.stream);
		return s;
	}
	e->body.earlyExit = true;   // the PCG stops enqueuing once it has converged (sparse_solver.cpp:63-73)
	*out = e.release();
	return B3G_OK;
}

b3g_status b3g_be_destroy(b3g_be e) {
	if (!e) return B3G_OK;
	(void)hipSetDevice(e->body.device);
	(void)hipStreamSynchronize(e->body.stream);
	dfree_all(e->body);
	free_exit_flags(e->body);
	e->stage.release();
	if (e->body.stream) (void)hipStreamDestroy(e->body.stream);
	delete e;
	return B3G_OK;
}

b3g_status b3g_be_set_reduction(b3g_be e, int mode) {
	if (!e) { set_error("null solver"); return B3G_INVALID_ARGUMENT; }
	if (mode != B3G_REDUCE_FP64 && mode != B3G_REDUCE_SEQUENTIAL && mode != B3G_REDUCE_FP64_FUSED &&
	    mode != B3G_REDUCE_FP64_SEPARATE) {
		set_error("unknown reduction mode");
		return B3G_INVALID_ARGUMENT;
	}
	e->body.reduction = mode;
	return B3G_OK;
}

b3g_status b3g_be_input_defaults(b3g_be_input* in) {
	if (!in) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	std::memset(in, 0, sizeof(*in));
	in->max_iterations = 1;
	in->tolerance = kEps;
	in->max_sub_iterations = 20;
	in->sub_tolerance = kEps;
	return B3G_OK;
}

b3g_status b3g_solve_be(b3g_be e, b3g_be_output* out, const b3g_be_input* in) {
	if (!e || !out || !in) { set_error("null argument"); return B3G_INVALID_ARGUMENT; }
	const uint32_t n = in->n;
	if (n > 0 && (!in->x0 || !in->v0 || !in->fe || !in->M || !in->y || !in->S || !out->x || !out->v)) {
		set_error("null array");
		return B3G_INVALID_ARGUMENT;
	}
	if (in->max_iterations > 0 && !in->force_model) { set_error("null force model"); return B3G_INVALID_ARGUMENT; }
	if (in->max_sub_iterations > 0x3FFFFFFFu) { set_error("too many PCG iterations"); return B3G_INVALID_ARGUMENT; }
	// no Newton iteration: x = x0, v = v0, error 0 (sparse_force_solver.cpp:136-140, 226-231)
	if (in->max_iterations == 0 || n == 0) {
		if (n == 0 && in->max_iterations > 0) {
			// the reference still calls the model once on the empty system; its
			// PCG then returns at once (delta_0 = 0), and error = 0 ends the loop
			b3g_matrix_s mx, mv;
			mx.reset(0);
			mv.reset(0);
			b3g_force_data data{in->h, in->inv_h, 0u, nullptr, nullptr, nullptr, &mx, &mv};
			const b3g_status ms = in->force_model(in->user, &data);
			if (ms != B3G_OK) {
				set_error("the force model returned status " + std::to_string((int)ms));
				return ms;
			}
			const uint32_t subIt = 0;
			out->min_sub_iterations = std::min(out->min_sub_iterations, subIt);
			out->max_sub_iterations = std::max(out->max_sub_iterations, subIt);
		}
		if (n) {
			std::memcpy(out->x, in->x0, (size_t)n * sizeof(V3));
			std::memcpy(out->v, in->v0, (size_t)n * sizeof(V3));
		}
		out->iterations = 0;
		out->error = 0.0f;
		return B3G_OK;
	}
	return solve(*e, out, in);
}

} // extern "C"
