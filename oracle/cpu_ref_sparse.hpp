/*
 * TEST INFRASTRUCTURE (oracle). Block-sparse 3x3 algebra with the reference's
 * list semantics (sparse/sparse_mat33.h, dense_vec3.h, diag_mat33.h).
 *
 * A reference row is a doubly linked list; new entries are inserted at the
 * head (sparse_mat33.h:50-60) and Copy() walks the source from its head while
 * inserting at the destination head, so every copy reverses a row
 * (sparse_mat33.h:210-230). Here a row is a std::vector whose back() is the
 * list head: insertion is push_back, head-to-tail traversal walks the vector
 * backwards and Copy() produces the reversed vector. Summation orders of all
 * products and sums therefore match the reference term for term.
 */
#ifndef B3O_CPU_REF_SPARSE_HPP
#define B3O_CPU_REF_SPARSE_HPP

#include "cpu_ref_math.hpp"
#include <vector>

namespace b3o {

template <class T> using DVec = std::vector<V3<T>>;   // b3DenseVec3
template <class T> using DMat = std::vector<M33<T>>;  // b3DiagMat33

template <class T> struct SpMat {
	struct Entry {
		uint32_t col;
		M33<T> val;
	};
	std::vector<std::vector<Entry>> rows; // back() == list head

	explicit SpMat(uint32_t n) : rows(n) {}
	uint32_t n() const { return (uint32_t)rows.size(); }

	// sparse_mat33.h:268-288 (search from head, insert a zero block at head)
	M33<T>& at(uint32_t i, uint32_t j) {
		std::vector<Entry>& r = rows[i];
		for (size_t k = r.size(); k-- > 0;)
			if (r[k].col == j) return r[k].val;
		r.push_back({j, M33<T>::Zero()});
		return r.back().val;
	}
	// sparse_mat33.h:258-266
	M33<T> get(uint32_t i, uint32_t j) const {
		const std::vector<Entry>& r = rows[i];
		for (size_t k = r.size(); k-- > 0;)
			if (r[k].col == j) return r[k].val;
		return M33<T>::Zero();
	}
	// Copy (sparse_mat33.h:210-230): reversed rows.
	static SpMat CopyOf(const SpMat& s) {
		SpMat d(s.n());
		for (uint32_t i = 0; i < s.n(); ++i) {
			const std::vector<Entry>& src = s.rows[i];
			std::vector<Entry>& dst = d.rows[i];
			dst.reserve(src.size());
			for (size_t k = src.size(); k-- > 0;) dst.push_back(src[k]);
		}
		return d;
	}
	// sparse_mat33.h:334-366
	void AddAssign(const SpMat& m) {
		for (uint32_t i = 0; i < m.n(); ++i)
			for (size_t k = m.rows[i].size(); k-- > 0;) {
				const Entry& e = m.rows[i][k];
				at(i, e.col) += e.val;
			}
	}
	void SubAssign(const SpMat& m) {
		for (uint32_t i = 0; i < m.n(); ++i)
			for (size_t k = m.rows[i].size(); k-- > 0;) {
				const Entry& e = m.rows[i][k];
				at(i, e.col) -= e.val;
			}
	}
};

// operator*(scalar, Sparse): sparse_mat33.h:514-533 (s == 0 leaves it empty)
template <class T> SpMat<T> Scale(T s, const SpMat<T>& B) {
	SpMat<T> out(B.n());
	if (s == T(0)) return out;
	out = SpMat<T>::CopyOf(B);
	for (auto& r : out.rows)
		for (auto& e : r) e.val = s * e.val;
	return out;
}
// operator-(Diag, Sparse): sparse_mat33.h:482-492
template <class T> SpMat<T> DiagMinus(const DMat<T>& a, const SpMat<T>& b) {
	SpMat<T> out(b.n());
	for (uint32_t i = 0; i < (uint32_t)a.size(); ++i) out.at(i, i) = a[i];
	out.SubAssign(b);
	return out;
}
// operator-(Sparse, Sparse): sparse_mat33.h:458-462
template <class T> SpMat<T> Minus(const SpMat<T>& a, const SpMat<T>& b) {
	SpMat<T> out = SpMat<T>::CopyOf(a);
	out.SubAssign(b);
	return out;
}
// operator*(Diag, Sparse): sparse_mat33.h:535-554
template <class T> SpMat<T> DiagTimes(const DMat<T>& A, const SpMat<T>& B) {
	SpMat<T> out = SpMat<T>::CopyOf(B);
	for (uint32_t i = 0; i < out.n(); ++i)
		for (auto& e : out.rows[i]) e.val = A[i] * e.val;
	return out;
}
// operator*(Sparse, Diag): sparse_mat33.h:556-576
template <class T> SpMat<T> TimesDiag(const SpMat<T>& A, const DMat<T>& B) {
	SpMat<T> out = SpMat<T>::CopyOf(A);
	for (uint32_t i = 0; i < out.n(); ++i)
		for (auto& e : out.rows[i]) e.val = e.val * B[e.col];
	return out;
}
// operator+(Sparse, Diag) / operator-(Sparse, Diag): sparse_mat33.h:464-474
template <class T> SpMat<T> PlusDiag(const SpMat<T>& a, const DMat<T>& b) {
	SpMat<T> out = SpMat<T>::CopyOf(a);
	for (uint32_t i = 0; i < (uint32_t)b.size(); ++i) out.at(i, i) += b[i];
	return out;
}
template <class T> SpMat<T> MinusDiag(const SpMat<T>& a, const DMat<T>& b) {
	SpMat<T> out = SpMat<T>::CopyOf(a);
	for (uint32_t i = 0; i < (uint32_t)b.size(); ++i) out.at(i, i) -= b[i];
	return out;
}
// b3Mul(DenseVec, Sparse, DenseVec): sparse_mat33.h:494-512
template <class T> DVec<T> SpMV(const SpMat<T>& A, const DVec<T>& v) {
	DVec<T> out(v.size());
	for (auto& o : out) o.SetZero();
	for (uint32_t i = 0; i < A.n(); ++i) {
		const auto& r = A.rows[i];
		for (size_t k = r.size(); k-- > 0;) out[i] += r[k].val * v[r[k].col];
	}
	return out;
}

// dense_vec3.h / diag_mat33.h elementwise operators
template <class T> DVec<T> VAdd(const DVec<T>& a, const DVec<T>& b) {
	DVec<T> o(a.size());
	for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] + b[i];
	return o;
}
template <class T> DVec<T> VSub(const DVec<T>& a, const DVec<T>& b) {
	DVec<T> o(a.size());
	for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] - b[i];
	return o;
}
template <class T> DVec<T> VScale(T s, const DVec<T>& b) {
	DVec<T> o(b.size());
	for (size_t i = 0; i < b.size(); ++i) o[i] = s * b[i];
	return o;
}
template <class T> DVec<T> DiagMul(const DMat<T>& a, const DVec<T>& b) {
	DVec<T> o(b.size());
	for (size_t i = 0; i < b.size(); ++i) o[i] = a[i] * b[i];
	return o;
}
// dense_vec3.h:151-163 (sequential accumulation)
template <class T> T VDot(const DVec<T>& a, const DVec<T>& b) {
	T r(0);
	for (size_t i = 0; i < a.size(); ++i) r += Dot(a[i], b[i]);
	return r;
}
// dense_vec3.h:165-175
template <class T> T VLengthSquared(const DVec<T>& v) {
	T r(0);
	for (size_t i = 0; i < v.size(); ++i) r += LengthSquared(v[i]);
	return r;
}
// The same two reductions with the accumulator widened to double (each term
// still formed in T, the sum rounded to T once): the dot products of the
// device's FP64 reduction mode, restated. Not a reference operation.
template <class T> T VDotWide(const DVec<T>& a, const DVec<T>& b) {
	double r = 0.0;
	for (size_t i = 0; i < a.size(); ++i) r += double(Dot(a[i], b[i]));
	return T(r);
}
template <class T> T VLengthSquaredWide(const DVec<T>& v) {
	double r = 0.0;
	for (size_t i = 0; i < v.size(); ++i) r += double(LengthSquared(v[i]));
	return T(r);
}

} // namespace b3o

#endif
