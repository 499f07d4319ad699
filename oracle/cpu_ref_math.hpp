/*
 * TEST INFRASTRUCTURE (oracle). Scalar math of the reference restated with
 * the same operation order, templated on the scalar type so that the same
 * code serves as the FP32 oracle (bit-exact target) and the FP64 oracle.
 *
 * Every operator below follows the reference expression tree exactly; the
 * translation unit is compiled with -ffp-contract=off so that no product is
 * fused into an add. Citations are to the reference (include/bounce_softbody/).
 */
#ifndef B3O_CPU_REF_MATH_HPP
#define B3O_CPU_REF_MATH_HPP

#include <cmath>
#include <cstdint>
#include <limits>

namespace b3o {

template <class T> struct Eps { static T value() { return std::numeric_limits<T>::epsilon(); } }; // settings.h:54

template <class T> inline T Abs(T x) { return std::abs(x); }
template <class T> inline T Sqrt(T x) { return std::sqrt(x); }
template <class T> inline T Min(T a, T b) { return a < b ? a : b; }            // math.h:53-56
template <class T> inline T Max(T a, T b) { return a > b ? a : b; }            // math.h:59-62
template <class T> inline T Clamp(T a, T lo, T hi) { return Max(lo, Min(a, hi)); } // math.h:65-68
template <class T> inline T Sign(T x) { return T(T(0) < x) - T(T(x) < T(0)); }  // math.h:85-88
template <class T> inline T Lerp(T x, T a, T b) { return (T(1) - x) * a + x * b; } // math.h:71-74

// The reference calls the double-precision libm sin/cos on the (promoted)
// half angle, then narrows (quat.h:83-88).
template <class T> inline T SinNarrow(T t) { return T(std::sin(double(t))); }
template <class T> inline T CosNarrow(T t) { return T(std::cos(double(t))); }

// ---------------------------------------------------------------- vec2.h
template <class T> struct V2 {
	T x, y;
	void SetZero() { x = y = T(0); }
	// vec2.h Normalize divides by the length.
	T Normalize() {
		T len = Sqrt(x * x + y * y);
		if (len > Eps<T>::value()) { x /= len; y /= len; }
		return len;
	}
};
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline V2<T> operator-(V2<T> a) { return {-a.x, -a.y}; }
template <class T> inline V2<T> operator*(T s, V2<T> v) { return {s * v.x, s * v.y}; }
template <class T> inline T Dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T Det(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }

// ---------------------------------------------------------------- mat22.h (column major)
template <class T> struct M22 {
	V2<T> x, y;
	T& at(int r, int c) { return (&x.x)[r + 2 * c]; }
	T at(int r, int c) const { return (&x.x)[r + 2 * c]; }
};
template <class T> inline V2<T> operator*(const M22<T>& A, V2<T> v) { return v.x * A.x + v.y * A.y; }
template <class T> inline M22<T> operator*(const M22<T>& A, const M22<T>& B) { return {A * B.x, A * B.y}; }
template <class T> inline M22<T> operator+(const M22<T>& A, const M22<T>& B) { return {A.x + B.x, A.y + B.y}; }
template <class T> inline M22<T> operator*(T s, const M22<T>& A) { return {s * A.x, s * A.y}; }
template <class T> inline M22<T> Transpose(const M22<T>& A) { return {{A.x.x, A.y.x}, {A.x.y, A.y.y}}; }
template <class T> inline M22<T> Adjugate(const M22<T>& A) { return {{A.y.y, -A.x.y}, {-A.y.x, A.x.x}}; } // mat22.h

// ---------------------------------------------------------------- vec3.h
template <class T> struct V3 {
	T x, y, z;
	T operator[](int i) const { return (&x)[i]; }
	T& operator[](int i) { return (&x)[i]; }
	void SetZero() { x = y = z = T(0); }
	void operator+=(V3 b) { x += b.x; y += b.y; z += b.z; }
	void operator-=(V3 b) { x -= b.x; y -= b.y; z -= b.z; }
	void operator*=(T s) { x *= s; y *= s; z *= s; }
	// vec3.h:70-76: multiplies by the reciprocal
	void operator/=(T a) { T s = T(1) / a; x *= s; y *= s; z *= s; }
	// vec3.h:93-104
	T Normalize() {
		T length = Sqrt(x * x + y * y + z * z);
		if (length > Eps<T>::value()) { T s = T(1) / length; x *= s; y *= s; z *= s; }
		return length;
	}
};
template <class T> inline V3<T> MakeV3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> inline V3<T> operator-(V3<T> v) { return {-v.x, -v.y, -v.z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> v) { return {s * v.x, s * v.y, s * v.z}; }
template <class T> inline V3<T> operator/(V3<T> v, T s) { return {v.x / s, v.y / s, v.z / s}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline T Dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> Cross(V3<T> a, V3<T> b) {
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T Det(V3<T> a, V3<T> b, V3<T> c) { return Dot(a, Cross(b, c)); }
template <class T> inline T Length(V3<T> v) { return Sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
template <class T> inline T LengthSquared(V3<T> v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
// vec3.h:195-204 (free function: returns v unchanged when short)
template <class T> inline V3<T> Normalize(V3<T> v) {
	T length = Length(v);
	if (length > Eps<T>::value()) { T s = T(1) / length; return s * v; }
	return v;
}
template <class T> inline V3<T> Min(V3<T> a, V3<T> b) { return {Min(a.x, b.x), Min(a.y, b.y), Min(a.z, b.z)}; }
template <class T> inline V3<T> Max(V3<T> a, V3<T> b) { return {Max(a.x, b.x), Max(a.y, b.y), Max(a.z, b.z)}; }
template <class T> inline V3<T> Clamp(V3<T> v, V3<T> lo, V3<T> hi) { return Max(lo, Min(v, hi)); }
template <class T> inline V3<T> Abs(V3<T> v) { return {Abs(v.x), Abs(v.y), Abs(v.z)}; }
// vec3.h:255-271
template <class T> inline V3<T> Perp(V3<T> v) {
	const T sqrt_inv3 = Sqrt(T(1) / T(3));
	V3<T> u;
	if (Abs(v.x) >= sqrt_inv3) u = MakeV3(v.y, -v.x, T(0));
	else u = MakeV3(T(0), v.z, -v.y);
	u.Normalize();
	return u;
}

// ---------------------------------------------------------------- mat33.h (column major)
template <class T> struct M33 {
	V3<T> x, y, z;
	T& at(int r, int c) { return (&x.x)[r + 3 * c]; }
	T at(int r, int c) const { return (&x.x)[r + 3 * c]; }
	void operator+=(const M33& B) { x += B.x; y += B.y; z += B.z; }
	void operator-=(const M33& B) { x -= B.x; y -= B.y; z -= B.z; }
	static M33 Zero() { M33 m; m.x.SetZero(); m.y.SetZero(); m.z.SetZero(); return m; }
	static M33 Identity() { return Diagonal(T(1)); }
	static M33 Diagonal(T s) { return {{s, T(0), T(0)}, {T(0), s, T(0)}, {T(0), T(0), s}}; }
	static M33 Diagonal(T a, T b, T c) { return {{a, T(0), T(0)}, {T(0), b, T(0)}, {T(0), T(0), c}}; }
};
template <class T> inline M33<T> operator+(const M33<T>& A, const M33<T>& B) { return {A.x + B.x, A.y + B.y, A.z + B.z}; }
template <class T> inline M33<T> operator-(const M33<T>& A, const M33<T>& B) { return {A.x - B.x, A.y - B.y, A.z - B.z}; }
template <class T> inline M33<T> operator*(T s, const M33<T>& A) { return {s * A.x, s * A.y, s * A.z}; }
template <class T> inline M33<T> operator-(const M33<T>& A) { return T(-1) * A; }                 // mat33.h:123-126
template <class T> inline V3<T> operator*(const M33<T>& A, V3<T> v) { return v.x * A.x + v.y * A.y + v.z * A.z; } // mat33.h:129-132
template <class T> inline M33<T> operator*(const M33<T>& A, const M33<T>& B) { return {A * B.x, A * B.y, A * B.z}; }
template <class T> inline M33<T> Transpose(const M33<T>& A) {
	return {{A.x.x, A.y.x, A.z.x}, {A.x.y, A.y.y, A.z.y}, {A.x.z, A.y.z, A.z.z}};
}
template <class T> inline M33<T> Outer(V3<T> a, V3<T> b) { return {b.x * a, b.y * a, b.z * a}; } // mat33.h:303-306
// mat33.h:204-279
template <class T> inline M33<T> Adjugate(const M33<T>& A) {
	M33<T> C;
	C.x = Cross(A.y, A.z);
	C.y = Cross(A.z, A.x);
	C.z = Cross(A.x, A.y);
	return Transpose(C);
}
template <class T> inline M33<T> Abs(const M33<T>& A) { return {Abs(A.x), Abs(A.y), Abs(A.z)}; }
// mat33.h:310-329
template <class T> inline void ComputeBasis(V3<T> a, V3<T>& b, V3<T>& c) {
	const T sqrt_inv3 = Sqrt(T(1) / T(3));
	if (Abs(a.x) >= sqrt_inv3) b = MakeV3(a.y, -a.x, T(0));
	else b = MakeV3(T(0), a.z, -a.y);
	b.Normalize();
	c = Cross(a, b);
}

// math.h:135-151 : generic column-major product, C = A(AMxAN) * B(BMxBN)
template <class T> inline void GMul(T* C, const T* A, int AM, int AN, const T* B, int BM, int BN) {
	(void)BM;
	for (int i = 0; i < AM; ++i)
		for (int j = 0; j < BN; ++j) {
			C[i + AM * j] = T(0);
			for (int k = 0; k < AN; ++k) C[i + AM * j] += A[i + AM * k] * B[k + BM * j];
		}
}
// math.h:155-164
template <class T> inline void GTranspose(T* B, const T* A, int AM, int AN) {
	for (int i = 0; i < AM; ++i)
		for (int j = 0; j < AN; ++j) B[j + AN * i] = A[i + AM * j];
}

// ---------------------------------------------------------------- quat.h
template <class T> struct Q4 {
	V3<T> v;
	T s;
	void SetIdentity() { v.SetZero(); s = T(1); }
	// quat.h:63-72: vector part by reciprocal multiply, scalar part divides
	T Normalize() {
		T len = Sqrt(Dot(v, v) + s * s);
		if (len > Eps<T>::value()) { v /= len; s /= len; }
		return len;
	}
	V3<T> Rotate(V3<T> vIn) const { // quat.h:75-79
		V3<T> t = T(2) * Cross(v, vIn);
		return vIn + s * t + Cross(v, t);
	}
	void SetAxisAngle(V3<T> axis, T angle) { // quat.h:83-88
		T theta = T(0.5) * angle;
		v = SinNarrow(theta) * axis;
		s = CosNarrow(theta);
	}
	M33<T> GetRotationMatrix() const { // quat.h:133-146
		T x = v.x, y = v.y, z = v.z, w = s;
		T x2 = x + x, y2 = y + y, z2 = z + z;
		T xx = x * x2, xy = x * y2, xz = x * z2;
		T yy = y * y2, yz = y * z2, zz = z * z2;
		T wx = w * x2, wy = w * y2, wz = w * z2;
		return {{T(1) - (yy + zz), xy + wz, xz - wy},
		        {xy - wz, T(1) - (xx + zz), yz + wx},
		        {xz + wy, yz - wx, T(1) - (xx + yy)}};
	}
};
template <class T> inline Q4<T> QMul(const Q4<T>& a, const Q4<T>& b) { // quat.h:204-210
	Q4<T> r;
	r.v = Cross(a.v, b.v) + a.s * b.v + b.s * a.v;
	r.s = a.s * b.s - Dot(a.v, b.v);
	return r;
}
template <class T> inline Q4<T> Conjugate(const Q4<T>& q) { Q4<T> r; r.v = -q.v; r.s = q.s; return r; }

// transform.h: translation + rotation
template <class T> struct Xf {
	V3<T> p;
	Q4<T> q;
};
template <class T> inline V3<T> XMul(const Xf<T>& X, V3<T> v) { return X.q.Rotate(v) + X.p; }               // transform.h b3Mul
template <class T> inline V3<T> XMulT(const Xf<T>& X, V3<T> v) { return Conjugate(X.q).Rotate(v - X.p); } // transform.h b3MulT

// aabb.h
template <class T> struct AABB {
	V3<T> lo, hi;
	static AABB FromCenterRadius(V3<T> c, T r) { V3<T> rv = MakeV3(r, r, r); return {c - rv, c + rv}; }
	static AABB FromCenterRadius(V3<T> c, V3<T> r) { return {c - r, c + r}; }
	void Extend(T s) { V3<T> r = MakeV3(s, s, s); lo -= r; hi += r; }
	V3<T> Dimensions() const { return MakeV3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z); }
	V3<T> Relative(V3<T> p) const { return (p - lo) / Dimensions(); }
	// aabb.h:202-217
	void Transform(const Xf<T>& X) {
		V3<T> lc = T(0.5) * (lo + hi);
		V3<T> lr = T(0.5) * (hi - lo);
		V3<T> c = XMul(X, lc);
		M33<T> ar = Abs(X.q.GetRotationMatrix());
		V3<T> r;
		r.x = Dot(lr, ar.x);
		r.y = Dot(lr, ar.y);
		r.z = Dot(lr, ar.z);
		lo = c - r;
		hi = c + r;
	}
};
// aabb.h:252-272
template <class T> inline bool TestOverlap(const AABB<T>& a, const AABB<T>& b) {
	if (a.hi.x < b.lo.x || a.lo.x > b.hi.x) return false;
	if (a.hi.y < b.lo.y || a.lo.y > b.hi.y) return false;
	if (a.hi.z < b.lo.z || a.lo.z > b.hi.z) return false;
	return true;
}

} // namespace b3o

#endif
