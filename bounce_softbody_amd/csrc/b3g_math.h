// Scalar math of the implicit step, usable on host and device.
//
// Every operator keeps the expression tree of the reference's math library
// (include/bounce_softbody/common/math/{vec2,vec3,mat22,mat33,quat,transform}.h)
// so that, compiled with -ffp-contract=off and IEEE division/sqrt, device
// results equal the reference's FP32 results bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define B3G_HD __host__ __device__ __forceinline__

namespace b3g {

constexpr float kEps = 1.1920928955078125e-7f;   // FLT_EPSILON (settings.h:54)
constexpr float kLinearSlop = 0.005f;             // B3_LINEAR_SLOP (settings.h:70)

B3G_HD float fabs_(float x) { return x < 0.0f ? -x : (x == 0.0f ? 0.0f : x); }
B3G_HD float fmin_(float a, float b) { return a < b ? a : b; }   // math.h:53-56
B3G_HD float fmax_(float a, float b) { return a > b ? a : b; }   // math.h:59-62
B3G_HD float clamp_(float a, float lo, float hi) { return fmax_(lo, fmin_(a, hi)); }
B3G_HD float sign_(float x) { return float(0.0f < x) - float(x < 0.0f); }
B3G_HD float lerp_(float x, float a, float b) { return (1.0f - x) * a + x * b; }
B3G_HD float sqrt_(float x) { return sqrtf(x); }

struct V2 { float x, y; };
B3G_HD V2 v2(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
B3G_HD V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
B3G_HD V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
B3G_HD V2 operator-(V2 a) { return v2(-a.x, -a.y); }
B3G_HD V2 operator*(float s, V2 v) { return v2(s * v.x, s * v.y); }
B3G_HD float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
B3G_HD float det(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
// vec2.h Normalize: divides by the length
B3G_HD void normalize_div(V2& v) {
	float len = sqrt_(v.x * v.x + v.y * v.y);
	if (len > kEps) { v.x /= len; v.y /= len; }
}

struct M22 { V2 x, y; };
B3G_HD V2 operator*(const M22& A, V2 v) { return v.x * A.x + v.y * A.y; }
B3G_HD M22 operator*(const M22& A, const M22& B) { M22 r; r.x = A * B.x; r.y = A * B.y; return r; }
B3G_HD M22 operator+(const M22& A, const M22& B) { M22 r; r.x = A.x + B.x; r.y = A.y + B.y; return r; }
B3G_HD M22 operator*(float s, const M22& A) { M22 r; r.x = s * A.x; r.y = s * A.y; return r; }
B3G_HD M22 transpose(const M22& A) { M22 r; r.x = v2(A.x.x, A.y.x); r.y = v2(A.x.y, A.y.y); return r; }
B3G_HD M22 adjugate(const M22& A) { M22 r; r.x = v2(A.y.y, -A.x.y); r.y = v2(-A.y.x, A.x.x); return r; }

struct V3 { float x, y, z; };
B3G_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
B3G_HD V3 zero3() { return v3(0.0f, 0.0f, 0.0f); }
B3G_HD float comp(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
B3G_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
B3G_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
B3G_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
B3G_HD V3 operator*(float s, V3 v) { return v3(s * v.x, s * v.y, s * v.z); }
B3G_HD V3 operator/(V3 v, float s) { return v3(v.x / s, v.y / s, v.z / s); }
B3G_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
B3G_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
B3G_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
B3G_HD float det(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }
B3G_HD float length(V3 v) { return sqrt_(v.x * v.x + v.y * v.y + v.z * v.z); }
B3G_HD float length_sq(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
// vec3.h:93-104 (member Normalize: multiply by reciprocal)
B3G_HD void normalize_inplace(V3& v) {
	float len = sqrt_(v.x * v.x + v.y * v.y + v.z * v.z);
	if (len > kEps) { float s = 1.0f / len; v.x *= s; v.y *= s; v.z *= s; }
}
// vec3.h:195-204
B3G_HD V3 normalized(V3 v) {
	float len = length(v);
	if (len > kEps) { float s = 1.0f / len; return s * v; }
	return v;
}
B3G_HD V3 vmin(V3 a, V3 b) { return v3(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
B3G_HD V3 vmax(V3 a, V3 b) { return v3(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }
B3G_HD V3 vclamp(V3 v, V3 lo, V3 hi) { return vmax(lo, vmin(v, hi)); }
B3G_HD V3 vabs(V3 v) { return v3(fabs_(v.x), fabs_(v.y), fabs_(v.z)); }
// vec3.h:255-271
B3G_HD V3 perp(V3 v) {
	const float sqrt_inv3 = sqrt_(1.0f / 3.0f);
	V3 u = fabs_(v.x) >= sqrt_inv3 ? v3(v.y, -v.x, 0.0f) : v3(0.0f, v.z, -v.y);
	normalize_inplace(u);
	return u;
}

// Column-major 3x3 (mat33.h:25-96)
struct M33 { V3 x, y, z; };
B3G_HD M33 m33(V3 a, V3 b, V3 c) { M33 r; r.x = a; r.y = b; r.z = c; return r; }
B3G_HD M33 mzero() { return m33(zero3(), zero3(), zero3()); }
B3G_HD M33 mdiag(float s) { return m33(v3(s, 0.0f, 0.0f), v3(0.0f, s, 0.0f), v3(0.0f, 0.0f, s)); }
B3G_HD M33 mdiag3(float a, float b, float c) { return m33(v3(a, 0.0f, 0.0f), v3(0.0f, b, 0.0f), v3(0.0f, 0.0f, c)); }
B3G_HD float& at(M33& A, int r, int c) { return (&A.x.x)[r + 3 * c]; }
B3G_HD M33 operator+(const M33& A, const M33& B) { return m33(A.x + B.x, A.y + B.y, A.z + B.z); }
B3G_HD M33 operator-(const M33& A, const M33& B) { return m33(A.x - B.x, A.y - B.y, A.z - B.z); }
B3G_HD M33 operator*(float s, const M33& A) { return m33(s * A.x, s * A.y, s * A.z); }
B3G_HD M33 operator-(const M33& A) { return -1.0f * A; }
B3G_HD V3 operator*(const M33& A, V3 v) { return v.x * A.x + v.y * A.y + v.z * A.z; }   // mat33.h:129-132
B3G_HD M33 operator*(const M33& A, const M33& B) { return m33(A * B.x, A * B.y, A * B.z); }
B3G_HD M33 transpose(const M33& A) {
	return m33(v3(A.x.x, A.y.x, A.z.x), v3(A.x.y, A.y.y, A.z.y), v3(A.x.z, A.y.z, A.z.z));
}
B3G_HD M33 outer(V3 a, V3 b) { return m33(b.x * a, b.y * a, b.z * a); }              // mat33.h:303-306
B3G_HD M33 adjugate(const M33& A) { return transpose(m33(cross(A.y, A.z), cross(A.z, A.x), cross(A.x, A.y))); }
B3G_HD M33 mabs(const M33& A) { return m33(vabs(A.x), vabs(A.y), vabs(A.z)); }
// mat33.h:310-329
B3G_HD void compute_basis(V3 a, V3& b, V3& c) {
	const float sqrt_inv3 = sqrt_(1.0f / 3.0f);
	b = fabs_(a.x) >= sqrt_inv3 ? v3(a.y, -a.x, 0.0f) : v3(0.0f, a.z, -a.y);
	normalize_inplace(b);
	c = cross(a, b);
}

// math.h:135-151, generic column-major product C = A (AM x AN) * B (AN x BN)
template <int AM, int AN, int BN>
B3G_HD void gmul(float* C, const float* A, const float* B) {
	for (int i = 0; i < AM; ++i)
		for (int j = 0; j < BN; ++j) {
			float acc = 0.0f;
			for (int k = 0; k < AN; ++k) acc += A[i + AM * k] * B[k + AN * j];
			C[i + AM * j] = acc;
		}
}

struct Q4 { V3 v; float s; };
B3G_HD Q4 qidentity() { Q4 q; q.v = zero3(); q.s = 1.0f; return q; }
// quat.h:63-72
B3G_HD void qnormalize(Q4& q) {
	float len = sqrt_(dot(q.v, q.v) + q.s * q.s);
	if (len > kEps) {
		float s = 1.0f / len;
		q.v.x *= s; q.v.y *= s; q.v.z *= s;
		q.s /= len;
	}
}
B3G_HD V3 qrotate(const Q4& q, V3 vin) {   // quat.h:75-79
	V3 t = 2.0f * cross(q.v, vin);
	return vin + q.s * t + cross(q.v, t);
}
B3G_HD Q4 qmul(const Q4& a, const Q4& b) {  // quat.h:204-210
	Q4 r;
	r.v = cross(a.v, b.v) + a.s * b.v + b.s * a.v;
	r.s = a.s * b.s - dot(a.v, b.v);
	return r;
}
B3G_HD Q4 qconj(const Q4& q) { Q4 r; r.v = -q.v; r.s = q.s; return r; }
B3G_HD M33 qmatrix(const Q4& q) {           // quat.h:133-146
	float x = q.v.x, y = q.v.y, z = q.v.z, w = q.s;
	float x2 = x + x, y2 = y + y, z2 = z + z;
	float xx = x * x2, xy = x * y2, xz = x * z2;
	float yy = y * y2, yz = y * z2, zz = z * z2;
	float wx = w * x2, wy = w * y2, wz = w * z2;
	return m33(v3(1.0f - (yy + zz), xy + wz, xz - wy), v3(xy - wz, 1.0f - (xx + zz), yz + wx),
	           v3(xz + wy, yz - wx, 1.0f - (xx + yy)));
}

struct Xf { V3 p; Q4 q; };
B3G_HD V3 xmul(const Xf& X, V3 v) { return qrotate(X.q, v) + X.p; }
B3G_HD V3 xmulT(const Xf& X, V3 v) { return qrotate(qconj(X.q), v - X.p); }

// FP64 counterparts for the precision form's tetrahedron forces: the same
// expression trees in double (the reference with its scalar type set to
// double; normalisation threshold DBL_EPSILON).
constexpr double kEps64 = 2.220446049250313e-16;
struct D3 { double x, y, z; };
B3G_HD D3 d3(double x, double y, double z) { D3 r; r.x = x; r.y = y; r.z = z; return r; }
B3G_HD D3 d3(V3 v) { return d3(v.x, v.y, v.z); }
B3G_HD D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
B3G_HD D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
B3G_HD D3 operator-(D3 a) { return d3(-a.x, -a.y, -a.z); }
B3G_HD D3 operator*(double s, D3 v) { return d3(s * v.x, s * v.y, s * v.z); }
B3G_HD D3 operator/(D3 v, double s) { return d3(v.x / s, v.y / s, v.z / s); }
B3G_HD double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
B3G_HD D3 cross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
B3G_HD double det(D3 a, D3 b, D3 c) { return dot(a, cross(b, c)); }
B3G_HD double length(D3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
struct DM33 { D3 x, y, z; };
B3G_HD DM33 dm33(D3 a, D3 b, D3 c) { DM33 r; r.x = a; r.y = b; r.z = c; return r; }
B3G_HD DM33 dm33(const M33& A) { return dm33(d3(A.x), d3(A.y), d3(A.z)); }
B3G_HD D3 operator*(const DM33& A, D3 v) { return v.x * A.x + v.y * A.y + v.z * A.z; }
B3G_HD DM33 operator*(const DM33& A, const DM33& B) { return dm33(A * B.x, A * B.y, A * B.z); }
B3G_HD DM33 operator*(double s, const DM33& A) { return dm33(s * A.x, s * A.y, s * A.z); }
B3G_HD DM33 transpose(const DM33& A) {
	return dm33(d3(A.x.x, A.y.x, A.z.x), d3(A.x.y, A.y.y, A.z.y), d3(A.x.z, A.y.z, A.z.z));
}
B3G_HD DM33 adjugate(const DM33& A) { return transpose(dm33(cross(A.y, A.z), cross(A.z, A.x), cross(A.x, A.y))); }
struct DQ4 { D3 v; double s; };
B3G_HD void qnormalize(DQ4& q) {
	double len = sqrt(dot(q.v, q.v) + q.s * q.s);
	if (len > kEps64) {
		double s = 1.0 / len;
		q.v.x *= s; q.v.y *= s; q.v.z *= s;
		q.s /= len;
	}
}
B3G_HD DQ4 qmul(const DQ4& a, const DQ4& b) {
	DQ4 r;
	r.v = cross(a.v, b.v) + a.s * b.v + b.s * a.v;
	r.s = a.s * b.s - dot(a.v, b.v);
	return r;
}
B3G_HD DM33 qmatrix(const DQ4& q) {
	double x = q.v.x, y = q.v.y, z = q.v.z, w = q.s;
	double x2 = x + x, y2 = y + y, z2 = z + z;
	double xx = x * x2, xy = x * y2, xz = x * z2;
	double yy = y * y2, yz = y * z2, zz = z * z2;
	double wx = w * x2, wy = w * y2, wz = w * z2;
	return dm33(d3(1.0 - (yy + zz), xy + wz, xz - wy), d3(xy - wz, 1.0 - (xx + zz), yz + wx),
	            d3(xz + wy, yz - wx, 1.0 - (xx + yy)));
}

struct AABB { V3 lo, hi; };
B3G_HD AABB aabb_cr(V3 c, float r) { V3 rv = v3(r, r, r); AABB a; a.lo = c - rv; a.hi = c + rv; return a; }
B3G_HD AABB aabb_cr(V3 c, V3 r) { AABB a; a.lo = c - r; a.hi = c + r; return a; }
B3G_HD void aabb_extend(AABB& a, float s) { V3 r = v3(s, s, s); a.lo = a.lo - r; a.hi = a.hi + r; }
B3G_HD V3 aabb_dims(const AABB& a) { return v3(a.hi.x - a.lo.x, a.hi.y - a.lo.y, a.hi.z - a.lo.z); }
B3G_HD bool overlap(const AABB& a, const AABB& b) {   // aabb.h:252-272
	if (a.hi.x < b.lo.x || a.lo.x > b.hi.x) return false;
	if (a.hi.y < b.lo.y || a.lo.y > b.hi.y) return false;
	if (a.hi.z < b.lo.z || a.lo.z > b.hi.z) return false;
	return true;
}

} // namespace b3g
