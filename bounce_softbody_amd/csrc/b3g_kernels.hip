// Device kernels of the implicit step and their launch sequence.
//
// Data layout (HBM): particle vectors are arrays of 12-byte V3 in list order;
// element tables are planar (one plane per scalar, element-contiguous) so a
// wavefront reads 64 consecutive elements per load; pA is stored as 9 planes
// of its blocks in CSR order (the reference's row order); element
// contributions are staged once per linearisation and gathered per block in
// the reference's summation order (no atomics, deterministic).
#include "b3g_collide.h"
#include "b3g_tet.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace b3g {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kTileCap = 1024;     // 3x3 blocks per SpMV tile chunk (12 KB of LDS products)
constexpr uint32_t kFuseMaxPartials = 6144;   // default FP64 mode: fused reductions up to this many partials per instance
constexpr int kTileRows = 256;     // rows per SpMV tile (one phase-2 thread each)
constexpr int kRowPer = kTileCap / kBlock;   // chunk entries per thread

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G) {
	// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
	// contiguous range of tiles so neighbouring rows share one L2.
	uint32_t q = G / 8, r = G % 8, x = b % 8;
	uint32_t start = x * q + (x < r ? x : r);
	return start + b / 8;
}

// Holds the uniform values v... in scalar registers at this point, so their
// loads (independent of one another) are all issued ahead of it and waited
// for once; without it the compiler moves each load down to its first use,
// behind the branches before it, and the loads run one round trip each.
#define B3G_SETTLE(...) asm volatile("" ::__VA_ARGS__)

// The Newton gate of an element kernel: the iteration of the instance of the
// element's first particle is active. st null: the step's first Newton
// iteration, which every instance runs, so nothing is read.
__device__ __forceinline__ bool element_gate(const SolverState* st, const uint32_t* rowInst, const uint32_t* p,
                                             uint32_t e) {
	return !st || st[rowInst[p[e]]].newtonActive != 0;
}

// Work items (tiles or row blocks) [0, count) over the launch's workgroups.
// Multi = false: one workgroup per item, item = xcd_tile(block) (Xcd) or the
// block index, checked by every thread (the kernels of a single body).
// Multi = true: workgroup v = xcd_tile(block) takes items v, v + G, v + 2G, ...
// (G workgroups), so at any time the workgroups of one XCD work on one run
// of adjacent items, whose vector gathers share that XCD's L2, as with one
// workgroup per item. The items whose instance runs -- ok(instance of the
// item) -- are found 64 at a time with one ballot, so a batch whose instances
// have mostly converged costs a few loads per workgroup rather than a
// workgroup per item. ok() and the item's instance are uniform; body(item)
// may synchronise the workgroup.
template <bool Multi, bool Xcd, class Ok, class Body>
__device__ __forceinline__ void for_items(uint32_t count, const uint32_t* itemInst, Ok ok, Body body) {
	if (!Multi) {
		const uint32_t t = Xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
		if (ok(itemInst[t])) body(t);
		return;
	}
	const uint32_t G = gridDim.x, v = xcd_tile(blockIdx.x, G);
	__shared__ uint64_t mask;
	for (uint32_t p = v; p < count; p += 64 * G) {
		if (threadIdx.x < 64) {
			const uint32_t t = p + threadIdx.x * G;
			const uint64_t m = __ballot(t < count && ok(itemInst[t]));
			if (threadIdx.x == 0) mask = m;
		}
		__syncthreads();
		const uint64_t w = mask;   // uniform: kept in scalar registers
		uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(w >> 32)) << 32) |
		             (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)w);
		__syncthreads();   // the next pass rewrites the mask
		while (m) {
			const uint32_t b = (uint32_t)__ffsll((unsigned long long)m) - 1u;
			m &= m - 1;
			body(p + b * G);
		}
	}
}

// Whether the persistent kernel runs instance inst's PCG iterations (persist: per-instance flags, or null).
__device__ __forceinline__ bool persistent(const uint32_t* persist, uint32_t inst) { return persist && persist[inst]; }

__device__ __forceinline__ M33 ldA(const float* A, uint32_t stride, uint32_t k) {
	M33 a;
	float* p = &a.x.x;
#pragma unroll
	for (int c = 0; c < 9; ++c) p[c] = A[(size_t)c * stride + k];
	return a;
}
__device__ __forceinline__ void stA(float* A, uint32_t stride, uint32_t k, const M33& a) {
	const float* p = &a.x.x;
#pragma unroll
	for (int c = 0; c < 9; ++c) A[(size_t)c * stride + k] = p[c];
}
__device__ __forceinline__ M33 ldM(const float* P, uint32_t m, uint32_t plane0, uint32_t e) {
	M33 a;
	float* p = &a.x.x;
#pragma unroll
	for (int c = 0; c < 9; ++c) p[c] = P[(size_t)(plane0 + c) * m + e];
	return a;
}
__device__ __forceinline__ V3 ldv(const V3* p, uint32_t i) { return p[i]; }

// Wave64 + LDS block sum of doubles (fixed tree -> deterministic).
__device__ double block_sum(double v, double* sh) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();
	if (lane == 0) sh[wave] = v;
	__syncthreads();
	double t = 0.0;
	if (threadIdx.x == 0)
		for (int k = 0; k < kBlock / 64; ++k) t += sh[k];
	return t;   // valid in thread 0
}

enum Finalize : int { kFinAlpha = 0, kFinDelta = 1, kFinInit = 2, kFinNewton = 3, kFinRefine0 = 4, kFinRefine = 5 };

// The PCG's state machine (sparse_solver.cpp:57-103) over the per-iteration
// log: Init fills entry 0, Alpha(k) entry k's alpha, Delta(k) entry k from
// the reduction that follows iteration k - 1. `it` counts Delta steps, as the
// reference's iteration counter does.
__device__ void finalize(SolverState* st, CGLog* log, int kind, uint32_t k, double total) {
	float v = (float)total;
	switch (kind) {
	case kFinAlpha:   // sparse_solver.cpp:77
		log[k].alpha = log[k].delta / v;
		break;
	case kFinInit:    // sparse_solver.cpp:57-73 (a refinement sweep solves only while the sweeps are active)
		st->delta0 = v;
		st->it = 0;
		log[0].delta = v;
		log[0].beta = 0.0f;
		log[0].active = (st->maxIt != 0 && !(v <= st->eps2_sub * v) && (!st->refineOn || st->refineActive)) ? 1u : 0u;
		break;
	case kFinRefine0: // first FP64 residual of a Newton iteration: a sweep follows unless it is zero
		st->res0 = total;
		st->res = total;
		st->sweeps = 0;
		st->itSum = 0;
		st->refineActive = (st->maxSweeps != 0 && total > 0.0) ? 1u : 0u;
		break;
	case kFinRefine:  // residual after a sweep: count the sweep and its inner iterations, test convergence
		st->itSum += st->it;
		st->it = 0;
		if (st->refineActive) {
			st->res = total;
			st->sweeps += 1;
			st->refineActive = (st->sweeps < st->maxSweeps && !(total <= (double)st->refineTol2 * st->res0)) ? 1u : 0u;
		}
		break;
	case kFinDelta: { // sparse_solver.cpp:92-102
		log[k].delta = v;
		log[k].beta = v / log[k - 1].delta;
		log[k].active = (k != st->maxIt && !(v <= st->eps2_sub * st->delta0)) ? 1u : 0u;
		st->it = k;
		break;
	}
	case kFinNewton: { // sparse_force_solver.cpp:196-225
		st->error = v;
		const uint32_t its = st->refineOn ? st->itSum + st->it : st->it;
		st->statMinSub = min(st->statMinSub, its);
		st->statMaxSub = max(st->statMaxSub, its);
		if (st->refineOn) st->statSweeps = st->sweeps;
		bool conv;
		if (st->newtonIt == 0) { st->error0 = v; conv = v <= st->eps2; }
		else conv = v <= st->eps2 * st->error0;
		if (conv) st->newtonActive = 0;
		else { st->newtonIt += 1; st->newtonActive = st->newtonIt < st->newtonMax ? 1u : 0u; }
		st->statNewton = st->newtonIt;
		break;
	}
	}
}

// Reductions of one scalar per row over an instance. The producing kernel
// stores one FP64 partial per tile / row block (fixed tree) and, in
// sequential mode, each row's FP32 term. The sum is formed either by
// k_finalize (one workgroup per instance) or, for the PCG's alpha and delta
// in FP64 mode, by every workgroup of the consuming kernel in the same
// fixed order (wg_sum). The kernel boundary orders the stores before the
// loads, so no device-scope fence is needed (on a multi-XCD part such a
// fence writes back and invalidates a whole L2).
//   FP64 mode: deterministic, near-exact sum of the partials.
//   Sequential mode: the terms are added one by one in row order starting
//   from 0, which is the reference's b3Dot / b3LengthSquared accumulation
//   (dense_vec3.h:151-175).
struct Reduce {
	double* partial;
	float* terms;
	SolverState* st;         // per instance
	const uint32_t* range;   // per instance: tiles, row blocks, rows (begin, end)
	CGLog* log;              // per instance, logStride entries
	uint32_t logStride;
	const uint32_t* persist; // instances whose iterations the persistent kernel reduces itself, or null
	uint32_t* tileGate;      // batches: per-tile "instance done with this solve" words, reset at a solve's start
	uint32_t* blkGate;       // the same per row block of the vector update
};

// The workgroup's FP64 partial (each thread's `part` added in the block
// tree's fixed order) into partial[slot].
__device__ __forceinline__ void block_store(double part, double* partial, uint32_t slot) {
	__shared__ double sh[kBlock / 64];
	const double bs = block_sum(part, sh);
	if (threadIdx.x == 0) partial[slot] = bs;
}

// Row blocks of the vector kernels: kRowBlkRows rows per workgroup, thread x
// taking rows begin + x + u kBlock (u < kRowsPerThread), so a partial covers
// four times as many rows as a workgroup has threads and the kernels that
// re-read all of an instance's partials (fused reductions) read fewer.
constexpr int kRowsPerThread = 4;
constexpr uint32_t kRowBlkRows = kRowsPerThread * kBlock;

// One partial per tile / row block `slot` of instance `inst`.
__device__ __forceinline__ void block_partial(float term, bool has, uint32_t row, double* partial, float* terms,
                                              bool seq, uint32_t slot) {
	__shared__ double sh[kBlock / 64];
	if (seq && has) terms[row] = term;
	double bs = block_sum(has ? (double)term : 0.0, sh);
	if (threadIdx.x == 0) partial[slot] = bs;
}

constexpr int kFinBlock = 256;

// The first pass of wg_sum's loads (partials p0 + x + j kFinBlock, j < 16),
// issued on their own so that a kernel can put other loads in flight before
// it waits for them. A lane past the range loads nothing (a whole wave past
// it skips the load), so a short range costs few loads.
__device__ __forceinline__ void wg_sum_load(const double* partial, uint32_t p0, uint32_t p1, double (&p)[16]) {
#pragma unroll
	for (int u = 0; u < 16; ++u) {
		const uint32_t k = p0 + threadIdx.x + u * kFinBlock;
		p[u] = k < p1 ? partial[k] : 0.0;
	}
}

// Finishes wg_sum from its first pass's values (the same additions in the
// same order as wg_sum, so the same result).
__device__ double wg_sum_finish(const double* partial, uint32_t p0, uint32_t p1, const double (&first)[16]) {
	__shared__ double sh[kFinBlock / 64];
	__shared__ double total;
	// thread x adds partial[p0 + x + j kFinBlock] for j = 0, 1, ... in order;
	// 16 loads per pass keep up to 16 kFinBlock partials one L2 round trip
	double t = 0.0;
#pragma unroll
	for (int u = 0; u < 16; ++u) t += first[u];
	for (uint32_t base = p0 + 16 * kFinBlock; base < p1; base += 16 * kFinBlock) {
		double p[16];
#pragma unroll
		for (int u = 0; u < 16; ++u) {
			const uint32_t k = base + threadIdx.x + u * kFinBlock;
			p[u] = k < p1 ? partial[k] : 0.0;
		}
#pragma unroll
		for (int u = 0; u < 16; ++u) t += p[u];
	}
	for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();
	if (lane == 0) sh[wave] = t;
	__syncthreads();
	if (threadIdx.x == 0) {
		double tot = 0.0;
		for (int k = 0; k < kFinBlock / 64; ++k) tot += sh[k];
		total = tot;
	}
	__syncthreads();
	return total;
}

// Sum of partials [p0, p1) by a kFinBlock-thread workgroup in a fixed order
// (thread t adds p0 + t, p0 + t + 256, ..., sixteen loads in flight; then a
// wave tree and the four wave sums in order). Valid in every thread. An
// empty range sums to +0.
__device__ double wg_sum(const double* partial, uint32_t p0, uint32_t p1) {
	__shared__ double sh[kFinBlock / 64];
	__shared__ double total;
	double t = 0.0;
	for (uint32_t base = p0; base < p1; base += 16 * kFinBlock) {
		double p[16];
#pragma unroll
		for (int u = 0; u < 16; ++u) {
			const uint32_t k = base + threadIdx.x + u * kFinBlock;
			p[u] = k < p1 ? partial[k] : 0.0;
		}
#pragma unroll
		for (int u = 0; u < 16; ++u) t += p[u];
	}
	for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();
	if (lane == 0) sh[wave] = t;
	__syncthreads();
	if (threadIdx.x == 0) {
		double tot = 0.0;
		for (int k = 0; k < kFinBlock / 64; ++k) tot += sh[k];
		total = tot;
	}
	__syncthreads();
	return total;
}

// Gate of a finalize launch: the same condition under which its producer
// ran (a skipped producer leaves stale partials behind). Partials of the
// producer are per tile (kRangeTiles) or per row block (kRangeBlocks).
enum Gate : int { kGateNewton = 0, kGateIter = 1 };
enum RangeSel : int { kRangeTiles = 0, kRangeBlocks = 2 };

// A new solve: no iteration beyond the first is active until its Delta step
// says so (entries of an earlier solve must not gate this one).
__device__ __forceinline__ void clear_log(CGLog* log, uint32_t stride) {
	for (uint32_t j = 1 + threadIdx.x; j < stride; j += blockDim.x) log[j].active = 0u;
}

// One workgroup per instance. kGateIter: the producer belonged to PCG
// iteration `gk` and ran only if log[gk].active.
__global__ void __launch_bounds__(kFinBlock) k_finalize(Reduce rd, int sel, int kind, uint32_t k, int gate, uint32_t gk) {
	const uint32_t inst = blockIdx.x;
	SolverState* st = rd.st + inst;
	CGLog* log = rd.log + (size_t)inst * rd.logStride;
	if (!st->newtonActive) return;
	if ((kind == kFinAlpha || kind == kFinDelta) && persistent(rd.persist, inst)) return;
	if (gate == kGateIter && !log[gk].active) return;
	const uint32_t* R = rd.range + 6 * (size_t)inst;
	if (kind == kFinInit) {
		clear_log(log, rd.logStride);
		// the instance's tiles run again until the SpMV finds it done
		if (rd.tileGate)
			for (uint32_t t = R[0] + threadIdx.x; t < R[1]; t += blockDim.x) rd.tileGate[t] = 1u;
		if (rd.blkGate)
			for (uint32_t j = R[2] + threadIdx.x; j < R[3]; j += blockDim.x) rd.blkGate[j] = 1u;
	}
	__shared__ float chunk[4 * kFinBlock];
	if (st->seq && kind != kFinRefine0 && kind != kFinRefine) {
		const uint32_t r0 = R[4], r1 = R[5];
		float acc = 0.0f;
		for (uint32_t base = r0; base < r1; base += 4 * kFinBlock) {
			for (uint32_t j = threadIdx.x; j < 4 * kFinBlock && base + j < r1; j += kFinBlock) chunk[j] = rd.terms[base + j];
			__syncthreads();
			if (threadIdx.x == 0) {
				const uint32_t cnt = min((uint32_t)(4 * kFinBlock), r1 - base);
				for (uint32_t j = 0; j < cnt; ++j) acc += chunk[j];
			}
			__syncthreads();
		}
		if (threadIdx.x == 0) finalize(st, log, kind, k, (double)acc);
		return;
	}
	const double tot = wg_sum(rd.partial, R[sel], R[sel + 1]);
	if (threadIdx.x == 0) finalize(st, log, kind, k, tot);
}

// ---------------------------------------------------------------------------
// Set-up: tetrahedron rest data (tetrahedron_element_force.cpp:219-265)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_tet_setup(uint32_t m, const float* rest, const float* param, float4* rec,
                                                       uint32_t* degenerate) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m) return;
	V3 x[4];
	for (int k = 0; k < 4; ++k)
		x[k] = v3(rest[(size_t)(3 * k) * m + e], rest[(size_t)(3 * k + 1) * m + e], rest[(size_t)(3 * k + 2) * m + e]);
	float E = param[e], nu = param[(size_t)m + e];
	V3 e1 = x[1] - x[0], e2 = x[2] - x[0], e3 = x[3] - x[0];
	M33 Em = m33(e1, e2, e3);
	float dt = det(e1, e2, e3);
	if (dt == 0.0f) atomicOr(degenerate, 1u);
	float V = fabs_(dt) / 6.0f;
	dt = 1.0f / dt;
	M33 iE = dt * adjugate(Em);
	// the K0 blocks are recomputed from this record where they are used
	float r[kTetRecFloats];
	tet_record(iE, E, nu, V, r);
	for (int k = 0; k < 4; ++k) rec[4 * (size_t)e + k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

// ---------------------------------------------------------------------------
// Per-step particle kernels
// ---------------------------------------------------------------------------
struct StepArgs {
	uint32_t n;
	float h;
	const V3* gravity;          // per instance
	const uint32_t* rowInst;
	const uint32_t* rowBlk;     // row block boundaries
	const uint32_t* blkInst;
	const uint8_t* type;
	const float* mass;
	const float* damping;
	V3 *x, *v, *force, *translation, *x0, *v0, *fe, *y, *w, *px;
	SolverState* st;            // per instance
	const uint32_t* range;
	uint32_t* faultWord;        // fault bits of any instance
	uint32_t single;            // one instance: every row is instance 0
};

// force_solver.cpp:100-130 + sparse_force_solver.cpp:135-139. Every input is
// read (unconditionally, in bounds) before the first store, so the loads
// issue together.
__global__ void __launch_bounds__(kBlock) k_step_begin(StepArgs a) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n) return;
	// one instance: its gravity without the row's instance (no dependent load)
	auto run = [&](auto single) {
		const V3 x = a.x[i], v = a.v[i], force = a.force[i], translation = a.translation[i];
		const uint32_t type = a.type[i];
		const float mass = a.mass[i];
		const V3 g = a.gravity[decltype(single)::value ? 0u : a.rowInst[i]];
		// all of them in registers here: one wait, not a load sunk into the branch below
		B3G_SETTLE("v"(x.x), "v"(x.y), "v"(x.z), "v"(v.x), "v"(v.y), "v"(v.z), "v"(force.x), "v"(force.y), "v"(force.z),
		           "v"(translation.x), "v"(translation.y), "v"(translation.z), "v"(type), "v"(mass), "v"(g.x), "v"(g.y),
		           "v"(g.z));
		V3 fe = force;
		if (type == kDynamic) fe = fe + mass * g;
		a.x0[i] = x;
		a.v0[i] = v;
		a.fe[i] = fe;
		a.y[i] = translation;
		a.px[i] = zero3();
	};
	if (a.single) run(std::true_type{});
	else run(std::false_type{});
}

__global__ void __launch_bounds__(kBlock) k_state_begin(SolverState* sts, uint32_t count, uint32_t iters, uint32_t subIters,
                                                         float tol, float subTol, uint32_t seq, uint32_t refineOn,
                                                         uint32_t maxSweeps, float refineTol) {
	uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= count) return;
	SolverState* st = sts + k;
	st->seq = seq;
	st->refineOn = refineOn;
	st->maxSweeps = maxSweeps;
	st->refineTol2 = refineTol * refineTol;
	st->refineActive = 0;
	st->newtonIt = 0;
	st->newtonMax = iters;
	st->newtonActive = iters > 0 ? 1u : 0u;
	st->eps2 = tol * tol;
	st->eps2_sub = subTol * subTol;
	st->maxIt = subIters;
	st->it = 0;
}

// w = x0 - x + h v + y (sparse_force_solver.cpp:167)
__global__ void __launch_bounds__(kBlock) k_newton_prep(StepArgs a) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n) return;
	const V3 x0 = a.x0[i], x = a.x[i], v = a.v[i], y = a.y[i];   // in bounds whatever the gate says
	if (!a.st[a.single ? 0u : a.rowInst[i]].newtonActive) return;
	a.w[i] = ((x0 - x) + a.h * v) + y;
}

// ---------------------------------------------------------------------------
// Element kernels: stage this linearisation's contributions.
// ---------------------------------------------------------------------------
struct TetArgs {
	uint32_t m;
	const uint32_t* p;
	const float* rest;
	const float4* rec;       // b3g_tet.h record per element
	float4* q;
	const float* kd;
	const V3* x;
	const V3* v;
	V3* fstage;
	uint32_t f, fd;
	const SolverState* st;   // per instance
	const uint32_t* rowInst;
};

// tetrahedron_element_force.cpp:273-311 (b3ExtractRotation, 32 iterations)
__device__ Q4 extract_rotation(const M33& A, Q4 q) {
	for (int it = 0; it < 32; ++it) {
		M33 R = qmatrix(q);
		float den = fabs_(dot(R.x, A.x) + dot(R.y, A.y) + dot(R.z, A.z));
		if (den == 0.0f) break;
		V3 torque = cross(R.x, A.x) + cross(R.y, A.y) + cross(R.z, A.z);
		const float kTol = 1.0e-9f;
		float inv_den = 1.0f / den + kTol;
		V3 omega = inv_den * torque;
		float w = length(omega);
		if (w < kTol) break;
		V3 axis = omega / w;
		float theta = 0.5f * w;
		// the reference narrows the double-precision sin/cos (quat.h:83-88).
		// Below 2e-4 the narrowed values are exactly theta and 1 (theta^3 / 6
		// and theta^2 / 2 are under half a float ulp; every float there is
		// checked against the host's libm), so the double sincos is skipped
		float sf = theta, cf = 1.0f;
		if (!(theta < 2.0e-4f)) {
			double sd, cd;
			sincos((double)theta, &sd, &cd);
			sf = (float)sd;
			cf = (float)cd;
		}
		Q4 qw;
		qw.v = sf * axis;
		qw.s = cf;
		q = qmul(qw, q);
		qnormalize(q);
	}
	return q;
}

// tetrahedron_element_force.cpp:313-464. K0_ij comes from the element's
// record (b3g_tet.h). The element's Jacobian blocks -R K0_ij R^T are not
// staged: the slot assembly recomputes them from the record and the stored
// rotation (same operations, same values).
__global__ void __launch_bounds__(kBlock) k_tet_forces(TetArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	uint32_t id[4];
	// Only F and the warm start live through the polar iterations: the
	// positions, rest positions and the stiffness record are (re)loaded after
	// them (L2-hot), which keeps the loop's registers and the occupancy up.
	M33 F;
	{
		V3 p[4];
		for (int k = 0; k < 4; ++k) {
			id[k] = a.p[(size_t)k * m + e];
			p[k] = a.x[id[k]];
		}
		float rec[kTetRecFloats];
		for (int k = 0; k < 3; ++k) {
			const float4 r4 = a.rec[4 * (size_t)e + k];
			rec[4 * k] = r4.x; rec[4 * k + 1] = r4.y; rec[4 * k + 2] = r4.z; rec[4 * k + 3] = r4.w;
		}
		const M33 Em = m33(p[1] - p[0], p[2] - p[0], p[3] - p[0]);
		F = Em * tet_inv_e(rec);
	}
	const float4 q4 = a.q[e];
	Q4 q;
	q.v = v3(q4.x, q4.y, q4.z);
	q.s = q4.w;
	q = extract_rotation(F, q);
	a.q[e] = make_float4(q.v.x, q.v.y, q.v.z, q.s);
	V3 p[4], xr[4];
	for (int k = 0; k < 4; ++k) {
		p[k] = a.x[id[k]];
		xr[k] = v3(a.rest[(size_t)(3 * k) * m + e], a.rest[(size_t)(3 * k + 1) * m + e], a.rest[(size_t)(3 * k + 2) * m + e]);
	}
	float rec[kTetRecFloats];
	for (int k = 0; k < 4; ++k) {
		const float4 r4 = a.rec[4 * (size_t)e + k];
		rec[4 * k] = r4.x; rec[4 * k + 1] = r4.y; rec[4 * k + 2] = r4.z; rec[4 * k + 3] = r4.w;
	}
	M33 R = qmatrix(q);
	M33 RT = transpose(R);
	V3 us[4];
	for (int k = 0; k < 4; ++k) us[k] = RT * p[k] - xr[k];
	for (int i = 0; i < 4; ++i) {
		V3 fs = zero3();
		for (int j = 0; j < 4; ++j) fs = fs + tet_k0(rec, i, j) * us[j];   // m_K[i + 4 j]
		a.fstage[a.f + (size_t)i * m + e] = -(R * fs);      // f[vi] -= R fs
	}
	// Damping in a pass of its own (same sums per vertex), so that the
	// velocities and the displacements are never live together.
	const float kd = a.kd[e];
	if (kd > 0.0f) {
		// the record read again after the force stores: the blocks are
		// recomputed here, not kept live from the loop above
		for (int k = 0; k < 4; ++k) {
			const float4 r4 = a.rec[4 * (size_t)e + k];
			rec[4 * k] = r4.x; rec[4 * k + 1] = r4.y; rec[4 * k + 2] = r4.z; rec[4 * k + 3] = r4.w;
		}
		V3 vs[4];
		for (int k = 0; k < 4; ++k) vs[k] = a.v[id[k]];
		for (int i = 0; i < 4; ++i) {
			V3 fds = zero3();
			for (int j = 0; j < 4; ++j) {
				M33 K = R * tet_k0(rec, i, j) * RT;
				M33 kK = kd * K;
				fds = fds + kK * vs[j];
			}
			a.fstage[a.fd + (size_t)i * m + e] = -fds;
		}
	}
}

// ---- precision form: tetrahedron forces in FP64 ----------------------------
// Rest data in double (tetrahedron_element_force.cpp:219-265): 12 planes per
// tet: invE (9), volume, lambda, mu.
__global__ void __launch_bounds__(kBlock) k_tet_setup64(uint32_t m, const float* rest, const float* param, double* tet64) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m) return;
	D3 x[4];
	for (int k = 0; k < 4; ++k)
		x[k] = d3(rest[(size_t)(3 * k) * m + e], rest[(size_t)(3 * k + 1) * m + e], rest[(size_t)(3 * k + 2) * m + e]);
	D3 e1 = x[1] - x[0], e2 = x[2] - x[0], e3 = x[3] - x[0];
	double dt = det(e1, e2, e3);
	const double V = fabs(dt) / 6.0;
	DM33 iE = (1.0 / dt) * adjugate(dm33(e1, e2, e3));
	const double* pi = &iE.x.x;
	for (int c = 0; c < 9; ++c) tet64[(size_t)c * m + e] = pi[c];
	const double E = param[e], nu = param[(size_t)m + e];
	tet64[(size_t)9 * m + e] = V;
	tet64[(size_t)10 * m + e] = (nu * E) / ((1.0 + nu) * (1.0 - 2.0 * nu));   // b3ComputeD (:37-57)
	tet64[(size_t)11 * m + e] = E / (2.0 * (1.0 + nu));
}

// K0 u = V B^T D B u for the 4 vertex vectors u (K0 = V B^T D B, :121-159 and
// :247-261), evaluated through the strain: exact in real arithmetic, and in
// double free of K0's rounding. g[v] = shape-function gradients.
__device__ void k0_apply(const D3 g[4], double V, double lambda, double mu, const D3 u[4], D3 out[4]) {
	double e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0;
	for (int v = 0; v < 4; ++v) {
		e0 += g[v].x * u[v].x;
		e1 += g[v].y * u[v].y;
		e2 += g[v].z * u[v].z;
		e3 += g[v].y * u[v].x + g[v].x * u[v].y;
		e4 += g[v].z * u[v].y + g[v].y * u[v].z;
		e5 += g[v].z * u[v].x + g[v].x * u[v].z;
	}
	const double l2m = lambda + 2.0 * mu;
	const double s0 = l2m * e0 + lambda * e1 + lambda * e2;
	const double s1 = lambda * e0 + l2m * e1 + lambda * e2;
	const double s2 = lambda * e0 + lambda * e1 + l2m * e2;
	const double s3 = mu * e3, s4 = mu * e4, s5 = mu * e5;
	for (int v = 0; v < 4; ++v)
		out[v] = V * d3(g[v].x * s0 + g[v].y * s3 + g[v].z * s5, g[v].y * s1 + g[v].x * s3 + g[v].z * s4,
		                g[v].z * s2 + g[v].y * s4 + g[v].x * s5);
}

// b3ExtractRotation in double (tetrahedron_element_force.cpp:273-311).
__device__ DQ4 extract_rotation64(const DM33& A, DQ4 q) {
	for (int it = 0; it < 32; ++it) {
		DM33 R = qmatrix(q);
		double den = fabs(dot(R.x, A.x) + dot(R.y, A.y) + dot(R.z, A.z));
		if (den == 0.0) break;
		D3 torque = cross(R.x, A.x) + cross(R.y, A.y) + cross(R.z, A.z);
		const double kTol = 1.0e-9;
		double inv_den = 1.0 / den + kTol;
		D3 omega = inv_den * torque;
		double w = length(omega);
		if (w < kTol) break;
		D3 axis = omega / w;
		double sd, cd;
		sincos(0.5 * w, &sd, &cd);
		DQ4 qw;
		qw.v = sd * axis;
		qw.s = cd;
		q = qmul(qw, q);
		qnormalize(q);
	}
	return q;
}

// k_tet_forces in FP64: polar decomposition, u = R^T x - x_rest and the forces
// in double, written to the FP64 force arena (and, rounded, to the FP32 one).
// The rotation is kept (rounded) as the warm start; the Jacobian blocks the
// assembly recomputes from it stay FP32.
__global__ void __launch_bounds__(kBlock) k_tet_forces64(TetArgs a, const double* tet64, D3* fstage64) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	uint32_t id[4];
	// Only F and the warm start live through the polar iterations; the
	// positions, rest data and inverse rest shape are read again after them
	// (as k_tet_forces does), which keeps the loop's registers down.
	DM33 F;
	{
		D3 p[4];
		for (int k = 0; k < 4; ++k) {
			id[k] = a.p[(size_t)k * m + e];
			p[k] = d3(a.x[id[k]]);
		}
		DM33 iE;
		double* pi = &iE.x.x;
		for (int c = 0; c < 9; ++c) pi[c] = tet64[(size_t)c * m + e];
		F = dm33(p[1] - p[0], p[2] - p[0], p[3] - p[0]) * iE;
	}
	const float4 q4 = a.q[e];
	DQ4 q;
	q.v = d3(q4.x, q4.y, q4.z);
	q.s = q4.w;
	q = extract_rotation64(F, q);
	a.q[e] = make_float4((float)q.v.x, (float)q.v.y, (float)q.v.z, (float)q.s);
	D3 p[4], xr[4];
	for (int k = 0; k < 4; ++k) {
		p[k] = d3(a.x[id[k]]);
		xr[k] = d3(a.rest[(size_t)(3 * k) * m + e], a.rest[(size_t)(3 * k + 1) * m + e], a.rest[(size_t)(3 * k + 2) * m + e]);
	}
	DM33 iE;
	double* pi = &iE.x.x;
	for (int c = 0; c < 9; ++c) pi[c] = tet64[(size_t)c * m + e];
	const double V = tet64[(size_t)9 * m + e], lambda = tet64[(size_t)10 * m + e], mu = tet64[(size_t)11 * m + e];
	DM33 R = qmatrix(q);
	DM33 RT = transpose(R);
	// shape-function gradients (rows of invE; vertex 0 = minus their sum)
	D3 g[4];
	g[1] = d3(iE.x.x, iE.y.x, iE.z.x);
	g[2] = d3(iE.x.y, iE.y.y, iE.z.y);
	g[3] = d3(iE.x.z, iE.y.z, iE.z.z);
	g[0] = d3(-g[1].x - g[2].x - g[3].x, -g[1].y - g[2].y - g[3].y, -g[1].z - g[2].z - g[3].z);
	D3 us[4], fs[4];
	for (int k = 0; k < 4; ++k) us[k] = RT * p[k] - xr[k];
	k0_apply(g, V, lambda, mu, us, fs);
	const float kd = a.kd[e];
	D3 fds[4];
	if (kd > 0.0f) {   // f_d = -kd R K0 R^T v
		D3 uv[4];
		for (int k = 0; k < 4; ++k) uv[k] = RT * d3(a.v[id[k]]);
		k0_apply(g, V, lambda, mu, uv, fds);
	}
	for (int i = 0; i < 4; ++i) {
		const D3 f = -(R * fs[i]);
		fstage64[a.f + (size_t)i * m + e] = f;
		a.fstage[a.f + (size_t)i * m + e] = v3((float)f.x, (float)f.y, (float)f.z);
		if (kd > 0.0f) {
			const D3 fd = -((double)kd * (R * fds[i]));
			fstage64[a.fd + (size_t)i * m + e] = fd;
			a.fstage[a.fd + (size_t)i * m + e] = v3((float)fd.x, (float)fd.y, (float)fd.z);
		}
	}
}

struct TriArgs {
	uint32_t m;
	const uint32_t* p;
	const float* x2;
	const float* invS;
	const float* K0;
	const float* kd;
	const V3* x;
	const V3* v;
	M33* jstage;
	V3* fstage;
	uint32_t jx, jv, f, fd;
	const SolverState* st;   // per instance
	const uint32_t* rowInst;
};

__device__ M22 extract_rotation2(const M22& M) {   // triangle_element_force.cpp:131-150
	float m11 = M.x.x, m12 = M.y.x, m21 = M.x.y, m22 = M.y.y;
	float d = m11 * m22 - m12 * m21;
	M22 A;
	A.x.x = M.y.y; A.x.y = -M.y.x; A.y.x = -M.x.y; A.y.y = M.x.x;
	M22 Q = M + sign_(d) * A;
	normalize_div(Q.x);
	normalize_div(Q.y);
	return Q;
}

// triangle_element_force.cpp:285-468
__global__ void __launch_bounds__(kBlock) k_tri_forces(TriArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	uint32_t id[3];
	V3 p[3];
	for (int k = 0; k < 3; ++k) { id[k] = a.p[(size_t)k * m + e]; p[k] = a.x[id[k]]; }
	V3 n = cross(p[1] - p[0], p[2] - p[0]);
	normalize_inplace(n);
	V3 px, py;
	compute_basis(n, px, py);
	float P[6] = {px.x, py.x, px.y, py.y, px.z, py.z};
	float PT[6] = {px.x, px.y, px.z, py.x, py.y, py.z};
	V2 x[3];
	for (int k = 0; k < 3; ++k) gmul<2, 3, 1>(&x[k].x, P, &p[k].x);
	M22 T;
	T.x = x[1] - x[0];
	T.y = x[2] - x[0];
	M22 iS;
	iS.x = v2(a.invS[e], a.invS[(size_t)m + e]);
	iS.y = v2(a.invS[(size_t)2 * m + e], a.invS[(size_t)3 * m + e]);
	M22 R = extract_rotation2(T * iS);
	M22 RT = transpose(R);
	V2 us[3];
	for (int k = 0; k < 3; ++k) us[k] = RT * x[k] - v2(a.x2[(size_t)(2 * k) * m + e], a.x2[(size_t)(2 * k + 1) * m + e]);
	auto K0 = [&](int blk) {
		M22 k;
		k.x = v2(a.K0[(size_t)(4 * blk) * m + e], a.K0[(size_t)(4 * blk + 1) * m + e]);
		k.y = v2(a.K0[(size_t)(4 * blk + 2) * m + e], a.K0[(size_t)(4 * blk + 3) * m + e]);
		return k;
	};
	for (int i = 0; i < 3; ++i) {
		V2 fs = v2(0.0f, 0.0f);
		for (int j = 0; j < 3; ++j) fs = fs + K0(i + 3 * j) * us[j];
		fs = R * fs;
		V3 f3;
		gmul<3, 2, 1>(&f3.x, PT, &fs.x);
		a.fstage[a.f + (size_t)i * m + e] = -f3;   // f[vi] -= f
	}
	const float kd = a.kd[e];
	V3 vs[3];
	if (kd > 0.0f)
		for (int k = 0; k < 3; ++k) vs[k] = a.v[id[k]];
	V3 fds[3] = {zero3(), zero3(), zero3()};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			const int blk = i + 3 * j;
			M22 k = R * K0(blk) * RT;
			float PT_k[6];
			gmul<3, 2, 2>(PT_k, PT, &k.x.x);
			M33 Ke;
			gmul<3, 2, 3>(&Ke.x.x, PT_k, P);
			a.jstage[a.jx + (size_t)blk * m + e] = -Ke;
			if (kd > 0.0f) {
				M33 kK = kd * Ke;
				fds[i] = fds[i] + kK * vs[j];
				a.jstage[a.jv + (size_t)blk * m + e] = -kK;
			}
		}
	if (kd > 0.0f)
		for (int i = 0; i < 3; ++i) a.fstage[a.fd + (size_t)i * m + e] = -fds[i];
}

struct UVArgs {
	uint32_t m;
	const uint32_t* p;
	const float* c;
	const V3* x;
	const V3* v;
	float4* rec;             // per-element records the assembly rebuilds the blocks from
	V3* fstage;
	uint32_t f;
	const SolverState* st;   // per instance
	const uint32_t* rowInst;
};

__device__ __forceinline__ float coef(const UVArgs& a, int k, uint32_t e) { return a.c[(size_t)k * a.m + e]; }

// stretch_force.cpp:121-389: parts u (0) and v (1). Forces are staged; the
// Jacobian blocks of each part are rebuilt in the assembly (uv_block) from
// (n_w, |w|) written here and the part's constants.
__global__ void __launch_bounds__(kBlock) k_stretch_forces(UVArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	// every input in registers before any use (coefficients, positions,
	// velocities): the loads issue in two rounds (indices, then the rest)
	// instead of each one sinking to its first use behind a branch
	float cf[18];
#pragma unroll
	for (int k = 0; k < 18; ++k) cf[k] = coef(a, k, e);
	uint32_t id[3];
	V3 xs[3], vs[3];
	for (int k = 0; k < 3; ++k) { id[k] = a.p[(size_t)k * m + e]; xs[k] = a.x[id[k]]; vs[k] = a.v[id[k]]; }
	B3G_SETTLE("v"(cf[0]), "v"(cf[1]), "v"(cf[2]), "v"(cf[3]), "v"(cf[4]), "v"(cf[5]), "v"(cf[6]), "v"(cf[7]),
	           "v"(cf[8]), "v"(cf[9]), "v"(cf[10]), "v"(cf[11]), "v"(cf[12]), "v"(cf[13]), "v"(cf[14]), "v"(cf[15]),
	           "v"(cf[16]), "v"(cf[17]));
	B3G_SETTLE("v"(xs[0].x), "v"(xs[0].y), "v"(xs[0].z), "v"(xs[1].x), "v"(xs[1].y), "v"(xs[1].z), "v"(xs[2].x),
	           "v"(xs[2].y), "v"(xs[2].z), "v"(vs[0].x), "v"(vs[0].y), "v"(vs[0].z), "v"(vs[1].x), "v"(vs[1].y),
	           "v"(vs[1].z), "v"(vs[2].x), "v"(vs[2].y), "v"(vs[2].z));
	const float alpha = cf[0], du1 = cf[1], dv1 = cf[2], du2 = cf[3], dv2 = cf[4], inv_det = cf[5];
	const V3 dwudx = v3(cf[6], cf[7], cf[8]), dwvdx = v3(cf[9], cf[10], cf[11]);
	V3 dx1 = xs[1] - xs[0], dx2 = xs[2] - xs[0];
	V3 wu = inv_det * (dv2 * dx1 - dv1 * dx2);
	float len_wu = length(wu);
	V3 wv = inv_det * (-du2 * dx1 + du1 * dx2);
	float len_wv = length(wv);
	for (int part = 0; part < 2; ++part) {
		V3 wvec = part == 0 ? wu : wv;
		float len = part == 0 ? len_wu : len_wv;
		V3 dwdx = part == 0 ? dwudx : dwvdx;
		float ks = cf[12 + 3 * part], kd = cf[13 + 3 * part], bb = cf[14 + 3 * part];
		bool live = len > 0.0f;
		float inv_len = live ? 1.0f / len : 0.0f;
		V3 nw = inv_len * wvec;
		V3 dCdx[3];
		for (int i = 0; i < 3; ++i) dCdx[i] = alpha * comp(dwdx, i) * nw;
		if (ks > 0.0f) {
			float C = alpha * (len - bb);
			for (int i = 0; i < 3; ++i)
				a.fstage[a.f + (size_t)(6 * part + i) * m + e] = live ? -ks * C * dCdx[i] : zero3();
		}
		if (kd > 0.0f) {
			float dCdt = 0.0f;
			for (int i = 0; i < 3; ++i) dCdt += dot(dCdx[i], vs[i]);
			for (int i = 0; i < 3; ++i)
				a.fstage[a.f + (size_t)(6 * part + 3 + i) * m + e] = live ? -kd * dCdt * dCdx[i] : zero3();
		}
		a.rec[(size_t)(2 * e + part) * 3] = make_float4(nw.x, nw.y, nw.z, len);
	}
}

// shear_force.cpp:106-245
__global__ void __launch_bounds__(kBlock) k_shear_forces(UVArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	// every input in registers before any use (see k_stretch_forces)
	float cf[14];
#pragma unroll
	for (int k = 0; k < 14; ++k) cf[k] = coef(a, k, e);
	uint32_t id[3];
	V3 xs[3], vs[3];
	for (int k = 0; k < 3; ++k) { id[k] = a.p[(size_t)k * m + e]; xs[k] = a.x[id[k]]; vs[k] = a.v[id[k]]; }
	B3G_SETTLE("v"(cf[0]), "v"(cf[1]), "v"(cf[2]), "v"(cf[3]), "v"(cf[4]), "v"(cf[5]), "v"(cf[6]), "v"(cf[7]),
	           "v"(cf[8]), "v"(cf[9]), "v"(cf[10]), "v"(cf[11]), "v"(cf[12]), "v"(cf[13]));
	B3G_SETTLE("v"(xs[0].x), "v"(xs[0].y), "v"(xs[0].z), "v"(xs[1].x), "v"(xs[1].y), "v"(xs[1].z), "v"(xs[2].x),
	           "v"(xs[2].y), "v"(xs[2].z), "v"(vs[0].x), "v"(vs[0].y), "v"(vs[0].z), "v"(vs[1].x), "v"(vs[1].y),
	           "v"(vs[1].z), "v"(vs[2].x), "v"(vs[2].y), "v"(vs[2].z));
	const float alpha = cf[0], du1 = cf[1], dv1 = cf[2], du2 = cf[3], dv2 = cf[4], inv_det = cf[5];
	const V3 dwudx = v3(cf[6], cf[7], cf[8]), dwvdx = v3(cf[9], cf[10], cf[11]);
	const float ks = cf[12], kd = cf[13];
	V3 dx1 = xs[1] - xs[0], dx2 = xs[2] - xs[0];
	V3 wu = inv_det * (dv2 * dx1 - dv1 * dx2);
	V3 wv = inv_det * (-du2 * dx1 + du1 * dx2);
	V3 dCdx[3];
	for (int i = 0; i < 3; ++i) dCdx[i] = alpha * (comp(dwudx, i) * wv + comp(dwvdx, i) * wu);
	if (ks > 0.0f) {
		float C = alpha * dot(wu, wv);
		for (int i = 0; i < 3; ++i) a.fstage[a.f + (size_t)i * m + e] = -ks * C * dCdx[i];
	}
	if (kd > 0.0f) {
		float dCdt = 0.0f;
		for (int i = 0; i < 3; ++i) dCdt += dot(dCdx[i], vs[i]);
		for (int i = 0; i < 3; ++i) a.fstage[a.f + (size_t)(3 + i) * m + e] = -kd * dCdt * dCdx[i];
	}
	a.rec[4 * (size_t)e] = make_float4(wu.x, wu.y, wu.z, alpha);
	a.rec[4 * (size_t)e + 1] = make_float4(wv.x, wv.y, wv.z, ks);
}

// Action forces (e.g. spring_force.cpp:101-102, 128-129): per Newton
// iteration each vertex's accumulator takes the force's terms in the
// reference's order, stiffness then damping (stretch: u then v). A term
// whose coefficient is not positive is never added; a stiffness term the
// element switched off is staged as +0, and adding +0 to an accumulator that
// starts at +0 leaves it exactly as skipping the term would.
struct ActArgs {
	uint32_t m, np, nterm;
	uint32_t f;                  // staging base of the force type
	const V3* fstage;
	const float* c;              // coefficient planes
	uint32_t coef[4];            // per term: coefficient plane that must be positive
	uint32_t slot[4];            // per term: staging plane of vertex 0 (vertex i: slot + i)
	V3* act;
	uint32_t base;
	const uint32_t* p;           // particle planes (first vertex: instance)
	const SolverState* st;
	const uint32_t* rowInst;
};

__global__ void __launch_bounds__(kBlock) k_action_forces(ActArgs a) {
	const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	for (uint32_t i = 0; i < a.np; ++i) {
		V3 s = a.act[a.base + (size_t)i * m + e];
		for (uint32_t t = 0; t < a.nterm; ++t)
			if (a.c[(size_t)a.coef[t] * m + e] > 0.0f) s = s + a.fstage[a.f + (size_t)(a.slot[t] + i) * m + e];
		a.act[a.base + (size_t)i * m + e] = s;
	}
}

__global__ void __launch_bounds__(kBlock) k_clear_v3(size_t count, V3* p) {
	const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (i < count) p[i] = zero3();
}

struct PairArgs {
	uint32_t m;
	const uint32_t* p;
	const float* c;
	const V3* x;
	const V3* v;
	M33* jstage;
	V3* fstage;
	uint32_t jx, jv, f;
	const SolverState* st;   // per instance
	const uint32_t* rowInst;
};

// spring_force.cpp:61-143
__global__ void __launch_bounds__(kBlock) k_spring_forces(PairArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	uint32_t i1 = a.p[e], i2 = a.p[(size_t)m + e];
	float L0 = a.c[e], ks = a.c[(size_t)m + e], kd = a.c[(size_t)2 * m + e];
	V3 x1 = a.x[i1], v1 = a.v[i1], x2 = a.x[i2], v2_ = a.v[i2];
	M33 I = mdiag(1.0f);
	V3 dx = x1 - x2;
	float L = length(dx);
	bool live = L > 0.0f;
	V3 n = live ? dx / L : zero3();
	if (ks > 0.0f) {
		bool on = live && L > L0;
		V3 f1 = zero3(), f2 = zero3();
		M33 K11 = mzero(), K12 = mzero();
		if (on) {
			float C = L - L0;
			f1 = -ks * C * n;
			f2 = -f1;
			K11 = -ks * (outer(dx, dx) + (1.0f - L0 / L) * (I - outer(dx, dx)));
			K12 = -K11;
		}
		a.fstage[a.f + e] = f1;
		a.fstage[a.f + (size_t)m + e] = f2;
		a.jstage[a.jx + e] = K11;
		a.jstage[a.jx + (size_t)m + e] = K12;
		a.jstage[a.jx + (size_t)2 * m + e] = K12;
		a.jstage[a.jx + (size_t)3 * m + e] = K11;
	}
	if (kd > 0.0f) {
		V3 f1 = zero3(), f2 = zero3();
		M33 K11 = mzero(), K12 = mzero();
		if (live) {
			float dCdt = dot(n, v1 - v2_);
			f1 = -kd * dCdt * n;
			f2 = -f1;
			K11 = -kd * outer(n, n);
			K12 = -K11;
		}
		a.fstage[a.f + (size_t)2 * m + e] = f1;
		a.fstage[a.f + (size_t)3 * m + e] = f2;
		a.jstage[a.jv + e] = K11;
		a.jstage[a.jv + (size_t)m + e] = K12;
		a.jstage[a.jv + (size_t)2 * m + e] = K12;
		a.jstage[a.jv + (size_t)3 * m + e] = K11;
	}
}

// mouse_force.cpp:58-246
__global__ void __launch_bounds__(kBlock) k_mouse_forces(PairArgs a) {
	uint32_t e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= a.m) return;
	if (!element_gate(a.st, a.rowInst, a.p, e)) return;
	const uint32_t m = a.m;
	uint32_t id[4];
	V3 xs[4], vs[4];
	for (int k = 0; k < 4; ++k) { id[k] = a.p[(size_t)k * m + e]; xs[k] = a.x[id[k]]; vs[k] = a.v[id[k]]; }
	float w2 = a.c[e], w3 = a.c[(size_t)m + e], w4 = a.c[(size_t)2 * m + e];
	float L0 = a.c[(size_t)3 * m + e], ks = a.c[(size_t)4 * m + e], kd = a.c[(size_t)5 * m + e];
	M33 I = mdiag(1.0f);
	V3 c2 = w2 * xs[1] + w3 * xs[2] + w4 * xs[3];
	V3 d = xs[0] - c2;
	float L = length(d);
	bool live = L > 0.0f;
	float inv_L = live ? 1.0f / L : 0.0f;
	V3 n = inv_L * d;
	V3 dCdx[4] = {n, -w2 * n, -w3 * n, -w4 * n};
	if (ks > 0.0f) {
		bool on = live && L > L0;
		float C = L - L0;
		float L3 = L * L * L;
		float inv_L3 = L3 > 0.0f ? 1.0f / L3 : 0.0f;
		M33 A = inv_L * I - inv_L3 * outer(d, d);
		M33 B = inv_L3 * outer(d, d) - inv_L * I;
		const float wv[4] = {1.0f, w2, w3, w4};
		for (int i = 0; i < 4; ++i) a.fstage[a.f + (size_t)i * m + e] = on ? -ks * C * dCdx[i] : zero3();
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) {
				M33 d2;
				if (i == 0) d2 = j == 0 ? A : -wv[j] * A;
				else if (j == 0) d2 = wv[i] * B;
				else d2 = -wv[i] * wv[j] * B;
				if (i > 0 && j > 0 && i > j) d2 = -wv[j] * wv[i] * B;   // d2Cdx[i][j] = -w_j w_i B below the diagonal
				a.jstage[a.jx + (size_t)(4 * i + j) * m + e] = on ? -ks * (outer(dCdx[i], dCdx[j]) + C * d2) : mzero();
			}
	}
	if (kd > 0.0f) {
		V3 vc2 = w2 * vs[1] + w3 * vs[2] + w4 * vs[3];
		float dCdt = dot(vs[0] - vc2, n);
		for (int i = 0; i < 4; ++i) a.fstage[a.f + (size_t)(4 + i) * m + e] = live ? -kd * dCdt * dCdx[i] : zero3();
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j)
				a.jstage[a.jv + (size_t)(4 * i + j) * m + e] = live ? -kd * outer(dCdx[i], dCdx[j]) : mzero();
	}
}

// ---------------------------------------------------------------------------
// Contacts
// ---------------------------------------------------------------------------
struct ContactArgs {
	uint32_t nSpheres, nPairs;
	const WorldShape* worlds;
	ShapeData shapes;        // SDF voxels, mesh arrays
	const uint32_t* sphereParticle;
	const float* sphereRadius;
	const float* sphereFriction;
	const uint32_t* pairStart;    // per sphere + 1
	const uint32_t* pairSphere;   // per pair
	const uint32_t* pairWorld;    // per pair: world fixture (list order)
	const uint8_t* type;
	const float* invMass;
	V3* x;
	V3* v;
	uint32_t* active;
	uint32_t* stamp;
	uint32_t* applyFriction;
	float* normalForce;
	V3* manifold;   // 2 per pair: point, normal
	uint32_t* hit;
	V3* cF;
	M33* cK;
	uint32_t* stampCounter;       // device counter of FindNewContacts calls
	float dt;
	SolverState* st;              // per instance
	const uint32_t* rowInst;
};

// contact_manager.cpp:108-146
__global__ void __launch_bounds__(kBlock) k_update_contacts(ContactArgs a) {
	uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= a.nPairs) return;
	if (!a.active[k]) return;
	uint32_t s = a.pairSphere[k], w = a.pairWorld[k];
	uint32_t p = a.sphereParticle[s];
	bool keep = a.type[p] == kDynamic && overlap(aabb_cr(a.x[p], a.sphereRadius[s]), a.worlds[w].aabb);
	if (!keep) { a.active[k] = 0; return; }
	a.normalForce[k] = 0.0f;      // sphere_shape_contact.cpp:191-195
	a.applyFriction[k] = 0;
}

// contact_manager.cpp:65-82. Pairs are keyed (sphere, world); a stamp
// (FindNewContacts call, world position) orders a sphere's contact list.
__global__ void __launch_bounds__(kBlock) k_find_contacts(ContactArgs a) {
	uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= a.nPairs) return;
	if (a.active[k]) return;
	uint32_t s = a.pairSphere[k], w = a.pairWorld[k];
	uint32_t p = a.sphereParticle[s];
	if (!overlap(aabb_cr(a.x[p], a.sphereRadius[s]), a.worlds[w].aabb)) return;
	if (a.type[p] != kDynamic) return;
	const uint32_t p0 = a.pairStart[s], nw = a.pairStart[s + 1] - p0;
	a.active[k] = 1;
	a.stamp[k] = *a.stampCounter * nw + (k - p0);
	a.normalForce[k] = 0.0f;
	a.applyFriction[k] = 0;
	a.manifold[2 * k] = zero3();
	a.manifold[2 * k + 1] = zero3();
}

// One FindNewContacts call done.
__global__ void k_stamp_advance(uint32_t* counter) { *counter += 1u; }

// sphere_shape_contact.cpp:53-143
__global__ void __launch_bounds__(kBlock) k_contact_forces(ContactArgs a) {
	uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= a.nPairs) return;
	uint32_t s = a.pairSphere[k], w = a.pairWorld[k];
	uint32_t p = a.sphereParticle[s];
	if (!a.st[a.rowInst[p]].newtonActive) return;
	a.hit[k] = 0;
	if (!a.active[k]) return;
	const WorldShape& ws = a.worlds[w];
	V3 x1 = a.x[p];
	float r1 = a.sphereRadius[s], r2 = ws.radius;
	Manifold m2;
	if (!collide(ws, a.shapes, x1, r1, &m2)) return;
	if (!a.applyFriction[k]) {
		a.manifold[2 * k] = m2.point;
		a.manifold[2 * k + 1] = m2.normal;
		a.applyFriction[k] = 1;
	}
	V3 x2 = m2.point, n2 = m2.normal, n1 = -n2;
	V3 c1 = x1 + r1 * n1;
	V3 c2 = x2 + r2 * n2;
	float C = length(c2 - c1);
	C = fmin_(0.2f * C, 0.2f);                   // B3_BAUMGARTE, B3_MAX_CONTACT_LINEAR_CORRECTION
	V3 f1 = -1000.0f * C * n1;                   // B3_CONTACT_STIFFNESS
	M33 K11 = -1000.0f * (outer(n1, n1) + C * (mdiag(1.0f) - outer(n1, n1)));
	a.cF[k] = f1;
	a.cK[k] = K11;
	a.hit[k] = 1;
	a.normalForce[k] += length(f1);
}

// Contacts of one sphere in the reference's list order (newest first), as
// positions within the sphere's pairs.
// A sphere's contacts in list order are its active pairs by descending stamp
// (stamps of one sphere are distinct: the call number times the sphere's pair
// count plus the world position). next_contact returns the active pair with
// the largest stamp below `below`, or kNoContact: walking from kNoContact
// visits the list head to tail with no per-thread array, for any number of
// world fixtures.
constexpr uint32_t kNoContact = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t next_contact(const ContactArgs& a, uint32_t s, uint32_t below) {
	const uint32_t p0 = a.pairStart[s], p1 = a.pairStart[s + 1];
	uint32_t best = kNoContact, bestStamp = 0u;
	for (uint32_t k = p0; k < p1; ++k) {
		if (!a.active[k]) continue;
		const uint32_t st = a.stamp[k];
		if (st < below && (best == kNoContact || st > bestStamp)) { best = k; bestStamp = st; }
	}
	return best;
}
#define FOR_CONTACTS(a, s, k) \
	for (uint32_t k = next_contact((a), (s), kNoContact); k != kNoContact; k = next_contact((a), (s), (a).stamp[k]))

// sphere_shape_contact.cpp:145-189, contacts of a particle in list order
__global__ void __launch_bounds__(kBlock) k_friction(ContactArgs a) {
	uint32_t s = blockIdx.x * kBlock + threadIdx.x;
	if (s >= a.nSpheres) return;
	if (next_contact(a, s, kNoContact) == kNoContact) return;
	uint32_t p = a.sphereParticle[s];
	V3 v1 = a.v[p];
	float im1 = a.invMass[p];
	FOR_CONTACTS(a, s, k) {
		if (!a.applyFriction[k]) continue;
		V3 normal = a.manifold[2 * k + 1];
		float normalForce = a.normalForce[k];
		V3 t1 = perp(normal);
		V3 t2 = cross(t1, normal);
		float friction = sqrt_(a.sphereFriction[s] * a.worlds[a.pairWorld[k]].friction);   // contact.h:28-31
		float tangentMass = im1 > 0.0f ? 1.0f / im1 : 0.0f;
		V2 Cdot = v2(dot(v1, t1), dot(v1, t2));
		V2 impulse = tangentMass * -Cdot;
		float normalImpulse = a.dt * normalForce;
		float maxImpulse = friction * normalImpulse;
		if (dot(impulse, impulse) > maxImpulse * maxImpulse) {
			normalize_div(impulse);
			impulse.x *= maxImpulse;
			impulse.y *= maxImpulse;
		}
		V3 P = impulse.x * t1 + impulse.y * t2;
		v1 = v1 + im1 * P;
	}
	a.v[p] = v1;
}

// ---------------------------------------------------------------------------
// Assembly of pA, pb (sparse_force_solver.cpp:165-174)
//
// Both kernels work on tiles: the tile's contributions are loaded (or, for
// tetrahedra, computed) by all threads in parallel into LDS, then one thread
// per slot / row adds its own contributions in the reference's order. The
// sums are the sequential ones of the reference; only the loads run wide.
// ---------------------------------------------------------------------------
// Contributions per LDS chunk of the slot assembly (18 KB). Alone the kernel
// runs as fast with 1024 (36 KB); with two bodies stepping on two streams,
// 512 left room for the other stream's kernels and ran 14-24 % faster.
constexpr int kAsmCap = 512;

struct AsmArgs {
	uint32_t n, nnz, hasContacts;
	uint32_t stride;             // plane stride of A
	float h;
	const uint32_t* slotTiles;   // slot boundaries of the slot-assembly tiles
	const uint32_t* slotTileInst;
	const uint32_t* rowTiles;    // row boundaries of the row-assembly tiles (the SpMV tiles)
	const uint32_t* tileInst;
	const uint32_t* slotRow;
	const uint32_t* cols;
	const uint32_t* jxPtr;
	const uint32_t* jxIdx;
	const uint32_t* jvPtr;
	const uint32_t* jvIdx;
	const uint32_t* bPtr;
	const uint32_t* bSlot;
	const uint32_t* fPtr;
	const uint32_t* fIdx;
	const uint32_t* diagSlot;
	const M33* jstage;
	const V3* fstage;
	// tetrahedra: rest records (b3g_tet.h) and the element rotations
	const float* tetRec;
	uint32_t tetM;
	const float4* tetQ;
	const float* tetKd;
	// stretch / shear records (UVTable::rec)
	const float4* stRec;
	const float4* shRec;
	const float4* uvFirst;       // a stretch or shear record that exists (stRec or shRec), for loads that are discarded
	uint32_t anyJv;
	uint32_t single;             // one instance: every tile is instance 0
	const uint8_t* type;
	const float* mass;
	const float* damping;
	const int32_t* sphereOf;
	const V3* x0;
	const V3* v0;
	const V3* v;
	const V3* fe;
	const V3* w;
	const V3* y;
	float* A;
	V3* tB;
	V3* pb;
	V3* invM;
	V3* fi;
	ContactArgs c;
	const SolverState* st;       // per instance
	// precision form: FP64 right-hand side
	double* tB64;                // double3 per block, or null
	const D3* fstage64;
	uint32_t tetF0, tetF1, tetFd0, tetFd1;   // staging ranges of the FP64 tetrahedron forces
	double* pb64;
	const double* mass64;
	const V3* force;
	const V3* gravity;
	const uint32_t* rowInst;
	SolverState* stw;             // the same states, written: fault bits
	uint32_t* faultWord;          // fault bits of any instance
	const uint32_t* slotPos;      // storage position of each CSR slot (column-sorted storage), or null
};

// acc += the contributions [ptr[s], ptr[s + 1]) of slot s, in order, for the
// tile's slots [s0, s1). Each chunk of kAsmCap contributions is produced in
// three batched phases per thread (kAsmCap / kBlock contributions each):
// codes, then the dependent loads (a tetrahedron's two gradients, Lame
// parameters and rotation, or a record or staged block), then the products
// into LDS, so a thread has all its loads of a phase in flight at once.
constexpr int kAsmPer = kAsmCap / kBlock;

struct Contrib {
	const uint32_t* idx;    // codes
	const float* kd;        // per-tetrahedron scale (dfdv) or null (dfdx)
	bool dv;                // the dfdv list
};

// A stretch or shear Jacobian block from the element's record r (see
// UVTable), with the element kernel's operations in the reference's order.
__device__ __forceinline__ M33 uv_block(uint32_t variant, const float4* r, bool dv) {
	if (variant < kShearVariant) {   // stretch_force.cpp:163-207 (dfdx), :231-248 (dfdv)
		const int i = (int)(variant % 9) / 3, j = (int)(variant % 3);
		const float len = r[0].w;
		if (!(len > 0.0f)) return mzero();
		const V3 nw = v3(r[0].x, r[0].y, r[0].z);
		const V3 dwdx = v3(r[1].x, r[1].y, r[1].z);
		const float alpha = r[1].w;
		const V3 dCi = alpha * comp(dwdx, i) * nw;
		const V3 dCj = alpha * comp(dwdx, j) * nw;
		if (dv) return -r[2].y * outer(dCi, dCj);
		const float ks = r[2].x, bb = r[2].z;
		const float inv_len = 1.0f / len;
		const float C = alpha * (len - bb);
		M33 K = outer(dCi, dCj);
		if (len > bb) {
			const M33 d2 = (alpha * inv_len * comp(dwdx, i) * comp(dwdx, j)) * (mdiag(1.0f) - outer(nw, nw));
			K = K + C * d2;
		}
		return -ks * K;
	}
	// shear_force.cpp:161-244
	const int k = (int)(variant - kShearVariant), i = k / 3, j = k % 3;
	const V3 wu = v3(r[0].x, r[0].y, r[0].z), wv = v3(r[1].x, r[1].y, r[1].z);
	const V3 dwudx = v3(r[2].x, r[2].y, r[2].z), dwvdx = v3(r[3].x, r[3].y, r[3].z);
	const float alpha = r[0].w;
	const V3 dCi = alpha * (comp(dwudx, i) * wv + comp(dwvdx, i) * wu);
	const V3 dCj = alpha * (comp(dwudx, j) * wv + comp(dwvdx, j) * wu);
	return (dv ? -r[2].w : -r[1].w) * outer(dCi, dCj);
}

// Contribution kinds a body's slot lists hold (k_assemble_slots is instantiated
// per combination, so a body with one kind runs no test of the code's kind bits).
enum : int { kKindTet = 1, kKindUV = 2, kKindStaged = 4, kKindAll = 7 };

template <int Kinds>
__device__ __forceinline__ bool is_tet(uint32_t c) {
	return (Kinds & kKindTet) && (Kinds == kKindTet || (c & kTetCode));
}
template <int Kinds>
__device__ __forceinline__ bool is_uv(uint32_t c) {
	return (Kinds & kKindUV) && (Kinds == kKindUV || (Kinds == (kKindTet | kKindUV)) || (c & kUVCode));
}

template <int Kinds>
__device__ __forceinline__ void tile_accumulate(const AsmArgs& a, const Contrib& cb_, const uint32_t* ptr, uint32_t s0,
                                                uint32_t s1, uint32_t s, bool own, float* lds, M33& acc) {
	constexpr bool kTetOnly = Kinds == kKindTet;
	const uint32_t c0 = ptr[s0], c1 = ptr[s1];
	const uint32_t rb = ptr[own ? s : s0], re = ptr[own ? s + 1 : s0];   // used by owners only
	// A lane past the chunk takes the chunk's first contribution (in bounds)
	// and stores nothing: with no branch around them, the loads of all of a
	// thread's contributions issue together. The next chunk's codes are
	// requested with this chunk's records (the last chunk requests its own
	// again), so a chunk costs one memory round trip rather than two.
	uint32_t next[kAsmPer];
	auto load_codes = [&](uint32_t cb) {
		const uint32_t ce = min(cb + (uint32_t)kAsmCap, c1);
#pragma unroll
		for (int u = 0; u < kAsmPer; ++u) {
			const uint32_t k = cb + threadIdx.x + u * kBlock;
			next[u] = cb_.idx[k < ce ? k : cb];
		}
	};
	if (c0 < c1) load_codes(c0);
	for (uint32_t cb = c0; cb < c1; cb += kAsmCap) {
		const uint32_t ce = min(cb + (uint32_t)kAsmCap, c1);
		uint32_t code[kAsmPer];
#pragma unroll
		for (int u = 0; u < kAsmPer; ++u) code[u] = next[u];
		// per contribution: a tetrahedron's g_i | V, g_j | L2, (L, mu) in
		// r[0..2] and its rotation in r[3], a staged block in r[0..2], or a
		// stretch / shear record in r[0..3]
		float4 r[kAsmPer][4];
#pragma unroll
		for (int u = 0; u < kAsmPer; ++u) {
			const uint32_t c = code[u];
			if constexpr (Kinds == (kKindTet | kKindUV)) {
				// both kinds' loads for every contribution, the other kind's from its
				// first record (in bounds, cache-hot), and a select: no branch
				// around the loads when tetrahedra and cloth records mix in a tile
				const bool tet = (c & kTetCode) != 0;
				const uint32_t et = tet ? c & (kTetMaxElements - 1) : 0u;
				const uint32_t blk = tet ? (c & ~kTetCode) >> kTetBlockShift : 0u;
				const float* trec = a.tetRec + (size_t)kTetRecFloats * et;
				const V3 gi = tet_grad(trec, (int)(blk & 3u)), gj = tet_grad(trec, (int)(blk >> 2));
				const float4 s4 = *reinterpret_cast<const float4*>(trec + kTetRecV);
				const float kdv = cb_.kd ? cb_.kd[et] : 0.0f;
				const float4 q4 = a.tetQ[et];
				const uint32_t variant = (c & ~kUVCode) >> kUVShift, e = c & (kUVMaxElements - 1);
				const bool shear = variant >= kShearVariant;
				const float4* urec = tet ? a.uvFirst
				                         : shear ? a.shRec + 4 * (size_t)e : a.stRec + (size_t)(2 * e + variant / 9) * 3;
				const float4 u0 = urec[0], u1 = urec[1], u2 = urec[2], u3 = urec[shear && !tet ? 3 : 2];
				r[u][0] = tet ? make_float4(gi.x, gi.y, gi.z, s4.x) : u0;
				r[u][1] = tet ? make_float4(gj.x, gj.y, gj.z, s4.y) : u1;
				r[u][2] = tet ? make_float4(s4.z, s4.w, kdv, 0.0f) : u2;
				r[u][3] = tet ? q4 : u3;
			} else if (is_tet<Kinds>(c)) {
				const uint32_t e = c & (kTetMaxElements - 1), blk = (c & ~kTetCode) >> kTetBlockShift;
				const float* rec = a.tetRec + (size_t)kTetRecFloats * e;
				const V3 gi = tet_grad(rec, (int)(blk & 3u)), gj = tet_grad(rec, (int)(blk >> 2));
				const float4 s4 = *reinterpret_cast<const float4*>(rec + kTetRecV);
				r[u][0] = make_float4(gi.x, gi.y, gi.z, s4.x);
				r[u][1] = make_float4(gj.x, gj.y, gj.z, s4.y);
				r[u][2] = make_float4(s4.z, s4.w, cb_.kd ? cb_.kd[e] : 0.0f, 0.0f);   // z: the dfdv scale
				r[u][3] = a.tetQ[e];
			} else if (is_uv<Kinds>(c)) {
				// a stretch record is three float4, a shear record four: one
				// address per load, no branch (the stretch's fourth is unused)
				const uint32_t variant = (c & ~kUVCode) >> kUVShift, e = c & (kUVMaxElements - 1);
				const bool shear = variant >= kShearVariant;
				const float4* rec = shear ? a.shRec + 4 * (size_t)e : a.stRec + (size_t)(2 * e + variant / 9) * 3;
				r[u][0] = rec[0];
				r[u][1] = rec[1];
				r[u][2] = rec[2];
				r[u][3] = rec[shear ? 3 : 2];
			} else {
				const M33 m = a.jstage[c];
				r[u][0] = make_float4(m.x.x, m.x.y, m.x.z, m.y.x);
				r[u][1] = make_float4(m.y.y, m.y.z, m.z.x, m.z.y);
				r[u][2] = make_float4(m.z.z, 0.0f, 0.0f, 0.0f);
			}
		}
		load_codes(cb + kAsmCap < c1 ? cb + kAsmCap : cb);
#pragma unroll
		for (int u = 0; u < kAsmPer; ++u) {
			const uint32_t k = cb + threadIdx.x + u * kBlock;
			if (k >= ce) continue;
			const uint32_t c = code[u];
			M33 v;
			if (is_tet<Kinds>(c)) {
				// K0_ij (b3g_tet.h)
				v = tet_k0(v3(r[u][0].x, r[u][0].y, r[u][0].z), v3(r[u][1].x, r[u][1].y, r[u][1].z), r[u][0].w,
				           r[u][1].w, r[u][2].x, r[u][2].y);
			} else if (is_uv<Kinds>(c)) {
				v = uv_block((c & ~kUVCode) >> kUVShift, r[u], cb_.dv);
			} else {
				v.x = v3(r[u][0].x, r[u][0].y, r[u][0].z);
				v.y = v3(r[u][0].w, r[u][1].x, r[u][1].y);
				v.z = v3(r[u][1].z, r[u][1].w, r[u][2].x);
			}
			if (is_tet<Kinds>(c)) {
				// R K0_ij R^T of the tetrahedron (tetrahedron_element_force.cpp:400-420)
				Q4 q;
				q.v = v3(r[u][3].x, r[u][3].y, r[u][3].z);
				q.s = r[u][3].w;
				M33 R = qmatrix(q);
				M33 RT = transpose(R);
				M33 K = R * v * RT;
				// -K (or -(kd K)) goes in; a tetrahedron-only body stores K and
				// subtracts it in the sums below (a + (-b) is a - b, exactly)
				if (kTetOnly) v = cb_.kd ? r[u][2].z * K : K;
				else v = cb_.kd ? -(r[u][2].z * K) : -K;
			}
			const float* pv = &v.x.x;
#pragma unroll
			for (int c = 0; c < 9; ++c) lds[c * kAsmCap + (k - cb)] = pv[c];
		}
		__syncthreads();
		if (own) {
			// the slot's contributions in list order; two LDS reads ahead of their adds
			const uint32_t kb = max(rb, cb), ke = min(re, ce);
			auto at = [&](uint32_t k) {
				M33 v;
				float* pv = &v.x.x;
#pragma unroll
				for (int c = 0; c < 9; ++c) pv[c] = lds[c * kAsmCap + (k - cb)];
				return v;
			};
			uint32_t k = kb;
			for (; k + 2 <= ke; k += 2) {
				const M33 v0 = at(k), v1 = at(k + 1);
				acc = kTetOnly ? acc - v0 : acc + v0;
				acc = kTetOnly ? acc - v1 : acc + v1;
			}
			for (; k < ke; ++k) acc = kTetOnly ? acc - at(k) : acc + at(k);
		}
		__syncthreads();
	}
}

// Same over V3 contributions of rows (kTileCap per chunk): val[idx[k]]. The
// next chunk's indices are requested together with this chunk's values, so
// a chunk costs one memory round trip rather than two.
__device__ __forceinline__ void tile_accumulate3(const uint32_t* ptr, const uint32_t* idx, const V3* val, uint32_t r0,
                                                 uint32_t r1, uint32_t r, bool own, float* lds, V3& acc) {
	const uint32_t c0 = ptr[r0], c1 = ptr[r1];
	const uint32_t rb = ptr[own ? r : r0], re = ptr[own ? r + 1 : r0];   // used by owners only
	// a lane past the chunk reads the chunk's first entry (in bounds) and
	// stores nothing: no branch around the loads, which then issue together
	uint32_t id[kRowPer];
	auto load_idx = [&](uint32_t cb) {
		const uint32_t ce = min(cb + (uint32_t)kTileCap, c1);
#pragma unroll
		for (int u = 0; u < kRowPer; ++u) {
			const uint32_t k = cb + threadIdx.x + u * kBlock;
			id[u] = idx[k < ce ? k : cb];
		}
	};
	if (c0 < c1) load_idx(c0);
	for (uint32_t cb = c0; cb < c1; cb += kTileCap) {
		const uint32_t ce = min(cb + (uint32_t)kTileCap, c1);
		V3 v[kRowPer];
#pragma unroll
		for (int u = 0; u < kRowPer; ++u) v[u] = val[id[u]];
		if (cb + kTileCap < c1) load_idx(cb + kTileCap);
#pragma unroll
		for (int u = 0; u < kRowPer; ++u) {
			const uint32_t k = cb + threadIdx.x + u * kBlock;
			if (k >= ce) continue;
			lds[k - cb] = v[u].x;
			lds[kTileCap + (k - cb)] = v[u].y;
			lds[2 * kTileCap + (k - cb)] = v[u].z;
		}
		__syncthreads();
		if (own) {
			const uint32_t kb = max(rb, cb), ke = min(re, ce);
			for (uint32_t k = kb; k < ke; ++k)
				acc = acc + v3(lds[k - cb], lds[kTileCap + (k - cb)], lds[2 * kTileCap + (k - cb)]);
		}
		__syncthreads();
	}
}

template <int Kinds>
__global__ void __launch_bounds__(kBlock) k_assemble_slots(AsmArgs a) {
	const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
	__shared__ float lds[9 * kAsmCap];
	const uint32_t s0 = a.slotTiles[t], s1 = a.slotTiles[t + 1];
	// the tile's slots and its instance's gate in one round of loads
	const uint32_t active = !a.st ? 1u : a.single ? a.st[0].newtonActive : a.st[a.slotTileInst[t]].newtonActive;
	B3G_SETTLE("s"(s0), "s"(s1), "s"(active));
	if (!active) return;
	const uint32_t s = s0 + threadIdx.x;
	const bool own = s < s1;
	// per-slot data of the epilogue, loaded ahead of the contribution phase;
	// a thread past the tile reads slot s0's (in bounds, no branch around the
	// loads) and keeps none of it
	const uint32_t sc = own ? s : s0;
	const uint32_t i = a.slotRow[sc], j = a.cols[sc];
	const uint8_t type_i = a.type[i], type_j = a.type[j];
	const float mass_i = a.mass[i], damp_i = a.damping[i];
	const int32_t sph_i = a.hasContacts ? a.sphereOf[i] : -1;
	const V3 w_j = a.w[j];
	// dfdx(i, j): contributions in list order, then contacts (force_solver.cpp:70-86)
	M33 dx = mzero();
	tile_accumulate<Kinds>(a, Contrib{a.jxIdx, nullptr, false}, a.jxPtr, s0, s1, s, own, lds, dx);
	// dfdv(i, j): particle mass damping first, then contributions
	M33 dv = mzero();
	if (own && i == j && damp_i > 0.0f) dv = dv + mdiag(-damp_i * mass_i);
	if (a.anyJv) tile_accumulate<Kinds>(a, Contrib{a.jvIdx, a.tetKd, true}, a.jvPtr, s0, s1, s, own, lds, dv);
	if (!own) return;
	const bool dyn_i = type_i == kDynamic;
	if (i == j && sph_i >= 0 && dyn_i) {
		const uint32_t sph = (uint32_t)sph_i;
		FOR_CONTACTS(a.c, sph, k) {
			if (a.c.hit[k]) dx = dx + a.c.cK[k];
		}
	}
	// A = M - h dfdv - (h h) dfdx, block by block
	const float h = a.h, hh = h * h;
	M33 Mij = i == j ? mdiag(dyn_i ? mass_i : 1.0f) : mzero();
	M33 Aij = (Mij - h * dv) - hh * dx;
	// pA = S A S^T + I - S (S = I for dynamic particles, 0 otherwise)
	const bool dyn_j = type_j == kDynamic;
	M33 Si = dyn_i ? mdiag(1.0f) : mzero();
	M33 Sj = dyn_j ? mdiag(1.0f) : mzero();
	M33 P = (Si * Aij) * Sj;
	if (i == j) P = (P + mdiag(1.0f)) - Si;
	stA(a.A, a.stride, a.slotPos ? a.slotPos[s] : s, P);
	// (h dfdx_ij) w_j for b's product
	a.tB[s] = (h * dx) * w_j;
	if (a.tB64) {   // precision form: the same product in FP64 with w in FP64
		const D3 w64 = ((d3(a.x0[j]) - d3(a.c.x[j])) + (double)h * d3(a.v[j])) + d3(a.y[j]);
		const D3 t = ((double)h * dm33(dx)) * w64;
		a.tB64[3 * (size_t)s] = t.x;
		a.tB64[3 * (size_t)s + 1] = t.y;
		a.tB64[3 * (size_t)s + 2] = t.z;
	}
}

// Precision form: fi, b and pb in FP64 (sparse_force_solver.cpp:165-174),
// from the FP64 tetrahedron forces and the FP32 contributions of the other
// forces, in the reference's order per row.
__global__ void __launch_bounds__(kBlock) k_assemble_rows64(AsmArgs a) {
	const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
	if (a.st && !a.st[a.tileInst[t]].newtonActive) return;
	__shared__ double lds[3 * (kTileCap / 2)];
	constexpr uint32_t cap = kTileCap / 2;
	const uint32_t r0 = a.rowTiles[t], r1 = a.rowTiles[t + 1];
	const uint32_t i = r0 + threadIdx.x;
	const bool own = i < r1;
	const double mass = own ? a.mass64[i] : 0.0;
	const D3 vi = own ? d3(a.v[i]) : d3(0.0, 0.0, 0.0);
	D3 fi = d3(0.0, 0.0, 0.0);
	if (own && a.damping[i] > 0.0f) fi = fi + (-(double)a.damping[i] * mass) * vi;
	D3 hdw = d3(0.0, 0.0, 0.0);
	for (int pass = 0; pass < 2; ++pass) {
		const uint32_t* ptr = pass == 0 ? a.fPtr : a.bPtr;
		const uint32_t c0 = ptr[r0], c1 = ptr[r1];
		const uint32_t rb = own ? ptr[i] : 0, re = own ? ptr[i + 1] : 0;
		D3& acc = pass == 0 ? fi : hdw;
		for (uint32_t cb = c0; cb < c1; cb += cap) {
			const uint32_t ce = min(cb + cap, c1);
			for (uint32_t k = cb + threadIdx.x; k < ce; k += kBlock) {
				D3 v;
				if (pass == 0) {
					const uint32_t idx = a.fIdx[k];
					const bool tet = (idx >= a.tetF0 && idx < a.tetF1) || (idx >= a.tetFd0 && idx < a.tetFd1);
					v = tet ? a.fstage64[idx] : d3(a.fstage[idx]);
				} else {
					const double* tb = a.tB64 + 3 * (size_t)a.bSlot[k];
					v = d3(tb[0], tb[1], tb[2]);
				}
				lds[3 * (k - cb)] = v.x;
				lds[3 * (k - cb) + 1] = v.y;
				lds[3 * (k - cb) + 2] = v.z;
			}
			__syncthreads();
			if (own) {
				const uint32_t kb = max(rb, cb), ke = min(re, ce);
				for (uint32_t k = kb; k < ke; ++k) acc = acc + d3(lds[3 * (k - cb)], lds[3 * (k - cb) + 1], lds[3 * (k - cb) + 2]);
			}
			__syncthreads();
		}
	}
	if (!own) return;
	const bool dyn = a.type[i] == kDynamic;
	if (a.hasContacts && a.sphereOf[i] >= 0 && dyn) {
		const uint32_t sph = (uint32_t)a.sphereOf[i];
		FOR_CONTACTS(a.c, sph, k) {
			if (a.c.hit[k]) fi = fi + d3(a.c.cF[k]);
		}
	}
	const double h = a.h;
	const double mm = dyn ? mass : 1.0;
	// fe = f_ext + m g (force_solver.cpp:100-130) with the FP64 mass
	D3 fe = d3(a.force[i]);
	if (dyn) fe = fe + mass * d3(a.gravity[a.rowInst[i]]);
	D3 b = mm * (d3(a.v0[i]) - vi) + h * (fe + fi);
	b = b + hdw;
	if (!dyn) b = d3(0.0, 0.0, 0.0);   // pb = S b
	a.pb64[3 * (size_t)i] = b.x;
	a.pb64[3 * (size_t)i + 1] = b.y;
	a.pb64[3 * (size_t)i + 2] = b.z;
}

__global__ void __launch_bounds__(kBlock) k_assemble_rows(AsmArgs a) {
	const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
	__shared__ float lds[3 * kTileCap];
	const uint32_t r0 = a.rowTiles[t], r1 = a.rowTiles[t + 1];
	// the tile's rows and its instance's gate in one round of loads
	const uint32_t active = !a.st ? 1u : a.single ? a.st[0].newtonActive : a.st[a.tileInst[t]].newtonActive;
	B3G_SETTLE("s"(r0), "s"(r1), "s"(active));
	if (!active) return;
	const uint32_t i = r0 + threadIdx.x;
	const bool own = i < r1;
	const uint32_t ic = own ? i : r0;   // a thread past the tile reads row r0 (in bounds) and keeps nothing
	const float mass = a.mass[ic], damping = a.damping[ic];
	const V3 vi = a.v[ic];
	// fi: particle damping, forces in list order, contacts (force_solver.cpp:70-86)
	V3 fi = zero3();
	if (own && damping > 0.0f) fi = fi + (-damping * mass) * vi;
	tile_accumulate3(a.fPtr, a.fIdx, a.fstage, r0, r1, i, own, lds, fi);
	// h dfdx (x0 - x + h v + y) in the first-touch order of dfdx
	V3 hdw = zero3();
	tile_accumulate3(a.bPtr, a.bSlot, a.tB, r0, r1, i, own, lds, hdw);
	if (!own) return;
	const bool dyn = a.type[i] == kDynamic;
	if (a.hasContacts && a.sphereOf[i] >= 0 && dyn) {
		const uint32_t sph = (uint32_t)a.sphereOf[i];
		FOR_CONTACTS(a.c, sph, k) {
			if (a.c.hit[k]) fi = fi + a.c.cF[k];
		}
	}
	a.fi[i] = fi;
	// b = M (v0 - v) + h (fe + fi) + h dfdx (x0 - x + h v + y)
	const float h = a.h;
	M33 M = mdiag(dyn ? mass : 1.0f);
	V3 b = M * (a.v0[i] - vi) + h * (a.fe[i] + fi);
	b = b + hdw;
	// pb = S (b - A z), z = 0
	M33 S = dyn ? mdiag(1.0f) : mzero();
	a.pb[i] = S * b;
	// Jacobi preconditioner from pA(i, i) (sparse_solver.cpp:40-55); the
	// reference asserts each diagonal entry > 0, here the fault is recorded
	uint32_t ds = a.slotPos ? a.slotPos[a.diagSlot[i]] : a.diagSlot[i];
	float axx = a.A[ds], ayy = a.A[(size_t)4 * a.stride + ds], azz = a.A[(size_t)8 * a.stride + ds];
	a.invM[i] = v3(1.0f / axx, 1.0f / ayy, 1.0f / azz);
	if (!(axx > 0.0f) || !(ayy > 0.0f) || !(azz > 0.0f)) {
		atomicOr(&a.stw[a.rowInst[i]].faults, (uint32_t)B3G_FAULT_DIAGONAL);
		atomicOr(a.faultWord, (uint32_t)B3G_FAULT_DIAGONAL);
	}
}

// ---------------------------------------------------------------------------
// PCG (sparse_solver.cpp:28-109)
// ---------------------------------------------------------------------------
struct CGArgs {
	uint32_t n, nnz, nTiles, nRowBlk;
	uint32_t stride;            // plane stride of A and length of cols (nnz rounded up to 4)
	const uint32_t* tiles;
	const uint32_t* tileInst;
	const uint32_t* tileBlk;    // block boundaries of the tiles (rowptr at tiles)
	const uint32_t* rowBlk;
	const uint32_t* blkInst;
	uint32_t single;            // one instance: every tile and row block is instance 0
	uint32_t* tileGate;         // batches: 0 once the tile's instance is done with the current solve, or null
	uint32_t* blkGate;          // the same per row block
	const uint32_t* rowptr;
	const uint32_t* cols;
	const float* A;
	const V3* pb;
	const V3* invM;
	V3* px;
	V3* r;
	V3* s;
	V3* q;
	const V3* dold;
	V3* dnew;
	double* partial;            // residual / vector-update partials
	double* partialQ;           // d.q partials of the SpMV kernel
	float* terms;
	SolverState* st;            // per instance
	const uint32_t* range;
	CGLog* log;
	uint32_t logStride;
	uint32_t fused;             // FP64 mode: alpha / delta formed in the consuming kernel
	// BF16 operator (precision form): off-diagonal blocks in BF16, diagonal blocks FP32 per row
	uint32_t bf16;
	const uint32_t* AbP;        // nine pair planes (Body::AbP)
	const uint32_t* pairBase;   // first pair of each tile
	uint32_t pairStride;
	const float* diagF;
	// column-sorted storage (Body::sortedA), or null: CSR storage with cols
	const uint32_t* code;
	const uint32_t* tileBase;
	// per instance, 1 where the persistent kernel runs the PCG iterations (the
	// launch sequence then skips the instance after the solve's start), or null
	const uint32_t* persist;
};


__device__ __forceinline__ V3 diag_apply(V3 im, V3 r) { return mdiag3(im.x, im.y, im.z) * r; }

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }


// One SpMV tile into acc (products in LDS, then row sums in CSR order).
// before() runs first, uniformly over the workgroup; false from it ends the
// tile (the function returns false).
// Pairs: the BF16 operator, whose pair planes give each thread the blocks
// 2l and 2l + 1 of the chunk (l = thread, thread + 256) from nine words.
// A tile's rows [r0, r1), blocks [b0, b1), first column (column-sorted
// storage) and first BF16 pair: independent loads, issued together.
struct TileRange {
	uint32_t r0, r1, b0, b1, base, pair;
};
template <bool Sorted, bool Pairs>
__device__ __forceinline__ TileRange tile_range_(const CGArgs& a, uint32_t t) {
	TileRange tr;
	tr.r0 = a.tiles[t];
	tr.r1 = a.tiles[t + 1];
	tr.b0 = a.tileBlk[t];
	tr.b1 = a.tileBlk[t + 1];
	tr.base = Sorted ? a.tileBase[t] : 0u;
	tr.pair = Pairs ? a.pairBase[t] : 0u;
	return tr;
}
template <int St>
__device__ __forceinline__ TileRange tile_range(const CGArgs& a, uint32_t t) {
	return tile_range_<(St & 1) != 0, (St & 2) != 0>(a, t);
}

template <bool Sorted, bool Pairs, class Gather, class Before>
__device__ __forceinline__ bool tile_rows_(const CGArgs& a, const TileRange& tr, uint32_t& row, Gather gather,
                                           float* lds, Before before, V3& acc) {
	const uint32_t r0 = tr.r0, r1 = tr.r1, b0 = tr.b0, b1 = tr.b1, base = tr.base;
	row = r0 + threadIdx.x;
	const bool own = row < r1;
	uint32_t rb = own ? a.rowptr[row] : 0, re = own ? a.rowptr[row + 1] : 0;
	auto put = [&](uint32_t i, const V3& p) {
		lds[3 * i] = p.x;
		lds[3 * i + 1] = p.y;
		lds[3 * i + 2] = p.z;
	};
	// The thread's columns first, so each block's gathers issue together with
	// its matrix loads; a lane past the chunk loads the chunk's first block
	// (in bounds) and stores nothing, which keeps the loads free of branches.
	// With the column-sorted storage the block at storage position k has its
	// column and its CSR position (where its product goes) in code[k].
	auto local = [](int u) -> uint32_t {
		return Pairs ? 2 * (threadIdx.x + (u >> 1) * kBlock) + (u & 1) : threadIdx.x + u * kBlock;
	};
	uint32_t raw[kRowPer];
	auto load_codes = [&](uint32_t cb, uint32_t ce) {
#pragma unroll
		for (int u = 0; u < kRowPer; ++u) {
			const uint32_t k = cb + local(u), kk = k < ce ? k : cb;
			raw[u] = Sorted ? a.code[kk] : a.cols[kk];
		}
	};
	// the first chunk's codes are requested before the prologue, whose loads
	// then share their round trip
	if (b0 < b1) load_codes(b0, min(b0 + (uint32_t)kTileCap, b1));
	acc = zero3();
	if (!before()) return false;
	for (uint32_t cb = b0; cb < b1; cb += kTileCap) {
		const uint32_t ce = min(cb + (uint32_t)kTileCap, b1);
		if (cb != b0) load_codes(cb, ce);
		uint32_t col[kRowPer], dst[kRowPer];
#pragma unroll
		for (int u = 0; u < kRowPer; ++u) {
			const uint32_t k = cb + local(u), kk = k < ce ? k : cb;
			if (Sorted) {
				col[u] = base + (raw[u] >> 10);
				dst[u] = raw[u] & 1023u;
			} else {
				col[u] = raw[u];
				dst[u] = kk - cb;
			}
		}
		// two blocks' loads in flight per thread at a time
		auto products = [&](auto ld) {
#pragma unroll
			for (int u0 = 0; u0 < kRowPer; u0 += 2) {
				M33 m[2];
				V3 g[2];
#pragma unroll
				for (int w = 0; w < 2; ++w) {
					const uint32_t k = cb + threadIdx.x + (u0 + w) * kBlock;
					g[w] = gather(col[u0 + w]);
					m[w] = ld(k < ce ? k : cb);
				}
				B3G_SETTLE("v"(g[0].x), "v"(g[0].y), "v"(g[0].z), "v"(g[1].x), "v"(g[1].y), "v"(g[1].z), "v"(m[0].x.x),
				           "v"(m[0].x.y), "v"(m[0].x.z), "v"(m[0].y.x), "v"(m[0].y.y), "v"(m[0].y.z), "v"(m[0].z.x),
				           "v"(m[0].z.y), "v"(m[0].z.z), "v"(m[1].x.x), "v"(m[1].x.y), "v"(m[1].x.z), "v"(m[1].y.x),
				           "v"(m[1].y.y), "v"(m[1].y.z), "v"(m[1].z.x), "v"(m[1].z.y), "v"(m[1].z.z));
#pragma unroll
				for (int w = 0; w < 2; ++w) {
					const uint32_t k = cb + threadIdx.x + (u0 + w) * kBlock;
					const V3 p = m[w] * g[w];
					if (k < ce) put(dst[u0 + w], p);
				}
			}
		};
		if (Pairs) {
			const uint32_t p0 = tr.pair + ((cb - b0) >> 1);   // cb - b0 is a multiple of kTileCap
#pragma unroll
			for (int j = 0; j < kRowPer / 2; ++j) {
				const uint32_t l = threadIdx.x + j * kBlock, k = cb + 2 * l;
				const bool in0 = k < ce, in1 = k + 1 < ce;
				const uint32_t pp = p0 + (in0 ? l : 0u);
				uint32_t w[9];
#pragma unroll
				for (int c = 0; c < 9; ++c) w[c] = a.AbP[(size_t)c * a.pairStride + pp];
				M33 m0, m1;
				float* q0 = &m0.x.x;
				float* q1 = &m1.x.x;
#pragma unroll
				for (int c = 0; c < 9; ++c) {
					q0[c] = bf16_lo(w[c]);
					q1[c] = bf16_hi(w[c]);
				}
				const V3 g0 = gather(col[2 * j]), g1 = gather(col[2 * j + 1]);
				const V3 e0 = m0 * g0, e1 = m1 * g1;
				if (in0) put(dst[2 * j], e0);
				if (in1) put(dst[2 * j + 1], e1);
			}
		} else {
			products([&](uint32_t k) { return ldA(a.A, a.stride, k); });
		}
		__syncthreads();
		if (own) {
			// the row's products in CSR order; four LDS reads ahead of their adds
			const uint32_t kb = max(rb, cb), ke = min(re, ce);
			auto at = [&](uint32_t k) { return v3(lds[3 * (k - cb)], lds[3 * (k - cb) + 1], lds[3 * (k - cb) + 2]); };
			uint32_t k = kb;
			for (; k + 4 <= ke; k += 4) {
				const V3 v0 = at(k), v1 = at(k + 1), v2 = at(k + 2), v3_ = at(k + 3);
				acc = acc + v0;
				acc = acc + v1;
				acc = acc + v2;
				acc = acc + v3_;
			}
			for (; k < ke; ++k) acc = acc + at(k);
		}
		__syncthreads();
	}
	if (a.bf16 && own) acc = acc + ldM(a.diagF, a.n, 0, row) * gather(row);   // FP32 diagonal block (zero in the BF16 planes)
	return true;
}

// The storage of pA (St: bit 0 column-sorted chunks, bit 1 BF16 pair planes)
// is a property of the body, uniform over a launch, and each kernel that
// reads pA is instantiated per storage (tile_storage picks one at launch):
// one kernel with every storage inside ran the FP32 step 3.5 % slower.
template <int St, class Gather, class Before>
__device__ __forceinline__ bool tile_rows(const CGArgs& a, const TileRange& tr, uint32_t& row, Gather gather,
                                          float* lds, Before before, V3& acc) {
	return tile_rows_<(St & 1) != 0, (St & 2) != 0>(a, tr, row, gather, lds, before, acc);
}

__device__ __forceinline__ bool no_prologue() { return true; }

// r = b - A x ; s = invM r ; delta = r.s   (mode kFinInit: CG start, kFinDelta: true-residual
// refresh after iteration k, which ran only if log[k].active)
template <bool Multi, int St>
__global__ void __launch_bounds__(kBlock) k_cg_residual(CGArgs a, int mode, uint32_t k) {
	__shared__ float lds[3 * kTileCap];
	// One body: the tile's bounds and the instance's gate are read together,
	// before anything depends on them (every address is valid whether or not
	// the tile runs). A batch reads the gate first, so a tile of a converged
	// instance costs one round of three scalar loads.
	auto tile = [&](uint32_t t, uint32_t inst, auto single) {
		constexpr bool S = decltype(single)::value;
		const SolverState& st = a.st[inst];
		const uint32_t active = st.newtonActive;
		const uint32_t logActive = a.log[(size_t)inst * a.logStride + k].active;
		const uint32_t persist = a.persist ? a.persist[inst] : 0u;
		const bool run = active && (mode != kFinDelta || (logActive && !persist));
		if constexpr (!S) {
			B3G_SETTLE("s"(active), "s"(logActive), "s"(persist));
			if (!run) return;
		}
		const TileRange tr = tile_range<St>(a, t);
		const uint32_t seq = st.seq;
		if constexpr (S) {
			B3G_SETTLE("s"(tr.r0), "s"(tr.r1), "s"(tr.b0), "s"(tr.b1), "s"(tr.base), "s"(tr.pair), "s"(active),
			           "s"(seq), "s"(logActive), "s"(persist));
			if (!run) return;
		} else {
			B3G_SETTLE("s"(tr.r0), "s"(tr.r1), "s"(tr.b0), "s"(tr.b1), "s"(tr.base), "s"(tr.pair), "s"(seq));
		}
		uint32_t row;
		V3 ax;
		tile_rows<St>(a, tr, row, [&](uint32_t c) { return a.px[c]; }, lds, no_prologue, ax);
		float part = 0.0f;
		const bool has = row < tr.r1;
		if (has) {
			V3 r = a.pb[row] - ax;
			V3 s = diag_apply(a.invM[row], r);
			a.r[row] = r;
			a.s[row] = s;
			part = dot(r, s);
		}
		block_partial(part, has, row, a.partial, a.terms, seq != 0, t);
	};
	if (!Multi) {   // instance 0 as a constant: its reads need no load of the tile's instance first
		const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
		if (a.single) tile(t, 0u, std::true_type{});
		else tile(t, a.tileInst[t], std::false_type{});
		return;
	}
	auto ok = [&](uint32_t inst) {
		return a.st[inst].newtonActive &&
		       (mode != kFinDelta || (a.log[(size_t)inst * a.logStride + k].active && !persistent(a.persist, inst)));
	};
	for_items<Multi, true>(a.nTiles, a.tileInst, ok, [&](uint32_t t) { tile(t, a.tileInst[t], std::false_type{}); });
}

// PCG iteration k: q = A d with d = s + beta_k d_old formed on the fly
// (d = s at k = 0), d written for the tile's rows; partials of d.q.
// Fused mode: the workgroup first forms delta_k / beta_k (Delta(k)) from the
// partials of the reduction that ended iteration k - 1 (row blocks, or tiles
// after a residual refresh: sel); the instance's first tile records them.
// force = 1 runs it whatever the PCG state (timing pass on PCG scratch).
template <bool Multi, int St>
__global__ void __launch_bounds__(kBlock) k_cg_spmv_dq(CGArgs a, uint32_t k, int sel, int force) {
	__shared__ float lds[3 * kTileCap];
	// One body: the tile's bounds, the instance's gate and the scalars of the
	// fused prologue are read together, before anything depends on them
	// (every address is valid whether or not the tile runs), one round trip
	// ahead of the tile's column codes instead of a chain of dependent ones.
	// A batch reads the gate first, so a tile of a converged instance costs
	// one round of four scalar loads; the rest is read after it.
	auto tile = [&](uint32_t t, uint32_t inst, auto single) {
		constexpr bool S = decltype(single)::value;
		SolverState& st = a.st[inst];
		CGLog* log = a.log + (size_t)inst * a.logStride;
		const uint32_t km = k > 0 ? k - 1 : 0u;
		const uint32_t active = st.newtonActive, prevActive = log[km].active, curActive = log[k].active;
		const uint32_t persist = a.persist ? a.persist[inst] : 0u;
		// log[k].active is this launch's own output in fused mode, and read only otherwise
		const bool run = force || (active && !persist && prevActive && (k == 0 || a.fused || curActive));
		// a batch's tile of an instance that is done with this solve (converged,
		// out of the Newton loop or persistent) says so for the later launches
		auto done = [&]() {
			if (!S && a.tileGate && !force && threadIdx.x == 0) a.tileGate[t] = 0u;
		};
		if constexpr (!S) {
			B3G_SETTLE("s"(active), "s"(prevActive), "s"(curActive), "s"(persist));
			if (!run) {
				done();
				return;
			}
		}
		const TileRange tr = tile_range<St>(a, t);
		const uint32_t* R = a.range + 6 * (size_t)inst;
		const uint32_t seq = st.seq, maxIt = st.maxIt;
		const float eps2 = st.eps2_sub, delta0 = st.delta0;
		const float prevDelta = log[km].delta, logBeta = log[k].beta;
		const uint32_t p0 = R[sel], p1 = R[sel + 1], firstTile = R[0];
		if constexpr (S) {
			B3G_SETTLE("s"(tr.r0), "s"(tr.r1), "s"(tr.b0), "s"(tr.b1), "s"(tr.base), "s"(tr.pair), "s"(active),
			           "s"(seq), "s"(maxIt), "s"(eps2), "s"(delta0), "s"(prevActive), "s"(curActive), "s"(prevDelta),
			           "s"(logBeta), "s"(p0), "s"(p1), "s"(firstTile), "s"(persist));
			if (!run) return;
		} else {
			B3G_SETTLE("s"(tr.r0), "s"(tr.r1), "s"(tr.b0), "s"(tr.b1), "s"(tr.base), "s"(tr.pair), "s"(seq),
			           "s"(maxIt), "s"(eps2), "s"(delta0), "s"(prevDelta), "s"(logBeta), "s"(p0), "s"(p1),
			           "s"(firstTile));
		}
		float beta = 0.0f;
		if (k > 0 && !a.fused) beta = logBeta;
		// fused mode: delta_k / beta_k from the partials
		auto prologue = [&]() -> bool {
			if (k == 0 || !a.fused) return true;
			const float v = (float)wg_sum(a.partial, p0, p1);
			beta = v / prevDelta;
			const bool act = k != maxIt && !(v <= eps2 * delta0);
			if (!force && t == firstTile && threadIdx.x == 0) {
				log[k].delta = v;
				log[k].beta = beta;
				log[k].active = act ? 1u : 0u;
				st.it = k;
			}
			if (!act) done();
			return force || act;
		};
		// d = s + beta d_old, and d = s at k = 0: there dsrc is s and beta is +0,
		// and s + (+0) s = s exactly for finite s (the zero product carries the
		// sign of s), so every iteration gathers the same way, without a branch
		const V3* dsrc = k == 0 ? a.s : a.dold;
		auto dval = [&](uint32_t c) { return a.s[c] + beta * dsrc[c]; };
		uint32_t row;
		V3 q;
		if (!tile_rows<St>(a, tr, row, dval, lds, prologue, q)) return;
		float part = 0.0f;
		const bool has = row < tr.r1;
		if (has) {
			V3 d = dval(row);
			a.dnew[row] = d;
			a.q[row] = q;
			part = dot(d, q);
		}
		block_partial(part, has, row, a.partialQ, a.terms, seq != 0, t);
	};
	if (!Multi) {   // instance 0 as a constant: its reads need no load of the tile's instance first
		const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
		if (a.single) {
			tile(t, 0u, std::true_type{});
		} else {
			// the tile's gate word with its instance: a tile of an instance that
			// is done costs one round of two loads
			const uint32_t inst = a.tileInst[t], g = a.tileGate && !force ? a.tileGate[t] : 1u;
			B3G_SETTLE("s"(inst), "s"(g));
			if (!g) return;
			tile(t, inst, std::false_type{});
		}
		return;
	}
	auto ok = [&](uint32_t inst) -> bool {
		if (force) return true;
		const CGLog* log = a.log + (size_t)inst * a.logStride;
		if (!a.st[inst].newtonActive || persistent(a.persist, inst)) return false;
		if (k > 0) return log[k - 1].active && (a.fused || log[k].active);
		return log[0].active != 0;
	};
	for_items<Multi, true>(a.nTiles, a.tileInst, ok, [&](uint32_t t) { tile(t, a.tileInst[t], std::false_type{}); });
}

// alpha_k of a vector kernel of iteration k: fused mode sums the SpMV's d.q
// partials of the instance (Alpha(k)); the instance's first block records it.
__device__ __forceinline__ float cg_alpha(const CGArgs& a, uint32_t inst, uint32_t blk, uint32_t k) {
	CGLog* log = a.log + (size_t)inst * a.logStride;
	if (!a.fused) return log[k].alpha;
	const uint32_t* R = a.range + 6 * (size_t)inst;
	const float alpha = log[k].delta / (float)wg_sum(a.partialQ, R[0], R[1]);
	if (blk == R[2] && threadIdx.x == 0) log[k].alpha = alpha;
	return alpha;
}

// x += alpha d ; r -= alpha q ; s = invM r ; partials of r.s (iteration k).
// RPT rows per thread: row blocks of at most RPT x 256 rows (a thread adds its
// rows r0 + t, r0 + t + 256, ... in that order, whatever RPT, so the partials
// depend only on the block size); fewer rows per thread hold fewer registers.
template <bool Multi, int RPT>
__global__ void __launch_bounds__(kBlock) k_cg_update(CGArgs a, uint32_t k, int force) {
	// One body: the block's bounds, the instance's gate and the scalars of
	// alpha are read together before anything depends on them (every address
	// is valid whether or not the block runs), then alpha's partials and the
	// rows, all before the gate (in bounds whatever it says). A batch reads
	// the gate first, so a block of a converged instance costs one round of
	// three scalar loads and no row traffic.
	auto block = [&](uint32_t blk, uint32_t inst, auto single) {
		constexpr bool S = decltype(single)::value;
		const SolverState& st = a.st[inst];
		CGLog* log = a.log + (size_t)inst * a.logStride;
		const uint32_t active = st.newtonActive, logActive = log[k].active;
		const uint32_t persist = a.persist ? a.persist[inst] : 0u;
		const bool run = force || (active && logActive && !persist);
		if constexpr (!S) {
			B3G_SETTLE("s"(active), "s"(logActive), "s"(persist));
			if (!run) {
				// done with this solve (log[k] is final here): later launches skip the block
				if (a.blkGate && threadIdx.x == 0) a.blkGate[blk] = 0u;
				return;
			}
		}
		const uint32_t r0 = a.rowBlk[blk], r1 = a.rowBlk[blk + 1];
		const uint32_t* R = a.range + 6 * (size_t)inst;
		const uint32_t seq = st.seq;
		const float delta = log[k].delta, logAlpha = log[k].alpha;
		const uint32_t q0 = R[0], q1 = R[1], firstBlk = R[2];
		if constexpr (S)
			B3G_SETTLE("s"(r0), "s"(r1), "s"(active), "s"(seq), "s"(logActive), "s"(delta), "s"(logAlpha), "s"(q0),
			           "s"(q1), "s"(firstBlk), "s"(persist));
		else
			B3G_SETTLE("s"(r0), "s"(r1), "s"(seq), "s"(delta), "s"(logAlpha), "s"(q0), "s"(q1), "s"(firstBlk));
		// the SpMV's d.q partials for alpha, requested first: their sum then
		// runs while the rows are still in flight. Every input of the thread's
		// rows is loaded before the first store (the stores could alias them
		// as far as the compiler knows); a row past the block loads row r0 and
		// stores nothing.
		double pq[16];
		wg_sum_load(a.partialQ, q0, q1, pq);
		V3 d[RPT], x[RPT], r[RPT], q[RPT], im[RPT];
#pragma unroll
		for (int u = 0; u < RPT; ++u) {
			const uint32_t i = r0 + threadIdx.x + u * kBlock, ii = i < r1 ? i : r0;
			d[u] = a.dnew[ii];
			x[u] = a.px[ii];
			r[u] = a.r[ii];
			q[u] = a.q[ii];
			im[u] = a.invM[ii];
		}
		if constexpr (S)
			if (!run) return;
		// alpha_k (Alpha(k)): fused mode sums the SpMV's d.q partials of the
		// instance, and the instance's first block records it
		float alpha = logAlpha;
		if (a.fused) {
			alpha = delta / (float)wg_sum_finish(a.partialQ, q0, q1, pq);
			if (blk == firstBlk && threadIdx.x == 0) log[k].alpha = alpha;
		}
		double part = 0.0;
#pragma unroll
		for (int u = 0; u < RPT; ++u) {
			const uint32_t i = r0 + threadIdx.x + u * kBlock;
			if (i >= r1) break;
			const V3 rn = r[u] - alpha * q[u];
			const V3 sn = diag_apply(im[u], rn);
			a.px[i] = x[u] + alpha * d[u];
			a.r[i] = rn;
			a.s[i] = sn;
			const float term = dot(rn, sn);
			if (seq) a.terms[i] = term;
			part += (double)term;
		}
		block_store(part, a.partial, blk);
	};
	if (!Multi) {   // instance 0 as a constant: its reads need no load of the block's instance first
		const uint32_t blk = blockIdx.x;
		if (a.single) {
			block(blk, 0u, std::true_type{});
		} else {
			// the block's gate word with its instance: a block of an instance
			// that is done costs one round of two loads
			const uint32_t inst = a.blkInst[blk], g = a.blkGate && !force ? a.blkGate[blk] : 1u;
			B3G_SETTLE("s"(inst), "s"(g));
			if (!g) return;
			block(blk, inst, std::false_type{});
		}
		return;
	}
	auto ok = [&](uint32_t inst) {
		return force || (a.st[inst].newtonActive && a.log[(size_t)inst * a.logStride + k].active &&
		                 !persistent(a.persist, inst));
	};
	for_items<Multi, false>(a.nRowBlk, a.blkInst, ok,
	                        [&](uint32_t blk) { block(blk, a.blkInst[blk], std::false_type{}); });
}

// x += alpha d (iterations that refresh the true residual)
template <bool Multi>
__global__ void __launch_bounds__(kBlock) k_cg_update_x(CGArgs a, uint32_t k) {
	auto block = [&](uint32_t blk, uint32_t inst) {
		const uint32_t r0 = a.rowBlk[blk], r1 = a.rowBlk[blk + 1];
		const uint32_t active = a.st[inst].newtonActive, logActive = a.log[(size_t)inst * a.logStride + k].active;
		const uint32_t persist = a.persist ? a.persist[inst] : 0u;
		B3G_SETTLE("s"(r0), "s"(r1), "s"(active), "s"(logActive), "s"(persist));
		if (!(active && logActive && !persist)) return;
		V3 d[kRowsPerThread], x[kRowsPerThread];
#pragma unroll
		for (int u = 0; u < kRowsPerThread; ++u) {
			const uint32_t i = r0 + threadIdx.x + u * kBlock, ii = i < r1 ? i : r0;
			d[u] = a.dnew[ii];
			x[u] = a.px[ii];
		}
		const float alpha = cg_alpha(a, inst, blk, k);
#pragma unroll
		for (int u = 0; u < kRowsPerThread; ++u) {
			const uint32_t i = r0 + threadIdx.x + u * kBlock;
			if (i >= r1) break;
			a.px[i] = x[u] + alpha * d[u];
		}
	};
	if (!Multi) {   // instance 0 as a constant: its reads need no load of the block's instance first
		const uint32_t blk = blockIdx.x;
		if (a.single) block(blk, 0u);
		else block(blk, a.blkInst[blk]);
		return;
	}
	auto ok = [&](uint32_t inst) {
		return a.st[inst].newtonActive && a.log[(size_t)inst * a.logStride + k].active && !persistent(a.persist, inst);
	};
	for_items<Multi, false>(a.nRowBlk, a.blkInst, ok, [&](uint32_t blk) { block(blk, a.blkInst[blk]); });
}

// v += dv ; x = x0 + h v + y ; error = |dv|^2 (sparse_force_solver.cpp:193-223)
__global__ void __launch_bounds__(kBlock) k_newton_update(StepArgs s, double* partial, float* terms) {
	const uint32_t blk = blockIdx.x, inst = s.single ? 0u : s.blkInst[blk];
	const uint32_t r0 = s.rowBlk[blk], r1 = s.rowBlk[blk + 1];
	const uint32_t active = s.st[inst].newtonActive, seqw = s.st[inst].seq;
	B3G_SETTLE("s"(r0), "s"(r1), "s"(active), "s"(seqw));
	if (!active) return;
	const bool seq = seqw != 0;
	// every row's inputs before the first store; a row past the block reads row r0 and stores nothing
	V3 pxv[kRowsPerThread], vv[kRowsPerThread], x0v[kRowsPerThread], yv[kRowsPerThread];
#pragma unroll
	for (int u = 0; u < kRowsPerThread; ++u) {
		const uint32_t i = r0 + threadIdx.x + u * kBlock, ii = i < r1 ? i : r0;
		pxv[u] = s.px[ii];
		vv[u] = s.v[ii];
		x0v[u] = s.x0[ii];
		yv[u] = s.y[ii];
	}
	double part = 0.0;
#pragma unroll
	for (int u = 0; u < kRowsPerThread; ++u) {
		const uint32_t i = r0 + threadIdx.x + u * kBlock;
		if (i >= r1) break;
		V3 dv = pxv[u] + zero3();
		V3 v = vv[u] + dv;
		const V3 x = (x0v[u] + s.h * v) + yv[u];
		s.v[i] = v;
		s.x[i] = x;
		if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) {
			atomicOr(&s.st[inst].faults, (uint32_t)B3G_FAULT_NON_FINITE);
			atomicOr(s.faultWord, (uint32_t)B3G_FAULT_NON_FINITE);
		}
		const float term = length_sq(dv);
		if (seq) terms[i] = term;
		part += (double)term;
	}
	block_store(part, partial, blk);
}

__global__ void __launch_bounds__(kBlock) k_clear_external(uint32_t n, V3* force, V3* translation) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	force[i] = zero3();
	translation[i] = zero3();
}

// Persistent PCG (Schedule::persistTiles): each workgroup runs, in turn, the
// whole PCG solve of the listed instances (sparse_solver.cpp:62-103) from the
// state the solve's start left (r, s, delta_0, log[0]), iteration after
// iteration in one launch, until the instance's own convergence test stops it.
// Its SpMV tiles, row blocks, FP64 partials and their fixed-order sums are
// the launch sequence's, in the same order, with the fused reductions' scalar
// steps, and it writes the same log entries; so every instance's results are
// bit-identical to the launch sequence's. Only this workgroup touches the
// instance's rows: __syncthreads orders its stores before its loads. Its r.s
// partials go to arrays of its own, tiles to partialT and row blocks to
// partialK: its instances run out of step with one another and with the launch
// sequence, whose partial array holds tile and row-block partials by turns
// (one instance's tile indices overlap another's row-block indices).
template <int St>
__global__ void __launch_bounds__(kBlock) k_pcg_persistent(CGArgs a, const uint32_t* plist, uint32_t count, V3* d0,
                                                           V3* d1, double* partialT, double* partialK) {
	__shared__ float lds[3 * kTileCap];
	for (uint32_t j = blockIdx.x; j < count; j += gridDim.x) {
		const uint32_t inst = plist[j];
		SolverState& st = a.st[inst];
		CGLog* log = a.log + (size_t)inst * a.logStride;
		const uint32_t* R = a.range + 6 * (size_t)inst;
		const uint32_t t0 = R[0], t1 = R[1], k0 = R[2], k1 = R[3];
		if (!st.newtonActive || !log[0].active) continue;
		const uint32_t maxIt = st.maxIt;
		const float eps2 = st.eps2_sub, delta0 = st.delta0;
		float beta = 0.0f;
		for (uint32_t k = 0;; ++k) {
			V3* dold = (k & 1) ? d1 : d0;
			V3* dnew = (k & 1) ? d0 : d1;
			if (k > 0) {   // Delta(k), from the r.s partials of iteration k - 1 (tiles after a refresh)
				const bool refreshed = (k - 1) % 50 == 0;
				const float v = refreshed ? (float)wg_sum(partialT, t0, t1) : (float)wg_sum(partialK, k0, k1);
				beta = v / log[k - 1].delta;
				const bool act = k != maxIt && !(v <= eps2 * delta0);
				if (threadIdx.x == 0) {
					log[k].delta = v;
					log[k].beta = beta;
					log[k].active = act ? 1u : 0u;
					st.it = k;
				}
				if (!act) break;
			}
			// q = A d, d = s + beta d_old (d = s at k = 0), partials of d.q (k_cg_spmv_dq)
			const V3* dsrc = k == 0 ? a.s : dold;
			auto dval = [&](uint32_t c) { return a.s[c] + beta * dsrc[c]; };
			for (uint32_t t = t0; t < t1; ++t) {
				const TileRange tr = tile_range<St>(a, t);
				uint32_t row;
				V3 q;
				tile_rows<St>(a, tr, row, dval, lds, no_prologue, q);
				float part = 0.0f;
				const bool has = row < tr.r1;
				if (has) {
					const V3 d = dval(row);
					dnew[row] = d;
					a.q[row] = q;
					part = dot(d, q);
				}
				block_partial(part, has, row, a.partialQ, a.terms, false, t);
			}
			__syncthreads();
			const float alpha = log[k].delta / (float)wg_sum(a.partialQ, t0, t1);   // Alpha(k)
			if (threadIdx.x == 0) log[k].alpha = alpha;
			if (k % 50 == 0) {
				// x += alpha d, then the true residual r = b - A x, s = invM r (k_cg_update_x, k_cg_residual)
				for (uint32_t blk = k0; blk < k1; ++blk) {
					const uint32_t r0 = a.rowBlk[blk], r1 = a.rowBlk[blk + 1];
#pragma unroll
					for (int u = 0; u < kRowsPerThread; ++u) {
						const uint32_t i = r0 + threadIdx.x + u * kBlock;
						if (i >= r1) break;
						a.px[i] = a.px[i] + alpha * dnew[i];
					}
				}
				__syncthreads();
				for (uint32_t t = t0; t < t1; ++t) {
					const TileRange tr = tile_range<St>(a, t);
					uint32_t row;
					V3 ax;
					tile_rows<St>(a, tr, row, [&](uint32_t c) { return a.px[c]; }, lds, no_prologue, ax);
					float part = 0.0f;
					const bool has = row < tr.r1;
					if (has) {
						const V3 r = a.pb[row] - ax;
						const V3 s = diag_apply(a.invM[row], r);
						a.r[row] = r;
						a.s[row] = s;
						part = dot(r, s);
					}
					block_partial(part, has, row, partialT, a.terms, false, t);
				}
			} else {
				// x += alpha d; r -= alpha q; s = invM r; partials of r.s (k_cg_update)
				for (uint32_t blk = k0; blk < k1; ++blk) {
					const uint32_t r0 = a.rowBlk[blk], r1 = a.rowBlk[blk + 1];
					double part = 0.0;
#pragma unroll
					for (int u = 0; u < kRowsPerThread; ++u) {
						const uint32_t i = r0 + threadIdx.x + u * kBlock;
						if (i >= r1) break;
						const V3 rn = a.r[i] - alpha * a.q[i];
						const V3 sn = diag_apply(a.invM[i], rn);
						a.px[i] = a.px[i] + alpha * dnew[i];
						a.r[i] = rn;
						a.s[i] = sn;
						part += (double)dot(rn, sn);
					}
					block_store(part, partialK, blk);
				}
			}
			__syncthreads();
		}
		__syncthreads();   // the next instance reuses the shared memory
	}
}

// Plain SpMV y = A x (diagnostics / operator-level entry point), reference row order.
template <int St>
__global__ void __launch_bounds__(kBlock) k_spmv(CGArgs a, const V3* xin, V3* yout) {
	__shared__ float lds[3 * kTileCap];
	const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
	const TileRange tr = tile_range<St>(a, t);
	uint32_t row;
	V3 y;
	tile_rows<St>(a, tr, row, [&](uint32_t c) { return xin[c]; }, lds, no_prologue, y);
	if (row < tr.r1) yout[row] = y;
}

// ---------------------------------------------------------------------------
// Precision form: FP64 iterative refinement around the PCG
// ---------------------------------------------------------------------------
struct RefineArgs {
	uint32_t n, nnz;
	uint32_t stride;          // plane stride of A
	const uint32_t* tiles;
	const uint32_t* tileInst;
	const uint32_t* rowptr;
	const uint32_t* cols;
	const uint32_t* slotRow;
	const uint32_t* rowInst;
	const float* A;
	const V3* pb;
	const V3* invM;
	double* x64;     // 3 per row
	const double* pb64;   // 3 per row: the FP64 right-hand side
	V3* px;
	V3* dx;
	V3* rhs;
	V3* r;
	V3* s;
	double* partialR;
	double* partial;
	SolverState* st;
	uint16_t* AbP16;   // the pair planes as 16-bit entries: plane c, entry 2 pairBase[t] + (k - b0)
	const uint32_t* pairBase;
	uint32_t pairStride;
	float* diagF;
	// column-sorted storage (Body::sortedA), or null
	const uint32_t* code;
	const uint32_t* tileBase;
	const uint32_t* storeSlot;
};

// Round-to-nearest-even FP32 -> BF16 (NaN stays NaN).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
	uint32_t u = __float_as_uint(f);
	if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return u >> 16;
}

// BF16 copy of pA's off-diagonal blocks (diagonal slots zero) in pair
// planes and the FP32 diagonal blocks per row, once per linearisation; one
// workgroup per SpMV tile. A tile of odd length gets a zero high half in its
// last pair.
__global__ void __launch_bounds__(kBlock) k_split_bf16(RefineArgs a) {
	const uint32_t t = blockIdx.x;
	if (!a.st[a.tileInst[t]].newtonActive) return;
	const uint32_t b0 = a.rowptr[a.tiles[t]], b1 = a.rowptr[a.tiles[t + 1]];
	const size_t plane = 2 * (size_t)a.pairStride;
	const uint32_t e0 = 2 * a.pairBase[t];
	for (uint32_t k = b0 + threadIdx.x; k < b1; k += kBlock) {
		const uint32_t slot = a.storeSlot ? a.storeSlot[k] : k;   // the CSR slot stored at position k
		const uint32_t row = a.slotRow[slot];
		const uint32_t e = e0 + (k - b0);
		M33 m = ldA(a.A, a.stride, k);
		const float* p = &m.x.x;
		const bool diag = a.cols[slot] == row;
		if (diag)
			for (int c = 0; c < 9; ++c) a.diagF[(size_t)c * a.n + row] = p[c];
		for (int c = 0; c < 9; ++c) a.AbP16[c * plane + e] = diag ? (uint16_t)0 : (uint16_t)bf16_bits(p[c]);
		if (k + 1 == b1 && ((k - b0) & 1u) == 0)
			for (int c = 0; c < 9; ++c) a.AbP16[c * plane + e + 1] = 0;
	}
}

// r = pb - pA x in FP64 (FP32 blocks promoted, FP64 x and sums, the
// reference's row order), partials of |r|^2; then the sweep's inner system:
// rhs = (float) r, dx = 0, and the PCG start r_in = rhs, s_in = invM rhs with
// partials of r_in.s_in (the inner solve starts from dx = 0, so its first
// residual needs no product).
__global__ void __launch_bounds__(kBlock) k_resid64(RefineArgs a, int first) {
	const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
	const uint32_t inst = a.tileInst[t];
	const SolverState& st = a.st[inst];
	if (!st.newtonActive || (!first && !st.refineActive)) return;
	__shared__ double lds[3 * kTileCap];
	const uint32_t r0 = a.tiles[t], r1 = a.tiles[t + 1];
	const uint32_t b0 = a.rowptr[r0], b1 = a.rowptr[r1];
	const uint32_t base = a.code ? a.tileBase[t] : 0u;
	const uint32_t row = r0 + threadIdx.x;
	const bool own = row < r1;
	const uint32_t rb = own ? a.rowptr[row] : 0, re = own ? a.rowptr[row + 1] : 0;
	double ax = 0.0, ay = 0.0, az = 0.0;
	constexpr uint32_t cap = kTileCap;   // the chunks of the column-sorted storage
	for (uint32_t cb = b0; cb < b1; cb += cap) {
		const uint32_t ce = min(cb + cap, b1);
		for (uint32_t k = cb + threadIdx.x; k < ce; k += kBlock) {
			M33 m = ldA(a.A, a.stride, k);
			uint32_t col, dst;
			if (a.code) {
				const uint32_t c = a.code[k];
				col = base + (c >> 10);
				dst = c & 1023u;
			} else {
				col = a.cols[k];
				dst = k - cb;
			}
			const double* xv = a.x64 + 3 * (size_t)col;
			const double x0 = xv[0], x1 = xv[1], x2 = xv[2];
			lds[3 * dst] = (double)m.x.x * x0 + (double)m.y.x * x1 + (double)m.z.x * x2;
			lds[3 * dst + 1] = (double)m.x.y * x0 + (double)m.y.y * x1 + (double)m.z.y * x2;
			lds[3 * dst + 2] = (double)m.x.z * x0 + (double)m.y.z * x1 + (double)m.z.z * x2;
		}
		__syncthreads();
		if (own) {
			const uint32_t kb = max(rb, cb), ke = min(re, ce);
			for (uint32_t k = kb; k < ke; ++k) {
				ax += lds[3 * (k - cb)];
				ay += lds[3 * (k - cb) + 1];
				az += lds[3 * (k - cb) + 2];
			}
		}
		__syncthreads();
	}
	__shared__ double sh[kBlock / 64];
	double rr = 0.0, rs = 0.0;
	if (own) {
		const double* b = a.pb64 + 3 * (size_t)row;
		const double rx = b[0] - ax, ry = b[1] - ay, rz = b[2] - az;
		rr = rx * rx + ry * ry + rz * rz;
		const V3 rf = v3((float)rx, (float)ry, (float)rz);
		const V3 sf = diag_apply(a.invM[row], rf);
		a.rhs[row] = rf;
		a.r[row] = rf;
		a.s[row] = sf;
		a.dx[row] = zero3();
		rs = (double)dot(rf, sf);
	}
	const double brr = block_sum(rr, sh);
	if (threadIdx.x == 0) a.partialR[t] = brr;
	const double brs = block_sum(rs, sh);
	if (threadIdx.x == 0) a.partial[t] = brs;
}

// x64 = px (the warm start) at the start of a Newton iteration's solve.
__global__ void __launch_bounds__(kBlock) k_x64_init(RefineArgs a) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n || !a.st[a.rowInst[i]].newtonActive) return;
	const V3 x = a.px[i];
	a.x64[3 * (size_t)i] = x.x;
	a.x64[3 * (size_t)i + 1] = x.y;
	a.x64[3 * (size_t)i + 2] = x.z;
}

// x64 += dx after a sweep that solved.
__global__ void __launch_bounds__(kBlock) k_x64_add(RefineArgs a) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n) return;
	const SolverState& st = a.st[a.rowInst[i]];
	if (!st.newtonActive || !st.refineActive) return;
	const V3 d = a.dx[i];
	a.x64[3 * (size_t)i] += (double)d.x;
	a.x64[3 * (size_t)i + 1] += (double)d.y;
	a.x64[3 * (size_t)i + 2] += (double)d.z;
}

// px = (float) x64: the Newton update's dv.
__global__ void __launch_bounds__(kBlock) k_x64_finish(RefineArgs a) {
	uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n || !a.st[a.rowInst[i]].newtonActive) return;
	a.px[i] = v3((float)a.x64[3 * (size_t)i], (float)a.x64[3 * (size_t)i + 1], (float)a.x64[3 * (size_t)i + 2]);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline uint32_t nblk(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// The persistent PCG serves the plain solver in the FP64 reduction modes (its
// sums are the fused reductions'; the sequential mode keeps the launches).
static inline bool use_persistent(const Body& b) {
	return b.nPersist > 0 && b.persist.p && b.solver == B3G_SOLVER_PCG && b.reduction != B3G_REDUCE_SEQUENTIAL;
}

// Workgroups of a PCG launch over `items` tiles or row blocks (for_items):
// one per item, or the schedule's maxGroups when that is at most half the
// items. One per item is the default: on the 4096-body batch, 2048 ranged
// workgroups took 84 ms per step against 69 ms (a workgroup that starts
// late runs its whole range late, and the range loop holds more registers),
// more than the converged tiles' early returns cost.
static inline uint32_t groups(const Body& b, uint32_t items) {
	const uint32_t cap = b.sched.maxGroups ? b.sched.maxGroups : items;
	if (items / 2 < cap) return std::max(1u, items);
	return cap;
}

#define LAUNCH(kernel, grid, ...)                                                                    \
	do {                                                                                             \
		if ((grid) > 0) hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, b.stream, __VA_ARGS__); \
		B3G_HIP(hipGetLastError());                                                                  \
	} while (0)

// A PCG kernel over `items` tiles or row blocks: one workgroup per item, or
// groups(b, items) workgroups that each take a range (for_items).
#define LAUNCH_ITEMS(kernel, items, ...)                                     \
	do {                                                                     \
		const uint32_t n_ = (items), g_ = groups(b, n_);                     \
		if (g_ < n_) LAUNCH(kernel<true>, g_, __VA_ARGS__);                  \
		else LAUNCH(kernel<false>, n_, __VA_ARGS__);                         \
	} while (0)

// The vector update of iteration k with the fewest rows per thread that
// cover the schedule's row blocks.
template <int RPT>
static b3g_status launch_update_rpt(Body& b, const CGArgs& a, uint32_t k, int force) {
	const uint32_t n = b.nRowBlk, g = groups(b, n);
	if (n == 0) return B3G_OK;
	if (g < n) hipLaunchKernelGGL((k_cg_update<true, RPT>), dim3(g), dim3(kBlock), 0, b.stream, a, k, force);
	else hipLaunchKernelGGL((k_cg_update<false, RPT>), dim3(n), dim3(kBlock), 0, b.stream, a, k, force);
	B3G_HIP(hipGetLastError());
	return B3G_OK;
}

static b3g_status launch_update(Body& b, const CGArgs& a, uint32_t k, int force) {
	const uint32_t rows = std::min(b.sched.rowBlockRows, kRowBlkRows);
	if (rows <= (uint32_t)kBlock) return launch_update_rpt<1>(b, a, k, force);
	if (rows <= 2u * kBlock) return launch_update_rpt<2>(b, a, k, force);
	return launch_update_rpt<kRowsPerThread>(b, a, k, force);
}

// The storage instantiation of a kernel that reads pA (tile_rows).
static inline int tile_storage(const CGArgs& a) { return (a.code ? 1 : 0) | (a.bf16 ? 2 : 0); }

// A tile kernel (templated <Multi, St>) over `items` tiles for the storage of
// the CGArgs `cga`: one workgroup per tile, or groups(b, items) that each
// take a range (for_items). Leaves the launch status in hipGetLastError().
#define TILE_KERNEL(kernel, cga, items, ...)                                                             \
	do {                                                                                                 \
		const uint32_t n_ = (items), g_ = groups(b, n_);                                                 \
		if (n_ > 0) {                                                                                    \
			const bool m_ = g_ < n_;                                                                     \
			const dim3 gr_(m_ ? g_ : n_);                                                                \
			switch (tile_storage(cga) * 2 + (m_ ? 1 : 0)) {                                              \
			case 0: hipLaunchKernelGGL((kernel<false, 0>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break; \
			case 1: hipLaunchKernelGGL((kernel<true, 0>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break;  \
			case 2: hipLaunchKernelGGL((kernel<false, 1>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break; \
			case 3: hipLaunchKernelGGL((kernel<true, 1>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break;  \
			case 4: hipLaunchKernelGGL((kernel<false, 2>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break; \
			case 5: hipLaunchKernelGGL((kernel<true, 2>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break;  \
			case 6: hipLaunchKernelGGL((kernel<false, 3>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break; \
			default: hipLaunchKernelGGL((kernel<true, 3>), gr_, dim3(kBlock), 0, b.stream, __VA_ARGS__); break; \
			}                                                                                            \
		}                                                                                                \
	} while (0)
#define LAUNCH_TILES(kernel, cga, items, ...)       \
	do {                                            \
		TILE_KERNEL(kernel, cga, items, __VA_ARGS__); \
		B3G_HIP(hipGetLastError());                 \
	} while (0)

// Row decompositions, both cut at instance boundaries so that every tile /
// block belongs to one instance: SpMV tiles (<= sched.tileRows rows,
// <= sched.tileBlocks blocks unless one row has more) and row blocks
// (<= sched.rowBlockRows rows) for the vector kernels; plus each instance's
// ranges of tiles, blocks and rows. With tiles = false the SpMV tiles (and
// the column-sorted storage built on them) are kept and only the row blocks
// are cut again.
b3g_status build_tiles(Body& b, bool sorted, bool tiles) {
	const uint32_t K = std::max(1u, b.nInst);
	auto inst = [&](uint32_t i) { return b.hostRowInst.empty() ? 0u : b.hostRowInst[i]; };
	const uint32_t tileRows = std::min(b.sched.tileRows, (uint32_t)kTileRows);
	const uint32_t tileCap = std::min(b.sched.tileBlocks, (uint32_t)kTileCap);
	const uint32_t blkRows = std::min(b.sched.rowBlockRows, kRowBlkRows);
	std::vector<uint32_t>& t = b.hostTiles;
	std::vector<uint32_t>& ti = b.hostTileInst;
	std::vector<uint32_t> rb{0}, bi;
	if (tiles || t.empty()) {
		t.assign(1, 0u);
		ti.clear();
		tiles = true;
	}
	uint32_t rows = 0, blocks = 0, brows = 0;
	for (uint32_t i = 0; i < b.n; ++i) {
		const uint32_t len = b.hostRowptr[i + 1] - b.hostRowptr[i];
		const bool newInst = i > 0 && inst(i) != inst(i - 1);
		if (tiles && rows > 0 && (newInst || rows == tileRows || blocks + len > tileCap)) {
			t.push_back(i);
			ti.push_back(inst(i - 1));
			rows = 0;
			blocks = 0;
		}
		if (brows > 0 && (newInst || brows == blkRows)) {
			rb.push_back(i);
			bi.push_back(inst(i - 1));
			brows = 0;
		}
		rows += 1;
		blocks += len;
		brows += 1;
	}
	if (tiles) {
		t.push_back(b.n);
		ti.push_back(b.n ? inst(b.n - 1) : 0u);
	}
	rb.push_back(b.n);
	bi.push_back(b.n ? inst(b.n - 1) : 0u);
	b.nTiles = (uint32_t)t.size() - 1;
	b.nRowBlk = (uint32_t)rb.size() - 1;
	// per instance: [tile begin, end), [row block begin, end), [row begin, end)
	std::vector<uint32_t> range(6 * (size_t)K, 0);
	for (uint32_t k = 0; k < K; ++k) {
		range[6 * k] = range[6 * k + 2] = range[6 * k + 4] = 0xFFFFFFFFu;
	}
	auto extend = [&](uint32_t k, int slot, uint32_t lo, uint32_t hi) {
		uint32_t& a = range[6 * k + slot];
		uint32_t& z = range[6 * k + slot + 1];
		if (a == 0xFFFFFFFFu) { a = lo; z = hi; } else { a = std::min(a, lo); z = std::max(z, hi); }
	};
	for (uint32_t x = 0; x < b.nTiles; ++x) extend(ti[x], 0, x, x + 1);
	for (uint32_t x = 0; x < b.nRowBlk; ++x) extend(bi[x], 2, x, x + 1);
	for (uint32_t i = 0; i < b.n; ++i) extend(inst(i), 4, i, i + 1);
	for (uint32_t k = 0; k < K; ++k)
		for (int slot = 0; slot < 6; slot += 2)
			if (range[6 * k + slot] == 0xFFFFFFFFu) range[6 * k + slot] = range[6 * k + slot + 1] = 0;
	b.maxInstPartials = 0;
	for (uint32_t k = 0; k < K; ++k)
		b.maxInstPartials = std::max({b.maxInstPartials, range[6 * k + 1] - range[6 * k], range[6 * k + 3] - range[6 * k + 2]});
	b3g_status s;
	if (tiles) {
		// column-sorted storage within each chunk (see Body::sortedA); a tile whose
		// columns span 2^22 rows or more cannot be coded, and the CSR storage stays
		b.sortedA = false;
		if (sorted && b.nnz) {
			std::vector<uint32_t> code(b.aStride, 0u), store(b.aStride, 0u), pos(b.nnz), base(b.nTiles, 0u);
			bool ok = true;
			std::vector<uint32_t> idx;
			for (uint32_t x = 0; x < b.nTiles && ok; ++x) {
				const uint32_t b0 = b.hostRowptr[t[x]], b1 = b.hostRowptr[t[x + 1]];
				uint32_t lo = 0xFFFFFFFFu;
				for (uint32_t k = b0; k < b1; ++k) lo = std::min(lo, b.hostCols[k]);
				base[x] = b1 > b0 ? lo : 0u;
				for (uint32_t cb = b0; cb < b1 && ok; cb += kTileCap) {
					const uint32_t ce = std::min(cb + (uint32_t)kTileCap, b1);
					idx.resize(ce - cb);
					for (uint32_t k = cb; k < ce; ++k) idx[k - cb] = k;
					std::stable_sort(idx.begin(), idx.end(), [&](uint32_t u, uint32_t v) { return b.hostCols[u] < b.hostCols[v]; });
					for (uint32_t j = 0; j < (uint32_t)idx.size(); ++j) {
						const uint32_t src = idx[j], dst = cb + j, off = b.hostCols[src] - base[x];
						if (off >= (1u << 22)) { ok = false; break; }
						code[dst] = (off << 10) | (src - cb);
						pos[src] = dst;
						store[dst] = src;
					}
				}
			}
			if (ok) {
				if ((s = dupload(b, b.code, code.data(), code.size() * 4)) != B3G_OK) return s;
				if ((s = dupload(b, b.storeSlot, store.data(), store.size() * 4)) != B3G_OK) return s;
				if ((s = dupload(b, b.slotPos, pos.data(), pos.size() * 4)) != B3G_OK) return s;
				if ((s = dupload(b, b.tileBase, base.data(), base.size() * 4)) != B3G_OK) return s;
				b.sortedA = true;
			}
		}
		if ((s = dupload(b, b.tiles, t.data(), t.size() * 4)) != B3G_OK) return s;
		if ((s = dupload(b, b.tileInst, ti.data(), ti.size() * 4)) != B3G_OK) return s;
		std::vector<uint32_t> tb(t.size());
		for (size_t x = 0; x < t.size(); ++x) tb[x] = b.hostRowptr[t[x]];
		if ((s = dupload(b, b.tileBlk, tb.data(), tb.size() * 4)) != B3G_OK) return s;
		std::vector<uint32_t> gate(std::max(1u, b.nTiles), 1u);
		if ((s = dupload(b, b.tileGate, gate.data(), gate.size() * 4)) != B3G_OK) return s;
		// the BF16 pair planes: each tile's first pair
		std::vector<uint32_t> pairs(std::max(1u, b.nTiles), 0u);
		uint64_t total = 0;
		for (uint32_t x = 0; x < b.nTiles; ++x) {
			pairs[x] = (uint32_t)total;
			total += (b.hostRowptr[t[x + 1]] - b.hostRowptr[t[x]] + 1u) / 2u;
		}
		if (total > 0x7FFFFFC0ull) { set_error("too many blocks for the BF16 pair planes"); return B3G_UNSUPPORTED; }
		b.pairStride = (uint32_t)((total + 63u) & ~63ull);
		if (b.pairStride == 0) b.pairStride = 64;
		if ((s = dupload(b, b.pairBase, pairs.data(), pairs.size() * 4)) != B3G_OK) return s;
	}
	if ((s = dupload(b, b.rowBlk, rb.data(), rb.size() * 4)) != B3G_OK) return s;
	{
		std::vector<uint32_t> gate(std::max<size_t>(1, rb.size()), 1u);
		if ((s = dupload(b, b.blkGate, gate.data(), gate.size() * 4)) != B3G_OK) return s;
	}
	if ((s = dupload(b, b.blkInst, bi.data(), bi.size() * 4)) != B3G_OK) return s;
	if ((s = dupload(b, b.instRange, range.data(), range.size() * 4)) != B3G_OK) return s;
	// the persistent PCG's instances: at most sched.persistTiles tiles, largest first
	b.nPersist = 0;
	if (b.sched.persistTiles > 0) {
		std::vector<uint32_t> flag(K, 0u), list;
		for (uint32_t k = 0; k < K; ++k) {
			const uint32_t nt = range[6 * k + 1] - range[6 * k];
			if (nt > 0 && nt <= b.sched.persistTiles) {
				flag[k] = 1u;
				list.push_back(k);
			}
		}
		std::stable_sort(list.begin(), list.end(), [&](uint32_t u, uint32_t v) {
			return range[6 * u + 5] - range[6 * u + 4] > range[6 * v + 5] - range[6 * v + 4];
		});
		if (!list.empty()) {
			if ((s = dupload(b, b.persist, flag.data(), flag.size() * 4)) != B3G_OK) return s;
			if ((s = dupload(b, b.plist, list.data(), list.size() * 4)) != B3G_OK) return s;
			const size_t pk = ((size_t)b.nTiles + b.nRowBlk + 1) * sizeof(double);   // tile, then row-block partials
			if ((!b.partialP.p || b.partialP.bytes < pk) && (s = dalloc(b, b.partialP, pk)) != B3G_OK) return s;
			if (!b.stream2) B3G_HIP(hipStreamCreateWithFlags(&b.stream2, hipStreamNonBlocking));
			if (!b.forkEvent) B3G_HIP(hipEventCreateWithFlags(&b.forkEvent, hipEventDisableTiming));
			if (!b.joinEvent) B3G_HIP(hipEventCreateWithFlags(&b.joinEvent, hipEventDisableTiming));
			b.nPersist = (uint32_t)list.size();
		}
	}
	return ensure_partials(b, std::max(b.nTiles, b.nRowBlk));
}

b3g_status ensure_partials(Body& b, uint32_t blocks) {
	if (!b.terms.p || b.terms.bytes < (size_t)b.n * sizeof(float)) {
		b3g_status s = dalloc(b, b.terms, (size_t)std::max(1u, b.n) * sizeof(float));
		if (s != B3G_OK) return s;
	}
	if (b.partialCount >= blocks && b.partial.p && b.partialQ.p) return B3G_OK;
	b3g_status s = dalloc(b, b.partial, (size_t)blocks * sizeof(double));
	if (s != B3G_OK) return s;
	s = dalloc(b, b.partialQ, (size_t)blocks * sizeof(double));
	if (s != B3G_OK) return s;
	b.partialCount = blocks;
	return B3G_OK;
}

// Room for `entries` PCG log entries per instance (a solve of maxIt
// iterations uses maxIt + 1).
__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* p, size_t count, uint32_t value) {
	const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (i < count) p[i] = value;
}

// The per-instance PCG log (zeroed when it is (re)allocated). Steps allocate
// it before any capture, so a captured step never reaches the allocation.
b3g_status ensure_cg_log(Body& b, uint32_t entries) {
	if (b.cgLog.p && b.logStride >= entries) return B3G_OK;
	const uint32_t stride = std::max(entries, 64u);
	B3G_HIP(hipStreamSynchronize(b.stream));   // kernels in flight may still read the old log
	const size_t bytes = (size_t)std::max(1u, b.nInst) * stride * sizeof(CGLog);
	b3g_status s = dalloc(b, b.cgLog, bytes);
	if (s != B3G_OK) return s;
	b.logStride = stride;
	const size_t words = bytes / 4;
	hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((words + kBlock - 1) / kBlock)), dim3(kBlock), 0, b.stream,
	                   dptr<uint32_t>(b.cgLog), words, 0u);
	B3G_HIP(hipGetLastError());
	return B3G_OK;
}

b3g_status launch_tet_setup(Body& b) {
	DBuf flag;
	b3g_status s = dalloc(b, flag, 4);
	if (s != B3G_OK) return s;
	B3G_HIP(hipMemsetAsync(flag.p, 0, 4, b.stream));
	LAUNCH(k_tet_setup, nblk(b.tet.m), b.tet.m, dptr<const float>(b.tet.rest), dptr<const float>(b.tet.param),
	       dptr<float4>(b.tet.rec), dptr<uint32_t>(flag));
	uint32_t deg = 0;
	B3G_HIP(hipMemcpyAsync(&deg, flag.p, 4, hipMemcpyDeviceToHost, b.stream));
	B3G_HIP(hipStreamSynchronize(b.stream));
	(void)hipFree(flag.p);
	b.owned.erase(std::remove(b.owned.begin(), b.owned.end(), &flag), b.owned.end());
	if (deg) {
		set_error("degenerate tetrahedron element (zero volume)");
		return B3G_DEGENERATE_ELEMENT;
	}
	return B3G_OK;
}

static StepArgs step_args(Body& b, float h) {
	StepArgs a;
	a.n = b.n;
	a.h = h;
	a.gravity = dptr<V3>(b.gravity);
	a.rowInst = dptr<uint32_t>(b.rowInst);
	a.rowBlk = dptr<uint32_t>(b.rowBlk);
	a.blkInst = dptr<uint32_t>(b.blkInst);
	a.type = dptr<uint8_t>(b.type);
	a.mass = dptr<float>(b.mass);
	a.damping = dptr<float>(b.damping);
	a.x = dptr<V3>(b.x); a.v = dptr<V3>(b.v); a.force = dptr<V3>(b.force); a.translation = dptr<V3>(b.translation);
	a.x0 = dptr<V3>(b.x0); a.v0 = dptr<V3>(b.v0); a.fe = dptr<V3>(b.fe); a.y = dptr<V3>(b.y); a.w = dptr<V3>(b.w);
	a.px = dptr<V3>(b.px);
	a.st = dptr<SolverState>(b.state);
	a.range = dptr<uint32_t>(b.instRange);
	a.faultWord = dptr<uint32_t>(b.faultWord);
	a.single = b.nInst <= 1 ? 1u : 0u;
	return a;
}

static ContactArgs contact_args(Body& b, float dt) {
	ContactArgs c;
	std::memset(&c, 0, sizeof(c));
	c.nSpheres = b.nSpheres;
	c.nPairs = b.nPairs;
	c.worlds = dptr<WorldShape>(b.worlds);
	c.shapes.vox = dptr<float>(b.voxels);
	c.shapes.meshV = dptr<V3>(b.meshV);
	c.shapes.meshT = dptr<uint32_t>(b.meshT);
	c.shapes.meshLeaf = dptr<uint32_t>(b.meshLeaf);
	c.shapes.meshLeafBox = dptr<RtBox>(b.meshLeafBox);
	c.shapes.meshChild = dptr<uint2>(b.meshChild);
	c.shapes.meshNodeBox = dptr<RtBox>(b.meshNodeBox);
	c.sphereParticle = dptr<uint32_t>(b.sphereParticle);
	c.sphereRadius = dptr<float>(b.sphereRadius);
	c.sphereFriction = dptr<float>(b.sphereFriction);
	c.pairStart = dptr<uint32_t>(b.pairStart);
	c.pairSphere = dptr<uint32_t>(b.pairSphere);
	c.pairWorld = dptr<uint32_t>(b.pairWorld);
	c.type = dptr<uint8_t>(b.type);
	c.invMass = dptr<float>(b.invMass);
	c.x = dptr<V3>(b.x);
	c.v = dptr<V3>(b.v);
	c.active = dptr<uint32_t>(b.cActive);
	c.stamp = dptr<uint32_t>(b.cStamp);
	c.applyFriction = dptr<uint32_t>(b.cFriction);
	c.normalForce = dptr<float>(b.cNormalForce);
	c.manifold = dptr<V3>(b.cManifold);
	c.hit = dptr<uint32_t>(b.cHit);
	c.cF = dptr<V3>(b.cF);
	c.cK = dptr<M33>(b.cK);
	c.stampCounter = dptr<uint32_t>(b.stampCounter);
	c.dt = dt;
	c.st = dptr<SolverState>(b.state);
	c.rowInst = dptr<uint32_t>(b.rowInst);
	return c;
}

static CGArgs cg_args(Body& b) {
	CGArgs a;
	a.n = b.n;
	a.nnz = b.nnz;
	a.stride = b.aStride;
	a.nTiles = b.nTiles;
	a.nRowBlk = b.nRowBlk;
	a.tiles = dptr<uint32_t>(b.tiles);
	a.tileInst = dptr<uint32_t>(b.tileInst);
	a.tileBlk = dptr<uint32_t>(b.tileBlk);
	a.tileGate = b.nInst > 1 ? dptr<uint32_t>(b.tileGate) : nullptr;
	a.blkGate = b.nInst > 1 ? dptr<uint32_t>(b.blkGate) : nullptr;
	a.rowBlk = dptr<uint32_t>(b.rowBlk);
	a.blkInst = dptr<uint32_t>(b.blkInst);
	a.single = b.nInst <= 1 ? 1u : 0u;
	a.rowptr = dptr<uint32_t>(b.rowptr);
	a.cols = dptr<uint32_t>(b.cols);
	a.A = dptr<float>(b.A);
	a.pb = dptr<V3>(b.pb);
	a.invM = dptr<V3>(b.invM);
	a.px = dptr<V3>(b.px);
	a.r = dptr<V3>(b.r);
	a.s = dptr<V3>(b.s);
	a.q = dptr<V3>(b.q);
	a.dold = dptr<V3>(b.d0);
	a.dnew = dptr<V3>(b.d1);
	a.partial = dptr<double>(b.partial);
	a.partialQ = dptr<double>(b.partialQ);
	a.terms = dptr<float>(b.terms);
	a.st = dptr<SolverState>(b.state);
	a.range = dptr<uint32_t>(b.instRange);
	a.log = dptr<CGLog>(b.cgLog);
	a.logStride = b.logStride;
	// the default FP64 mode fuses while the consumers' re-reads of the
	// partials stay cheaper than two extra launches per iteration (measured
	// crossover between 3.5k and 11.5k partials per instance)
	bool fused = b.reduction == B3G_REDUCE_FP64_FUSED;
	if (b.reduction == B3G_REDUCE_FP64)
		fused = b.sched.fused >= 0 ? b.sched.fused != 0 : b.maxInstPartials <= kFuseMaxPartials;
	a.fused = fused ? 1u : 0u;
	a.bf16 = b.solver == B3G_SOLVER_REFINED_BF16 && b.AbP.p ? 1u : 0u;
	a.AbP = dptr<uint32_t>(b.AbP);
	a.pairBase = dptr<uint32_t>(b.pairBase);
	a.pairStride = b.pairStride;
	a.diagF = dptr<float>(b.diagF);
	a.code = b.sortedA ? dptr<uint32_t>(b.code) : nullptr;
	a.tileBase = b.sortedA ? dptr<uint32_t>(b.tileBase) : nullptr;
	a.persist = use_persistent(b) ? dptr<uint32_t>(b.persist) : nullptr;
	return a;
}

b3g_status enqueue_update_contacts(Body& b) {
	if (!b.nPairs) return B3G_OK;
	ContactArgs c = contact_args(b, 0.0f);
	LAUNCH(k_update_contacts, nblk(b.nPairs), c);
	return B3G_OK;
}

b3g_status enqueue_find_contacts(Body& b) {
	if (!b.nPairs) return B3G_OK;
	ContactArgs c = contact_args(b, 0.0f);
	LAUNCH(k_find_contacts, nblk(b.nPairs), c);
	hipLaunchKernelGGL(k_stamp_advance, dim3(1), dim3(1), 0, b.stream, c.stampCounter);
	B3G_HIP(hipGetLastError());
	return B3G_OK;
}

// One linearisation: element kernels, contacts, pA / pb assembly.
// first: the step's first Newton iteration, which every instance runs (its
// element and assembly kernels then read no gate).
static b3g_status enqueue_linearise(Body& b, float h, bool first) {
	const SolverState* st = first ? nullptr : dptr<SolverState>(b.state);
	const uint32_t* ri = dptr<uint32_t>(b.rowInst);
	M33* js = dptr<M33>(b.jstage);
	V3* fs = dptr<V3>(b.fstage);
	StepArgs sa = step_args(b, h);
	LAUNCH(k_newton_prep, nblk(b.n), sa);
	const bool precise = b.solver != B3G_SOLVER_PCG && b.pb64.p;
	if (b.tet.m) {
		TetArgs t{b.tet.m, dptr<uint32_t>(b.tet.p), dptr<float>(b.tet.rest), dptr<float4>(b.tet.rec), dptr<float4>(b.tet.q), dptr<float>(b.tet.kd), dptr<V3>(b.x), dptr<V3>(b.v), fs, b.tet.f, b.tet.fd, st, ri};
		if (precise) LAUNCH(k_tet_forces64, nblk(b.tet.m), t, dptr<double>(b.tet64), dptr<D3>(b.fstage64));
		else LAUNCH(k_tet_forces, nblk(b.tet.m), t);
	}
	if (b.tri.m) {
		TriArgs t{b.tri.m, dptr<uint32_t>(b.tri.p), dptr<float>(b.tri.x2), dptr<float>(b.tri.invS), dptr<float>(b.tri.K0),
		          dptr<float>(b.tri.kd), dptr<V3>(b.x), dptr<V3>(b.v), js, fs, b.tri.jx, b.tri.jv, b.tri.f, b.tri.fd, st, ri};
		LAUNCH(k_tri_forces, nblk(b.tri.m), t);
	}
	if (b.stretch.m) {
		UVArgs u{b.stretch.m, dptr<uint32_t>(b.stretch.p), dptr<float>(b.stretch.c), dptr<V3>(b.x), dptr<V3>(b.v),
		         dptr<float4>(b.stretch.rec), fs, b.stretch.f, st, ri};
		LAUNCH(k_stretch_forces, nblk(b.stretch.m), u);
	}
	if (b.shear.m) {
		UVArgs u{b.shear.m, dptr<uint32_t>(b.shear.p), dptr<float>(b.shear.c), dptr<V3>(b.x), dptr<V3>(b.v),
		         dptr<float4>(b.shear.rec), fs, b.shear.f, st, ri};
		LAUNCH(k_shear_forces, nblk(b.shear.m), u);
	}
	if (b.spring.m) {
		PairArgs p{b.spring.m, dptr<uint32_t>(b.spring.p), dptr<float>(b.spring.c), dptr<V3>(b.x), dptr<V3>(b.v), js, fs,
		           b.spring.jx, b.spring.jv, b.spring.f, st, ri};
		LAUNCH(k_spring_forces, nblk(b.spring.m), p);
	}
	if (b.mouse.m) {
		PairArgs p{b.mouse.m, dptr<uint32_t>(b.mouse.p), dptr<float>(b.mouse.c), dptr<V3>(b.x), dptr<V3>(b.v), js, fs,
		           b.mouse.jx, b.mouse.jv, b.mouse.f, st, ri};
		LAUNCH(k_mouse_forces, nblk(b.mouse.m), p);
	}
	if (b.trackActions && b.act.p) {
		auto acc = [&](uint32_t m, uint32_t np, const DBuf& p, const DBuf& cf, uint32_t f, uint32_t base,
		               std::initializer_list<std::pair<uint32_t, uint32_t>> terms) -> b3g_status {
			if (!m) return B3G_OK;
			ActArgs aa{};
			aa.m = m; aa.np = np; aa.nterm = (uint32_t)terms.size(); aa.f = f; aa.fstage = fs;
			aa.c = dptr<float>(cf);
			uint32_t t = 0;
			for (const auto& tc : terms) { aa.coef[t] = tc.first; aa.slot[t] = tc.second; ++t; }
			aa.act = dptr<V3>(b.act); aa.base = base; aa.p = dptr<uint32_t>(p); aa.st = st; aa.rowInst = ri;
			LAUNCH(k_action_forces, nblk(m), aa);
			return B3G_OK;
		};
		b3g_status s;
		// (coefficient plane, staging plane of vertex 0) per term, in the order the reference adds them
		if ((s = acc(b.stretch.m, 3, b.stretch.p, b.stretch.c, b.stretch.f, b.actStretch, {{12, 0}, {13, 3}, {15, 6}, {16, 9}})) != B3G_OK) return s;
		if ((s = acc(b.shear.m, 3, b.shear.p, b.shear.c, b.shear.f, b.actShear, {{12, 0}, {13, 3}})) != B3G_OK) return s;
		if ((s = acc(b.spring.m, 2, b.spring.p, b.spring.c, b.spring.f, b.actSpring, {{1, 0}, {2, 2}})) != B3G_OK) return s;
		if ((s = acc(b.mouse.m, 4, b.mouse.p, b.mouse.c, b.mouse.f, b.actMouse, {{4, 0}, {5, 4}})) != B3G_OK) return s;
	}
	ContactArgs c = contact_args(b, h);
	if (b.nPairs) LAUNCH(k_contact_forces, nblk(b.nPairs), c);
	AsmArgs a;
	a.n = b.n; a.nnz = b.nnz; a.stride = b.aStride; a.hasContacts = b.nPairs ? 1u : 0u; a.h = h;
	a.slotTiles = dptr<uint32_t>(b.slotTiles); a.slotTileInst = dptr<uint32_t>(b.slotTileInst);
	a.rowTiles = dptr<uint32_t>(b.tiles); a.tileInst = dptr<uint32_t>(b.tileInst);
	a.tetRec = dptr<float>(b.tet.rec);
	a.tetM = b.tet.m; a.tetQ = dptr<float4>(b.tet.q); a.tetKd = dptr<float>(b.tet.kd); a.anyJv = b.anyJv ? 1u : 0u;
	a.stRec = dptr<float4>(b.stretch.rec); a.shRec = dptr<float4>(b.shear.rec);
	a.uvFirst = b.stretch.m ? a.stRec : a.shRec;
	a.slotRow = dptr<uint32_t>(b.slotRow); a.cols = dptr<uint32_t>(b.cols);
	a.jxPtr = dptr<uint32_t>(b.jxPtr); a.jxIdx = dptr<uint32_t>(b.jxIdx);
	a.jvPtr = dptr<uint32_t>(b.jvPtr); a.jvIdx = dptr<uint32_t>(b.jvIdx);
	a.bPtr = dptr<uint32_t>(b.bPtr); a.bSlot = dptr<uint32_t>(b.bSlot);
	a.fPtr = dptr<uint32_t>(b.fPtr); a.fIdx = dptr<uint32_t>(b.fIdx);
	a.diagSlot = dptr<uint32_t>(b.diagSlot);
	a.jstage = js; a.fstage = fs;
	a.type = dptr<uint8_t>(b.type); a.mass = dptr<float>(b.mass); a.damping = dptr<float>(b.damping);
	a.sphereOf = dptr<int32_t>(b.sphereOf);
	a.x0 = dptr<V3>(b.x0); a.v0 = dptr<V3>(b.v0); a.v = dptr<V3>(b.v); a.fe = dptr<V3>(b.fe); a.w = dptr<V3>(b.w);
	a.A = dptr<float>(b.A); a.tB = dptr<V3>(b.tB); a.pb = dptr<V3>(b.pb); a.invM = dptr<V3>(b.invM); a.fi = dptr<V3>(b.fi);
	a.c = c;
	a.st = st;
	a.stw = dptr<SolverState>(b.state);
	a.faultWord = dptr<uint32_t>(b.faultWord);
	a.slotPos = b.sortedA ? dptr<uint32_t>(b.slotPos) : nullptr;
	a.y = dptr<V3>(b.y);
	a.tB64 = precise ? dptr<double>(b.tB64) : nullptr;
	a.fstage64 = dptr<D3>(b.fstage64);
	a.tetF0 = b.tet.f; a.tetF1 = b.tet.f + 4 * b.tet.m;
	a.tetFd0 = b.tet.anyDamping ? b.tet.fd : 0; a.tetFd1 = b.tet.anyDamping ? b.tet.fd + 4 * b.tet.m : 0;
	a.pb64 = dptr<double>(b.pb64);
	a.mass64 = dptr<double>(b.mass64);
	a.force = dptr<V3>(b.force);
	a.gravity = dptr<V3>(b.gravity);
	a.rowInst = dptr<uint32_t>(b.rowInst);
	a.single = b.nInst <= 1 ? 1u : 0u;
	// the contribution kinds of the body's slot lists: tetrahedra, stretch / shear records, staged blocks
	const int kinds = (b.tet.m ? kKindTet : 0) | (b.stretch.m + b.shear.m ? kKindUV : 0) | (b.nJStage ? kKindStaged : 0);
	switch (kinds) {
	case kKindTet: LAUNCH(k_assemble_slots<kKindTet>, b.nSlotTiles, a); break;
	case kKindUV: LAUNCH(k_assemble_slots<kKindUV>, b.nSlotTiles, a); break;
	case kKindTet | kKindUV: LAUNCH(k_assemble_slots<kKindTet | kKindUV>, b.nSlotTiles, a); break;
	default: LAUNCH(k_assemble_slots<kKindAll>, b.nSlotTiles, a); break;
	}
	LAUNCH(k_assemble_rows, b.nTiles, a);
	if (precise) LAUNCH(k_assemble_rows64, b.nTiles, a);
	return B3G_OK;
}

static b3g_status enqueue_finalize(Body& b, const DBuf& partial, int sel, int kind, uint32_t k, int gate, uint32_t gk) {
	Reduce rd{dptr<double>(partial), dptr<float>(b.terms), dptr<SolverState>(b.state), dptr<uint32_t>(b.instRange),
	          dptr<CGLog>(b.cgLog), b.logStride, use_persistent(b) ? dptr<uint32_t>(b.persist) : nullptr,
	          b.nInst > 1 ? dptr<uint32_t>(b.tileGate) : nullptr, b.nInst > 1 ? dptr<uint32_t>(b.blkGate) : nullptr};
	hipLaunchKernelGGL(k_finalize, dim3(std::max(1u, b.nInst)), dim3(kFinBlock), 0, b.stream, rd, sel, kind, k, gate, gk);
	B3G_HIP(hipGetLastError());
	return B3G_OK;
}

#define FINALIZE(partial, sel, kind, k, gate, gk)                                          \
	do {                                                                                   \
		b3g_status fs_ = enqueue_finalize(b, (partial), (sel), (kind), (k), (gate), (gk)); \
		if (fs_ != B3G_OK) return fs_;                                                     \
	} while (0)

// Partials that end PCG iteration k: the true-residual refresh (tiles) on
// every 50th iteration (sparse_solver.cpp:81-90), the vector update (row
// blocks) otherwise.
static inline int delta_range(uint32_t k) { return k % 50 == 0 ? kRangeTiles : kRangeBlocks; }

// Early exit: whether any instance ran PCG iteration k (its log entry is
// active and its Newton loop is), and the largest iteration count so far,
// written with the chunk's sequence number into pinned host memory.
__global__ void __launch_bounds__(kBlock) k_pcg_flag(const SolverState* st, const CGLog* log, uint32_t stride,
                                                     uint32_t nInst, uint32_t k, uint32_t seq, uint32_t* out,
                                                     const uint32_t* persist) {
	uint32_t any = 0u, its = 0u;
	for (uint32_t i = threadIdx.x; i < nInst; i += kBlock) {
		if (persistent(persist, i)) continue;   // the persistent kernel's instances need no launches
		if (st[i].newtonActive && log[(size_t)i * stride + k].active) any = 1u;
		its = max(its, st[i].it);
	}
	__shared__ uint32_t shAny[kBlock], shIts[kBlock];
	shAny[threadIdx.x] = any;
	shIts[threadIdx.x] = its;
	__syncthreads();
	for (int o = kBlock / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) {
			shAny[threadIdx.x] |= shAny[threadIdx.x + o];
			shIts[threadIdx.x] = max(shIts[threadIdx.x], shIts[threadIdx.x + o]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		out[1] = shIts[0];
		__threadfence_system();
		out[0] = (seq << 1) | shAny[0];
		__threadfence_system();
	}
}

b3g_status ensure_exit_flags(Body& b) {
	if (b.hostFlag) return B3G_OK;
	void* p = nullptr;
	B3G_HIP(hipHostMalloc(&p, 512, hipHostMallocDefault));   // two slots 256 bytes apart
	std::memset(p, 0, 512);
	b.hostFlag = static_cast<uint32_t*>(p);
	for (hipEvent_t& e : b.flagEvent) B3G_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	return B3G_OK;
}

void free_persistent(Body& b) {
	if (b.stream2) (void)hipStreamSynchronize(b.stream2);
	if (b.forkEvent) (void)hipEventDestroy(b.forkEvent);
	if (b.joinEvent) (void)hipEventDestroy(b.joinEvent);
	if (b.stream2) (void)hipStreamDestroy(b.stream2);
	b.forkEvent = b.joinEvent = nullptr;
	b.stream2 = nullptr;
}

void free_exit_flags(Body& b) {
	for (hipEvent_t& e : b.flagEvent) {
		if (e) (void)hipEventDestroy(e);
		e = nullptr;
	}
	if (b.hostFlag) (void)hipHostFree(b.hostFlag);
	b.hostFlag = nullptr;
}

// Waits for the flag of chunk `seq` in slot `slot`: spins on the pinned word
// (its sequence number says the kernel has written it) and checks the event
// every few thousand reads, so that a failed stream ends the wait.
static b3g_status wait_flag(Body& b, uint32_t slot, uint32_t seq, bool* any, uint32_t* its) {
	volatile uint32_t* f = b.hostFlag + 64 * slot;
	for (uint32_t spin = 0;; ++spin) {
		const uint32_t v = f[0];
		if ((v >> 1) == (seq & 0x7FFFFFFFu)) {
			*any = (v & 1u) != 0u;
			*its = f[1];
			return B3G_OK;
		}
		if ((spin & 4095u) == 4095u) {
			const hipError_t e = hipEventQuery(b.flagEvent[slot]);
			if (e == hipSuccess) {   // complete: the word is written
				const uint32_t w = f[0];
				if ((w >> 1) != (seq & 0x7FFFFFFFu)) {
					set_error("PCG early-exit flag was not written");
					return B3G_HIP_ERROR;
				}
				*any = (w & 1u) != 0u;
				*its = f[1];
				return B3G_OK;
			}
			if (e != hipErrorNotReady) {
				set_error(std::string("PCG early-exit wait: ") + hipGetErrorString(e));
				return B3G_HIP_ERROR;
			}
		}
	}
}

// PCG iterations per chunk after the first. Two rather than four: the chunk
// in flight when a flag reports convergence runs as no-op launches (64^3 tet
// body alone 1.397 -> 1.382 ms per step).
constexpr uint32_t kExitChunk = 2;

// Iterations 0 .. maxIt - 1 of the PCG (sparse_solver.cpp:75-103) after its
// Init reduction, then the reduction that ends the last one: the reference's
// fixed launch sequence; iterations after convergence return at once on their
// log entry. With early exit, the sequence is enqueued in chunks and stops
// once a chunk's last iteration ran in no instance: everything after it
// would have returned at once, so the results are the same.
static b3g_status enqueue_pcg_iterations(Body& b, const CGArgs& a, uint32_t maxIt) {
	const bool fused = a.fused != 0;
	const uint32_t rb = b.nRowBlk;
	const bool poll = b.earlyExit && !b.capturing && b.hostFlag;
	struct Pending { uint32_t slot, seq; };
	Pending pend[2];
	uint32_t npend = 0;
	// first chunk: one short of the previous solve's count, so its check lands near convergence
	uint32_t chunk = poll ? std::max(kExitChunk, b.lastConverged > 1 ? b.lastConverged - 1 : 0u) : maxIt;
	uint32_t k = 0, its = 0;
	bool stopped = false;
	while (k < maxIt && !stopped) {
		const uint32_t e = std::min(maxIt, k + chunk);
		for (; k < e; ++k) {
			CGArgs ak = a;
			V3* d0 = dptr<V3>(b.d0);
			V3* d1 = dptr<V3>(b.d1);
			ak.dold = (k & 1) ? d1 : d0;
			ak.dnew = (k & 1) ? d0 : d1;
			const int sel = k > 0 ? delta_range(k - 1) : kRangeTiles;
			if (!fused && k > 0) FINALIZE(b.partial, sel, kFinDelta, k, kGateIter, k - 1);
			LAUNCH_TILES(k_cg_spmv_dq, ak, b.nTiles, ak, k, sel, 0);
			if (!fused) FINALIZE(b.partialQ, kRangeTiles, kFinAlpha, k, kGateIter, k);
			if (k % 50 == 0) {
				LAUNCH_ITEMS(k_cg_update_x, rb, ak, k);
				LAUNCH_TILES(k_cg_residual, ak, b.nTiles, ak, (int)kFinDelta, k);
			} else {
				const b3g_status us = launch_update(b, ak, k, 0);
				if (us != B3G_OK) return us;
			}
		}
		chunk = kExitChunk;
		if (!poll || k == maxIt) continue;
		b.flagSeq = (b.flagSeq + 1u) & 0x7FFFFFFFu;
		if (b.flagSeq == 0u) b.flagSeq = 1u;   // the zeroed words never match a chunk
		const uint32_t seq = b.flagSeq, slot = seq & 1u;
		hipLaunchKernelGGL(k_pcg_flag, dim3(1), dim3(kBlock), 0, b.stream, a.st, a.log, a.logStride, std::max(1u, b.nInst),
		                   k - 1, seq, b.hostFlag + 64 * slot, a.persist);
		B3G_HIP(hipGetLastError());
		B3G_HIP(hipEventRecord(b.flagEvent[slot], b.stream));
		pend[npend++] = {slot, seq};
		if (npend == 2) {   // the older chunk's flag, while the newer one runs
			bool any = true;
			b3g_status s = wait_flag(b, pend[0].slot, pend[0].seq, &any, &its);
			if (s != B3G_OK) return s;
			pend[0] = pend[1];
			npend = 1;
			stopped = !any;
		}
	}
	// the next solve's first chunk: this one's count, or the cap when the sequence ran to the end
	if (poll) b.lastConverged = stopped ? its : maxIt;
	if (maxIt > 0 && !stopped) FINALIZE(b.partial, delta_range(maxIt - 1), kFinDelta, maxIt, kGateIter, maxIt - 1);
	return B3G_OK;
}

// Jacobi PCG on the assembled pA / pb with the warm start px.
b3g_status enqueue_solve_cg(Body& b, uint32_t maxIt, float eps) {
	(void)eps;
	b3g_status s = ensure_cg_log(b, maxIt + 1);
	if (s != B3G_OK) return s;
	CGArgs a = cg_args(b);
	// the solve's start for every instance, the persistent ones included
	CGArgs start = a;
	start.persist = nullptr;
	LAUNCH_TILES(k_cg_residual, start, b.nTiles, start, (int)kFinInit, 0u);
	FINALIZE(b.partial, kRangeTiles, kFinInit, 0, kGateNewton, 0);
	if (!a.persist) return enqueue_pcg_iterations(b, a, maxIt);
	// the persistent instances' iterations on the second stream, beside the launch sequence of the others
	B3G_HIP(hipEventRecord(b.forkEvent, b.stream));
	B3G_HIP(hipStreamWaitEvent(b.stream2, b.forkEvent, 0));
	const uint32_t g = b.sched.persistGroups ? std::min(b.sched.persistGroups, b.nPersist) : b.nPersist;
	const uint32_t* plist = dptr<uint32_t>(b.plist);
	V3* d0 = dptr<V3>(b.d0);
	V3* d1 = dptr<V3>(b.d1);
	double* pt = dptr<double>(b.partialP);
	double* pk = pt + b.nTiles;
	if (tile_storage(a) == 1)   // the plain solver: FP32 blocks, column-sorted or CSR storage
		hipLaunchKernelGGL(k_pcg_persistent<1>, dim3(g), dim3(kBlock), 0, b.stream2, a, plist, b.nPersist, d0, d1, pt, pk);
	else
		hipLaunchKernelGGL(k_pcg_persistent<0>, dim3(g), dim3(kBlock), 0, b.stream2, a, plist, b.nPersist, d0, d1, pt, pk);
	B3G_HIP(hipGetLastError());
	B3G_HIP(hipEventRecord(b.joinEvent, b.stream2));
	// every instance persistent: the launch sequence would only skip them
	const b3g_status s2 = b.nPersist < std::max(1u, b.nInst) ? enqueue_pcg_iterations(b, a, maxIt) : B3G_OK;
	B3G_HIP(hipStreamWaitEvent(b.stream, b.joinEvent, 0));   // joined even when the launch sequence failed
	return s2;
}

static RefineArgs refine_args(Body& b) {
	RefineArgs r;
	r.n = b.n;
	r.nnz = b.nnz;
	r.stride = b.aStride;
	r.tiles = dptr<uint32_t>(b.tiles);
	r.tileInst = dptr<uint32_t>(b.tileInst);
	r.rowptr = dptr<uint32_t>(b.rowptr);
	r.cols = dptr<uint32_t>(b.cols);
	r.slotRow = dptr<uint32_t>(b.slotRow);
	r.rowInst = dptr<uint32_t>(b.rowInst);
	r.A = dptr<float>(b.A);
	r.pb = dptr<V3>(b.pb);
	r.invM = dptr<V3>(b.invM);
	r.x64 = dptr<double>(b.x64);
	r.pb64 = dptr<double>(b.pb64);
	r.px = dptr<V3>(b.px);
	r.dx = dptr<V3>(b.dx);
	r.rhs = dptr<V3>(b.rhsF);
	r.r = dptr<V3>(b.r);
	r.s = dptr<V3>(b.s);
	r.partialR = dptr<double>(b.partialR);
	r.partial = dptr<double>(b.partial);
	r.st = dptr<SolverState>(b.state);
	r.AbP16 = dptr<uint16_t>(b.AbP);
	r.pairBase = dptr<uint32_t>(b.pairBase);
	r.pairStride = b.pairStride;
	r.diagF = dptr<float>(b.diagF);
	r.code = b.sortedA ? dptr<uint32_t>(b.code) : nullptr;
	r.tileBase = b.sortedA ? dptr<uint32_t>(b.tileBase) : nullptr;
	r.storeSlot = b.sortedA ? dptr<uint32_t>(b.storeSlot) : nullptr;
	return r;
}

// Buffers of the precision form (allocated on first use).
static b3g_status ensure_refine(Body& b) {
	b3g_status s;
	auto need = [](const DBuf& d, size_t bytes) { return !d.p || d.bytes < bytes; };
	if (need(b.x64, (size_t)b.n * 24) && (s = dalloc(b, b.x64, (size_t)b.n * 24)) != B3G_OK) return s;
	if (need(b.dx, (size_t)b.n * sizeof(V3)) && (s = dalloc(b, b.dx, (size_t)b.n * sizeof(V3))) != B3G_OK) return s;
	if (need(b.rhsF, (size_t)b.n * sizeof(V3)) && (s = dalloc(b, b.rhsF, (size_t)b.n * sizeof(V3))) != B3G_OK) return s;
	if (need(b.partialR, (size_t)b.partialCount * 8) && (s = dalloc(b, b.partialR, (size_t)b.partialCount * 8)) != B3G_OK)
		return s;
	if (b.solver == B3G_SOLVER_REFINED_BF16) {
		const size_t pairBytes = (size_t)9 * b.pairStride * 4;
		if (need(b.AbP, pairBytes) && (s = dalloc(b, b.AbP, pairBytes)) != B3G_OK) return s;
		if (need(b.diagF, (size_t)b.n * 36) && (s = dalloc(b, b.diagF, (size_t)b.n * 36)) != B3G_OK) return s;
	}
	// FP64 right-hand side: tetrahedron forces and (h dfdx) w in double
	if (need(b.pb64, (size_t)b.n * 24) && (s = dalloc(b, b.pb64, (size_t)b.n * 24)) != B3G_OK) return s;
	if (need(b.tB64, (size_t)b.nnz * 24) && (s = dalloc(b, b.tB64, (size_t)b.nnz * 24)) != B3G_OK) return s;
	if (need(b.fstage64, (size_t)std::max(1u, b.nFStage) * sizeof(D3)) &&
	    (s = dalloc(b, b.fstage64, (size_t)std::max(1u, b.nFStage) * sizeof(D3))) != B3G_OK)
		return s;
	if (b.tet.m && !b.tet64.p) {
		if ((s = dalloc(b, b.tet64, (size_t)b.tet.m * 12 * sizeof(double))) != B3G_OK) return s;
		LAUNCH(k_tet_setup64, nblk(b.tet.m), b.tet.m, dptr<float>(b.tet.rest), dptr<float>(b.tet.param),
		       dptr<double>(b.tet64));
	}
	if (!b.mass64.p) {
		if ((s = dupload(b, b.mass64, b.hostMass64.data(), b.hostMass64.size() * sizeof(double))) != B3G_OK) return s;
	}
	return B3G_OK;
}

// Precision form of one Newton iteration's solve: sweeps of an FP64 residual
// and an inner PCG of at most innerIt iterations on pA dx = residual (FP32 or
// BF16-off-diagonal blocks), x += dx in FP64; then px = x.
static b3g_status enqueue_solve_refined(Body& b, uint32_t sweeps) {
	b3g_status s = ensure_cg_log(b, b.innerIt + 1);
	if (s != B3G_OK) return s;
	RefineArgs ra = refine_args(b);
	CGArgs a = cg_args(b);
	a.pb = dptr<V3>(b.rhsF);
	a.px = dptr<V3>(b.dx);
	if (b.solver == B3G_SOLVER_REFINED_BF16) LAUNCH(k_split_bf16, b.nTiles, ra);
	LAUNCH(k_x64_init, nblk(b.n), ra);
	for (uint32_t sw = 0; sw < sweeps; ++sw) {
		LAUNCH(k_resid64, b.nTiles, ra, sw == 0 ? 1 : 0);
		FINALIZE(b.partialR, kRangeTiles, sw == 0 ? kFinRefine0 : kFinRefine, 0, kGateNewton, 0);
		FINALIZE(b.partial, kRangeTiles, kFinInit, 0, kGateNewton, 0);
		s = enqueue_pcg_iterations(b, a, b.innerIt);
		if (s != B3G_OK) return s;
		LAUNCH(k_x64_add, nblk(b.n), ra);
	}
	LAUNCH(k_resid64, b.nTiles, ra, 0);   // the residual after the last sweep (skipped once converged)
	FINALIZE(b.partialR, kRangeTiles, kFinRefine, 0, kGateNewton, 0);
	LAUNCH(k_x64_finish, nblk(b.n), ra);
	return B3G_OK;
}

b3g_status enqueue_spmv(Body& b, const DBuf& xin, DBuf& yout) {
	CGArgs a = cg_args(b);
	if (b.nTiles) {
		const dim3 gr(b.nTiles);
		switch (tile_storage(a)) {
		case 0: hipLaunchKernelGGL(k_spmv<0>, gr, dim3(kBlock), 0, b.stream, a, dptr<V3>(xin), dptr<V3>(yout)); break;
		case 1: hipLaunchKernelGGL(k_spmv<1>, gr, dim3(kBlock), 0, b.stream, a, dptr<V3>(xin), dptr<V3>(yout)); break;
		case 2: hipLaunchKernelGGL(k_spmv<2>, gr, dim3(kBlock), 0, b.stream, a, dptr<V3>(xin), dptr<V3>(yout)); break;
		default: hipLaunchKernelGGL(k_spmv<3>, gr, dim3(kBlock), 0, b.stream, a, dptr<V3>(xin), dptr<V3>(yout)); break;
		}
	}
	B3G_HIP(hipGetLastError());
	return B3G_OK;
}

// Back-to-back launches of the PCG's SpMV kernel (an iteration's full work,
// on the current pA and the CG scratch vectors) between one event pair.
b3g_status time_pcg_spmv(Body& b, uint32_t launches, double* mean_ms) {
	b3g_status s0 = ensure_cg_log(b, 2);   // the kernel reads entry 1's beta (the log exists only after a step)
	if (s0 != B3G_OK) return s0;
	CGArgs a = cg_args(b);
	hipEvent_t e0, e1;
	B3G_HIP(hipEventCreate(&e0));
	B3G_HIP(hipEventCreate(&e1));
	b3g_status st = B3G_OK;
	float ms = 0.0f;
	do {
		if (hipEventRecord(e0, b.stream) != hipSuccess) { st = B3G_HIP_ERROR; break; }
		for (uint32_t k = 0; k < launches && st == B3G_OK; ++k) {
			CGArgs ak = a;
			ak.dold = (k & 1) ? dptr<V3>(b.d1) : dptr<V3>(b.d0);
			ak.dnew = (k & 1) ? dptr<V3>(b.d0) : dptr<V3>(b.d1);
			TILE_KERNEL(k_cg_spmv_dq, ak, b.nTiles, ak, 1u, (int)kRangeBlocks, 1);
			if (hipGetLastError() != hipSuccess) st = B3G_HIP_ERROR;
		}
		if (st != B3G_OK) break;
		if (hipEventRecord(e1, b.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
		    hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
			st = B3G_HIP_ERROR;
	} while (0);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (st != B3G_OK) {
		set_error("SpMV timing pass failed");
		return st;
	}
	*mean_ms = launches ? (double)ms / launches : 0.0;
	return B3G_OK;
}

// Streaming reads over a buffer larger than the L2s and the MALL together:
// afterwards a kernel's loads of anything else come from HBM. Reads, not
// writes: a write stream would leave the memory-side cache full of dirty
// lines, whose write-back would then run during the measured kernel.
__global__ void __launch_bounds__(kBlock) k_evict(const uint4* p, size_t count, uint32_t v, uint32_t* sink) {
	uint32_t acc = 0u;
	for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) {
		const uint4 w = p[i];
		acc ^= w.x ^ w.y ^ w.z ^ w.w;
	}
	if (acc == v + 0x9E3779B9u) sink[threadIdx.x] = acc;   // keeps the loads; the buffer is zero, so never taken
}

// The SpMV pass from HBM: `launches` x (eviction, SpMV) between one event
// pair, then `launches` evictions alone between another; the difference per
// launch is the SpMV with one launch gap.
b3g_status time_pcg_spmv_cold(Body& b, uint32_t launches, size_t evictBytes, double* mean_ms) {
	b3g_status s0 = ensure_cg_log(b, 2);
	if (s0 != B3G_OK) return s0;
	void* buf = nullptr;
	if (hipMalloc(&buf, evictBytes) != hipSuccess) {
		(void)hipGetLastError();
		set_error("no device memory for the eviction buffer");
		return B3G_OUT_OF_MEMORY;
	}
	const size_t count = (evictBytes - 4096) / sizeof(uint4);
	uint32_t* sink = reinterpret_cast<uint32_t*>(static_cast<char*>(buf) + count * sizeof(uint4));
	if (hipMemsetAsync(buf, 0, evictBytes, b.stream) != hipSuccess) {
		(void)hipGetLastError();
		(void)hipFree(buf);
		set_error("clearing the eviction buffer failed");
		return B3G_HIP_ERROR;
	}
	CGArgs a = cg_args(b);
	hipEvent_t e[4];
	for (hipEvent_t& x : e) x = nullptr;
	b3g_status st = B3G_OK;
	float both = 0.0f, alone = 0.0f;
	do {
		bool ok = true;
		for (hipEvent_t& x : e) ok = ok && hipEventCreate(&x) == hipSuccess;
		if (!ok) { st = B3G_HIP_ERROR; break; }
		// the first eviction kernel and SpMV launch warm up the code objects
		hipLaunchKernelGGL(k_evict, dim3(4096), dim3(kBlock), 0, b.stream, (const uint4*)buf, count, 0u, sink);
		ok = hipEventRecord(e[0], b.stream) == hipSuccess;
		for (uint32_t k = 0; k < launches && ok; ++k) {
			hipLaunchKernelGGL(k_evict, dim3(4096), dim3(kBlock), 0, b.stream, (const uint4*)buf, count, k, sink);
			CGArgs ak = a;
			ak.dold = (k & 1) ? dptr<V3>(b.d1) : dptr<V3>(b.d0);
			ak.dnew = (k & 1) ? dptr<V3>(b.d0) : dptr<V3>(b.d1);
			TILE_KERNEL(k_cg_spmv_dq, ak, b.nTiles, ak, 1u, (int)kRangeBlocks, 1);
			ok = hipGetLastError() == hipSuccess;
		}
		ok = ok && hipEventRecord(e[1], b.stream) == hipSuccess && hipEventRecord(e[2], b.stream) == hipSuccess;
		for (uint32_t k = 0; k < launches && ok; ++k) {
			hipLaunchKernelGGL(k_evict, dim3(4096), dim3(kBlock), 0, b.stream, (const uint4*)buf, count, k, sink);
			ok = hipGetLastError() == hipSuccess;
		}
		ok = ok && hipEventRecord(e[3], b.stream) == hipSuccess && hipEventSynchronize(e[3]) == hipSuccess &&
		     hipEventElapsedTime(&both, e[0], e[1]) == hipSuccess && hipEventElapsedTime(&alone, e[2], e[3]) == hipSuccess;
		if (!ok) st = B3G_HIP_ERROR;
	} while (0);
	for (hipEvent_t& x : e)
		if (x) (void)hipEventDestroy(x);
	(void)hipStreamSynchronize(b.stream);
	(void)hipFree(buf);
	if (st != B3G_OK) {
		set_error("cold SpMV timing pass failed");
		return st;
	}
	*mean_ms = ((double)both - (double)alone) / launches;
	return B3G_OK;
}

b3g_status time_pcg_iteration(Body& b, uint32_t reps, double* ms) {
	b3g_status s0 = ensure_cg_log(b, 3);
	if (s0 != B3G_OK) return s0;
	CGArgs a = cg_args(b);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	B3G_HIP(hipEventCreate(&e0));
	B3G_HIP(hipEventCreate(&e1));
	b3g_status st = B3G_OK;
	float el = 0.0f;
	auto body = [&]() -> b3g_status {
		B3G_HIP(hipEventRecord(e0, b.stream));
		for (uint32_t r = 0; r < reps; ++r) {
			CGArgs ak = a;
			ak.dold = (r & 1) ? dptr<V3>(b.d1) : dptr<V3>(b.d0);
			ak.dnew = (r & 1) ? dptr<V3>(b.d0) : dptr<V3>(b.d1);
			LAUNCH_TILES(k_cg_spmv_dq, ak, b.nTiles, ak, 1u, (int)kRangeBlocks, 1);
			if (!a.fused) FINALIZE(b.partialQ, kRangeTiles, kFinAlpha, 1, kGateIter, 1);
			{
				b3g_status us = launch_update(b, ak, 1u, 1);
				if (us != B3G_OK) return us;
			}
			if (!a.fused) FINALIZE(b.partial, kRangeBlocks, kFinDelta, 2, kGateIter, 1);
		}
		B3G_HIP(hipEventRecord(e1, b.stream));
		B3G_HIP(hipEventSynchronize(e1));
		B3G_HIP(hipEventElapsedTime(&el, e0, e1));
		return B3G_OK;
	};
	st = body();
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (st != B3G_OK) return st;
	*ms = reps ? (double)el / reps : 0.0;
	return B3G_OK;
}

// Candidate schedules: SpMV tile shapes (blocks, rows), vector row blocks and
// the place of the FP64 sums, then workgroup caps, each timed by `timer` (or as PCG iterations on
// the body's own system). The default is timed first; a candidate faster
// than it is timed again against it, alternately, and kept only if it is
// faster by more than 1% both times. Any captured step is dropped with each change of tiles.
b3g_status autotune(Body& b, const std::function<b3g_status(double*)>& timer, double* tunedMs, double* defaultMs) {
	const uint32_t shapes[][2] = {{1024, 256}, {512, 256}, {1024, 128}, {512, 128}, {256, 64}};
	const uint32_t blocks[] = {1024, 512, 256};
	auto timeIt = [&](double* ms) -> b3g_status {
		drop_graph(b);
		if (timer) return timer(ms);
		double probe = 0.0;
		b3g_status s = time_pcg_iteration(b, 2, &probe);   // warm-up and size estimate
		if (s != B3G_OK) return s;
		const uint32_t reps = (uint32_t)std::min(200.0, std::max(3.0, 20.0 / std::max(probe, 1e-3)));
		return time_pcg_iteration(b, reps, ms);
	};
	auto apply = [&](const Schedule& sc) -> b3g_status {
		b.sched = sc;
		return build_tiles(b, true, true);
	};
	b3g_status s;
	const Schedule dflt;
	if ((s = apply(dflt)) != B3G_OK) return s;
	double dfltMs = 0.0;
	if ((s = timeIt(&dfltMs)) != B3G_OK) return s;
	Schedule best = dflt;
	double bestMs = dfltMs;
	for (const auto& sh : shapes) {
		b.sched.tileBlocks = sh[0];
		b.sched.tileRows = sh[1];
		b.sched.rowBlockRows = 1024;
		if ((s = build_tiles(b, true, true)) != B3G_OK) return s;
		for (uint32_t rb : blocks) {
			b.sched.rowBlockRows = rb;
			if ((s = build_tiles(b, true, false)) != B3G_OK) return s;
			for (int fused = 0; fused < 2; ++fused) {
				b.sched.fused = fused;
				double ms = 0.0;
				if ((s = timeIt(&ms)) != B3G_OK) return s;
				if (ms < bestMs) { bestMs = ms; best = b.sched; }
			}
		}
	}
	// the workgroup cap for the best shape (one per item, or 1024 / 2048 / 4096)
	for (uint32_t cap : {0xFFFFFFFFu, 1024u, 2048u, 4096u}) {
		Schedule sc = best;
		sc.maxGroups = cap == 0xFFFFFFFFu ? std::max(1u, std::max(b.nTiles, b.nRowBlk)) : cap;
		if (sc.maxGroups < 256) continue;
		if ((s = apply(sc)) != B3G_OK) return s;
		double ms = 0.0;
		if ((s = timeIt(&ms)) != B3G_OK) return s;
		if (ms < bestMs) { bestMs = ms; best = sc; }
	}
	// the persistent PCG for a batch's smaller instances; whole steps only (the
	// PCG-iteration timing before any step does not run it)
	if (timer && b.nInst > 1 && b.solver == B3G_SOLVER_PCG && b.reduction != B3G_REDUCE_SEQUENTIAL) {
		for (uint32_t pt : {4u, 16u, 64u}) {
			Schedule sc = best;
			sc.persistTiles = pt;
			sc.persistGroups = 0;
			if ((s = apply(sc)) != B3G_OK) return s;
			double ms = 0.0;
			if ((s = timeIt(&ms)) != B3G_OK) return s;
			if (ms < bestMs) { bestMs = ms; best = sc; }
		}
	}
	// the confirmation's means are reported (a search minimum is biased low)
	bool keep = bestMs < dfltMs;
	if (keep) {
		double dsum = 0.0, csum = 0.0;
		for (int round = 0; round < 2 && keep; ++round) {
			double d = 0.0, c = 0.0;
			if ((s = apply(dflt)) != B3G_OK || (s = timeIt(&d)) != B3G_OK) return s;
			if ((s = apply(best)) != B3G_OK || (s = timeIt(&c)) != B3G_OK) return s;
			keep = c < 0.99 * d;   // by more than the timings' noise
			dsum += d;
			csum += c;
		}
		if (keep) {
			bestMs = csum / 2;
			dfltMs = dsum / 2;
		}
	}
	if (!keep) bestMs = dfltMs;
	if ((s = apply(keep ? best : dflt)) != B3G_OK) return s;
	drop_graph(b);
	*tunedMs = bestMs;
	*defaultMs = dfltMs;
	return B3G_OK;
}

// Everything a step allocates or sets up once, so that the step itself only
// enqueues kernels (and can be captured into a graph).
b3g_status prepare_step(Body& b, uint32_t subIters) {
	if (b.solver != B3G_SOLVER_PCG) {
		b3g_status s = ensure_refine(b);
		if (s != B3G_OK) return s;
		return ensure_cg_log(b, b.innerIt + 1);
	}
	return ensure_cg_log(b, subIters + 1);
}

b3g_status ensure_actions(Body& b) {
	if (!b.trackActions) return B3G_OK;
	b.actStretch = 0;
	b.actShear = b.actStretch + 3 * b.stretch.m;
	b.actSpring = b.actShear + 3 * b.shear.m;
	b.actMouse = b.actSpring + 2 * b.spring.m;
	b.actCount = b.actMouse + 4 * b.mouse.m;
	if (b.act.p && b.act.bytes >= (size_t)std::max(1u, b.actCount) * sizeof(V3)) return B3G_OK;
	b3g_status s = dalloc(b, b.act, (size_t)std::max(1u, b.actCount) * sizeof(V3));
	if (s != B3G_OK) return s;
	LAUNCH(k_clear_v3, (uint32_t)(((size_t)std::max(1u, b.actCount) + kBlock - 1) / kBlock), (size_t)std::max(1u, b.actCount),
	       dptr<V3>(b.act));
	return B3G_OK;
}

void drop_graph(Body& b) {
	if (b.graphExec) (void)hipGraphExecDestroy(b.graphExec);
	b.graphExec = nullptr;
	b.graphKey.clear();
}

b3g_status enqueue_step(Body& b, float dt, uint32_t iters, uint32_t subIters, float tol, float subTol,
                        bool assembleOnly) {
	SolverState* st = dptr<SolverState>(b.state);
	b3g_status s = enqueue_update_contacts(b);
	if (s != B3G_OK) return s;
	// b3Force::ClearForces runs on every step, before the dt > 0 test (body.cpp:563-574)
	if (b.trackActions && b.act.p && !assembleOnly)
		LAUNCH(k_clear_v3, (uint32_t)(((size_t)b.actCount + kBlock - 1) / kBlock), (size_t)b.actCount, dptr<V3>(b.act));
	if (dt > 0.0f) {
		StepArgs sa = step_args(b, dt);
		LAUNCH(k_step_begin, nblk(b.n), sa);
		const bool refined = b.solver != B3G_SOLVER_PCG && !assembleOnly;
		const uint32_t sweeps = refined ? std::max(1u, (subIters + b.innerIt - 1) / std::max(1u, b.innerIt)) : 0u;
		if (refined) {
			s = ensure_refine(b);
			if (s != B3G_OK) return s;
		}
		// refined: the PCG runs innerIt iterations at innerTol per sweep, subTol is the sweeps' target
		LAUNCH(k_state_begin, nblk(b.nInst), st, b.nInst, assembleOnly ? 1u : iters, refined ? b.innerIt : subIters, tol,
		       refined ? b.innerTol : subTol, b.reduction == B3G_REDUCE_SEQUENTIAL ? 1u : 0u, refined ? 1u : 0u, sweeps,
		       subTol);
		for (uint32_t it = 0; it < (assembleOnly ? 1u : iters); ++it) {
			s = enqueue_linearise(b, dt, it == 0);
			if (s != B3G_OK) return s;
			if (assembleOnly) return B3G_OK;
			s = refined ? enqueue_solve_refined(b, sweeps) : enqueue_solve_cg(b, subIters, subTol);
			if (s != B3G_OK) return s;
			LAUNCH(k_newton_update, b.nRowBlk, sa, dptr<double>(b.partial), dptr<float>(b.terms));
			FINALIZE(b.partial, kRangeBlocks, kFinNewton, 0, kGateNewton, 0);
		}
		if (b.nPairs) {
			ContactArgs c = contact_args(b, dt);
			LAUNCH(k_friction, nblk(b.nSpheres), c);
		}
	}
	LAUNCH(k_clear_external, nblk(b.n), b.n, dptr<V3>(b.force), dptr<V3>(b.translation));
	return enqueue_find_contacts(b);
}

// ---------------------------------------------------------------------------
// Operator-level backward Euler (b3SparseSolveBE with a host force model,
// sparse_force_solver.cpp:104-232). The symbolic phase on the host gives every
// slot of pA (the reference's row order of A) its dfdx / dfdv entry; the
// kernels below form the values operation for operation.
// ---------------------------------------------------------------------------

// Solver state for a solve: Newton and PCG limits and tolerances, the running
// min / max of the PCG counts the caller passed in.
__global__ void k_be_state(SolverState* st, uint32_t maxIt, float eps2, uint32_t maxSub, float subEps2, uint32_t seq,
                           uint32_t minSub, uint32_t maxSubStat) {
	uint32_t* words = reinterpret_cast<uint32_t*>(st);
	for (size_t k = 0; k < sizeof(SolverState) / 4; ++k) words[k] = 0;
	st->maxIt = maxSub;
	st->eps2_sub = subEps2;
	st->eps2 = eps2;
	st->newtonMax = maxIt;
	st->newtonActive = maxIt > 0 ? 1u : 0u;
	st->seq = seq;
	st->statMinSub = minSub;
	st->statMaxSub = maxSubStat;
}

// w = x0 - x + h v + y, the vector h dfdx multiplies in b
// (sparse_force_solver.cpp:169: VAdd(VAdd(VSub(x0, x), h v), y)).
__global__ void __launch_bounds__(kBlock) k_be_w(BEArgs a) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n) return;
	a.w[i] = ((a.x0[i] - a.x[i]) + a.h * a.v[i]) + a.y[i];
}

// One thread per slot of pA (row i, column j):
//   A_ij = ((j == i ? M_i : 0) - h dfdv_ij) - (h h) dfdx_ij, each term only
//   where the reference's lists hold it (b3Sub(Diag, Sparse) then
//   b3Sub(Sparse, Sparse), sparse_mat33.h:458-492; b3Mul(s, B) leaves B's
//   copy empty for s == 0, sparse_mat33.h:518-521);
//   pA_ij = (S_i A_ij) S_j^T, plus I minus S_i on the diagonal
//   (sparse_mat33.h:535-576, 464-474).
// A is kept per slot for the product A z.
__global__ void __launch_bounds__(kBlock) k_be_slots(BEArgs a) {
	const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= a.nnz) return;
	const uint32_t i = a.slotRow[k], j = a.cols[k];
	M33 t = j == i ? a.M[i] : mzero();
	const int32_t kv = a.slotV[k], kx = a.slotX[k];
	if (kv >= 0) t = t - a.h * a.valV[kv];
	if (kx >= 0) t = t - a.hh * a.valX[kx];
	a.Araw[k] = t;
	M33 p = (a.S[i] * t) * transpose(a.S[j]);
	if (j == i) p = (p + mdiag(1.0f)) - a.S[i];
	const uint32_t pos = a.slotPos ? a.slotPos[k] : k;
	const float* pe = &p.x.x;
#pragma unroll
	for (int c = 0; c < 9; ++c) a.A[(size_t)c * a.stride + pos] = pe[c];
}

// One thread per row: b = M (v0 - v) + h (fe + fi) + (h dfdx) w with the
// product over reverse(dfdx row) (the copy Scale makes), pb = S (b - A z)
// (sparse_force_solver.cpp:168-178), and the Jacobi preconditioner from pA's
// diagonal block (sparse_solver.cpp:41-54), whose entries must be > 0.
__global__ void __launch_bounds__(kBlock) k_be_rows(BEArgs a) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= a.n) return;
	const V3 t1 = a.M[i] * (a.v0[i] - a.v[i]);
	const V3 t2 = a.h * (a.fe[i] + a.fi[i]);
	V3 t3 = zero3();
	if (a.useV)
		for (uint32_t k = a.rowptrX[i + 1]; k-- > a.rowptrX[i];) t3 = t3 + (a.h * a.valX[k]) * a.w[a.colsX[k]];
	const V3 bb = (t1 + t2) + t3;
	V3 az = zero3();
	for (uint32_t k = a.rowptr[i]; k < a.rowptr[i + 1]; ++k) az = az + a.Araw[k] * a.z[a.cols[k]];
	a.pb[i] = a.S[i] * (bb - az);
	const uint32_t d = a.slotPos ? a.slotPos[a.diagSlot[i]] : a.diagSlot[i];
	const float axx = a.A[d], ayy = a.A[(size_t)4 * a.stride + d], azz = a.A[(size_t)8 * a.stride + d];
	if (!(axx > 0.0f) || !(ayy > 0.0f) || !(azz > 0.0f)) atomicOr(&a.st->faults, (uint32_t)B3G_FAULT_DIAGONAL);
	a.invM[i] = v3(1.0f / axx, 1.0f / ayy, 1.0f / azz);
}

// dv = py + z; v += dv; x = x0 + h v + y; error = |dv|^2
// (sparse_force_solver.cpp:180-204), one partial per row block.
__global__ void __launch_bounds__(kBlock) k_be_update(BEArgs a, const uint32_t* rowBlk, double* partial, float* terms) {
	if (!a.st->newtonActive) return;
	const uint32_t blk = blockIdx.x, r0 = rowBlk[blk], r1 = rowBlk[blk + 1];
	const bool seq = a.st->seq != 0;
	double part = 0.0;
#pragma unroll
	for (int u = 0; u < kRowsPerThread; ++u) {
		const uint32_t i = r0 + threadIdx.x + u * kBlock;
		if (i >= r1) break;
		const V3 dv = a.px[i] + a.z[i];
		const V3 v = a.v[i] + dv;
		const V3 x = (a.x0[i] + a.h * v) + a.y[i];
		a.v[i] = v;
		a.x[i] = x;
		if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z)))
			atomicOr(&a.st->faults, (uint32_t)B3G_FAULT_NON_FINITE);
		const float term = length_sq(dv);
		if (seq) terms[i] = term;
		part += (double)term;
	}
	block_store(part, partial, blk);
}

b3g_status enqueue_be_begin(Body& b, const BEArgs& a, uint32_t maxIt, float tol, uint32_t maxSub, float subTol,
                            uint32_t minSub, uint32_t maxSubStat) {
	const float eps2 = tol * tol, subEps2 = subTol * subTol;
	hipLaunchKernelGGL(k_be_state, dim3(1), dim3(1), 0, b.stream, a.st, maxIt, eps2, maxSub, subEps2,
	                   b.reduction == B3G_REDUCE_SEQUENTIAL ? 1u : 0u, minSub, maxSubStat);
	B3G_HIP(hipGetLastError());
	LAUNCH(k_clear_v3, nblk(b.n), (size_t)b.n, dptr<V3>(b.px));   // py = 0 (sparse_force_solver.cpp:133-134)
	return B3G_OK;
}

b3g_status enqueue_be_iteration(Body& b, const BEArgs& a, uint32_t maxSub, float subTol) {
	LAUNCH(k_be_w, nblk(b.n), a);
	LAUNCH(k_be_slots, nblk(b.nnz), a);
	LAUNCH(k_be_rows, nblk(b.n), a);
	b3g_status s = enqueue_solve_cg(b, maxSub, subTol);
	if (s != B3G_OK) return s;
	LAUNCH(k_be_update, b.nRowBlk, a, dptr<uint32_t>(b.rowBlk), dptr<double>(b.partial), dptr<float>(b.terms));
	FINALIZE(b.partial, kRangeBlocks, kFinNewton, 0, kGateNewton, 0);
	return B3G_OK;
}

} // namespace b3g
