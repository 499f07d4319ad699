// b3g_solve_cg_static: the operator-level Jacobi PCG (b3SparseSolveCG,
// sparse_solver.cpp:28-109) on device arrays with no allocation. Its PCG
// scratch (preconditioner, residual, directions, reduction partials, the
// per-iteration log, the solver state and the row decompositions) lives in
// the file-scope device array `scratch` below; each GPU holds one copy of it
// from the library's load until the process ends.
//
// Host code never passes `scratch` itself to a kernel or a copy: in host code
// that name denotes a host-side stand-in, not the device copy. The device
// address is looked up once per device with hipGetSymbolAddress (with that
// device current) and kept in g_scratch[device].
#include "b3g_internal.h"

#include <algorithm>
#include <mutex>

namespace b3g {

namespace {

constexpr uint32_t kMaxRows = B3G_CG_STATIC_MAX_ROWS;
constexpr uint32_t kMaxIterations = B3G_CG_STATIC_MAX_ITERATIONS;
constexpr uint32_t kTileRowsStatic = 256;     // rows per SpMV tile (one phase-2 thread each)
constexpr uint32_t kRowBlkRowsStatic = 1024;  // rows per vector-kernel row block
constexpr uint32_t kMaxTiles = (kMaxRows + kTileRowsStatic - 1) / kTileRowsStatic;
constexpr uint32_t kMaxRowBlks = (kMaxRows + kRowBlkRowsStatic - 1) / kRowBlkRowsStatic;
constexpr uint32_t kLogEntries = kMaxIterations + 1;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the parts of the scratch array.
constexpr size_t kVec = align256((size_t)kMaxRows * sizeof(V3));
constexpr size_t kOffInvM = 0;
constexpr size_t kOffR = kOffInvM + kVec;
constexpr size_t kOffS = kOffR + kVec;
constexpr size_t kOffQ = kOffS + kVec;
constexpr size_t kOffD0 = kOffQ + kVec;
constexpr size_t kOffD1 = kOffD0 + kVec;
constexpr size_t kOffTerms = kOffD1 + kVec;
constexpr size_t kOffPartial = kOffTerms + align256((size_t)kMaxRows * 4);
constexpr size_t kPartials = kMaxTiles > kMaxRowBlks ? kMaxTiles : kMaxRowBlks;
constexpr size_t kOffPartialQ = kOffPartial + align256(kPartials * 8);
constexpr size_t kOffLog = kOffPartialQ + align256(kPartials * 8);
constexpr size_t kOffState = kOffLog + align256((size_t)kLogEntries * sizeof(CGLog));
constexpr size_t kOffTiles = kOffState + align256(sizeof(SolverState));
constexpr size_t kOffTileInst = kOffTiles + align256((kMaxTiles + 1) * 4);
constexpr size_t kOffTileBlk = kOffTileInst + align256(kMaxTiles * 4);
constexpr size_t kOffRowBlk = kOffTileBlk + align256((kMaxTiles + 1) * 4);
constexpr size_t kOffBlkInst = kOffRowBlk + align256((kMaxRowBlks + 1) * 4);
constexpr size_t kOffRange = kOffBlkInst + align256(kMaxRowBlks * 4);
constexpr size_t kOffFlag = kOffRange + align256(6 * 4);
constexpr size_t kScratchBytes = kOffFlag + 256;

} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of this device's scratch array (the device is current).
b3g_status scratch_of(int device, unsigned char** out) {
	std::lock_guard<std::mutex> lk(g_lock);
	if (!g_scratch[device]) {
		void* p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			(void)hipGetLastError();
			set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
			return B3G_HIP_ERROR;
		}
		g_scratch[device] = static_cast<unsigned char*>(p);
	}
	*out = g_scratch[device];
	return B3G_OK;
}

// Structure check of the caller's CSR (rowptr non-decreasing from 0 to nnz,
// columns < n). Reads only within rowptr[0..n] and cols[0..nnz); any fault
// found sets *bad, which stops the solve before anything else reads them.
__global__ void __launch_bounds__(256) k_static_check(uint32_t n, uint32_t nnz, const uint32_t* rowptr,
                                                      const uint32_t* cols, uint32_t* bad) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	const uint32_t b0 = rowptr[i], b1 = rowptr[i + 1];
	bool ok = b0 <= b1 && b1 <= nnz && (i != 0 || b0 == 0) && (i != n - 1 || b1 == nnz);
	if (ok)
		for (uint32_t k = b0; k < b1; ++k)
			if (cols[k] >= n) { ok = false; break; }
	if (!ok) *bad = 1u;
}

// Clears the structure flag (the array keeps the previous call's contents).
__global__ void k_static_clear(uint32_t* bad) { *bad = 0u; }

// Row decompositions (fixed 256-row tiles, 1024-row blocks, one instance),
// the solver state (sparse_solver.cpp:38-56: maxIterations, tolerance^2) and
// the Jacobi preconditioner invM_i = 1 / diag(A_ii) (sparse_solver.cpp:45-52,
// first diagonal entry from the head of the row).
struct StaticSetup {
	uint32_t n, nnz, nTiles, nRowBlk, maxIt;
	float eps2;
	const uint32_t* rowptr;
	const uint32_t* cols;
	const float* A;
	const uint32_t* bad;
	uint32_t* tiles;
	uint32_t* tileInst;
	uint32_t* tileBlk;
	uint32_t* rowBlk;
	uint32_t* blkInst;
	uint32_t* range;
	SolverState* st;
	V3* invM;
};

__global__ void __launch_bounds__(256) k_static_setup(StaticSetup s) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	const bool bad = *s.bad != 0u;
	if (i <= s.nTiles) {
		s.tiles[i] = min(i * kTileRowsStatic, s.n);
		s.tileBlk[i] = s.rowptr[min(i * kTileRowsStatic, s.n)];   // read only after the structure check passed
	}
	if (i < s.nTiles) s.tileInst[i] = 0u;
	if (i <= s.nRowBlk) s.rowBlk[i] = min(i * kRowBlkRowsStatic, s.n);
	if (i < s.nRowBlk) s.blkInst[i] = 0u;
	if (i == 0) {
		const uint32_t r[6] = {0u, s.nTiles, 0u, s.nRowBlk, 0u, s.n};
		for (int k = 0; k < 6; ++k) s.range[k] = r[k];
		uint32_t* words = reinterpret_cast<uint32_t*>(s.st);
		for (size_t k = 0; k < sizeof(SolverState) / 4; ++k) words[k] = 0u;
		s.st->maxIt = s.maxIt;
		s.st->eps2_sub = s.eps2;
		s.st->newtonMax = 1u;
		s.st->newtonActive = bad ? 0u : 1u;   // a malformed matrix runs no PCG kernel past its gate
	}
	if (i < s.n && !bad) {
		float axx = 0.0f, ayy = 0.0f, azz = 0.0f;
		for (uint32_t k = s.rowptr[i]; k < s.rowptr[i + 1]; ++k)
			if (s.cols[k] == i) { axx = s.A[k]; ayy = s.A[(size_t)4 * s.nnz + k]; azz = s.A[(size_t)8 * s.nnz + k]; break; }
		s.invM[i] = v3(1.0f / axx, 1.0f / ayy, 1.0f / azz);
	}
}

// The iteration count (b3SolveCGOutput::iterations), or B3G_CG_STATIC_INVALID for a malformed matrix.
__global__ void k_static_iterations(const SolverState* st, const uint32_t* bad, uint32_t* out) {
	*out = *bad ? B3G_CG_STATIC_INVALID : st->it;
}

} // namespace

} // namespace b3g

using namespace b3g;

extern "C" b3g_status b3g_solve_cg_static(int device, void* stream, uint32_t n, uint32_t nnz, const uint32_t* rowptr,
                                          const uint32_t* cols, const float* vals, const float* b, float* x,
                                          uint32_t max_iterations, float tolerance, uint32_t* iterations) {
	if (!rowptr || !cols || !vals || !b || !x) { set_error("b3g_solve_cg_static: null array"); return B3G_INVALID_ARGUMENT; }
	if (n == 0 || n > kMaxRows) {
		set_error("b3g_solve_cg_static: rows must be in [1, B3G_CG_STATIC_MAX_ROWS]");
		return B3G_INVALID_ARGUMENT;
	}
	if (max_iterations > kMaxIterations) {
		set_error("b3g_solve_cg_static: max_iterations exceeds B3G_CG_STATIC_MAX_ITERATIONS");
		return B3G_INVALID_ARGUMENT;
	}
	if (nnz == 0) { set_error("b3g_solve_cg_static: the matrix has no blocks"); return B3G_INVALID_ARGUMENT; }
	if (!(tolerance > 0.0f && tolerance < 1.0f)) { set_error("tolerance must be in (0, 1)"); return B3G_INVALID_ARGUMENT; }
	int count = 0;
	B3G_HIP(hipGetDeviceCount(&count));
	if (device < 0 || device >= count || device >= kMaxDevices) { set_error("device index out of range"); return B3G_INVALID_ARGUMENT; }
	B3G_HIP(hipSetDevice(device));
	const void* ptrs[6] = {rowptr, cols, vals, b, x, iterations};
	for (const void* p : ptrs) {
		if (!p) continue;
		hipPointerAttribute_t attr;
		if (hipPointerGetAttributes(&attr, p) != hipSuccess || attr.type != hipMemoryTypeDevice) {
			(void)hipGetLastError();
			set_error("b3g_solve_cg_static: an array is not device memory");
			return B3G_INVALID_ARGUMENT;
		}
	}
	unsigned char* base = nullptr;
	b3g_status s = scratch_of(device, &base);
	if (s != B3G_OK) return s;

	// A body over the caller's arrays and the scratch array: nothing here allocates.
	Body t;
	t.device = device;
	t.stream = static_cast<hipStream_t>(stream);
	t.built = true;
	t.n = n;
	t.nnz = nnz;
	t.aStride = nnz;   // the caller's nine planes are nnz apart
	t.nInst = 1;
	t.nTiles = (n + kTileRowsStatic - 1) / kTileRowsStatic;
	t.nRowBlk = (n + kRowBlkRowsStatic - 1) / kRowBlkRowsStatic;
	t.maxInstPartials = t.nTiles > t.nRowBlk ? t.nTiles : t.nRowBlk;
	t.partialCount = (uint32_t)kPartials;
	t.logStride = kLogEntries;
	t.reduction = B3G_REDUCE_FP64;
	auto at = [&](DBuf& d, size_t off, size_t bytes) { d.p = base + off; d.bytes = bytes; };
	at(t.invM, kOffInvM, kVec);
	at(t.r, kOffR, kVec);
	at(t.s, kOffS, kVec);
	at(t.q, kOffQ, kVec);
	at(t.d0, kOffD0, kVec);
	at(t.d1, kOffD1, kVec);
	at(t.terms, kOffTerms, (size_t)kMaxRows * 4);
	at(t.partial, kOffPartial, kPartials * 8);
	at(t.partialQ, kOffPartialQ, kPartials * 8);
	at(t.cgLog, kOffLog, (size_t)kLogEntries * sizeof(CGLog));
	at(t.state, kOffState, sizeof(SolverState));
	at(t.tiles, kOffTiles, (kMaxTiles + 1) * 4);
	at(t.tileInst, kOffTileInst, kMaxTiles * 4);
	at(t.tileBlk, kOffTileBlk, (kMaxTiles + 1) * 4);
	at(t.rowBlk, kOffRowBlk, (kMaxRowBlks + 1) * 4);
	at(t.blkInst, kOffBlkInst, kMaxRowBlks * 4);
	at(t.instRange, kOffRange, 6 * 4);
	t.A.p = const_cast<float*>(vals);
	t.rowptr.p = const_cast<uint32_t*>(rowptr);
	t.cols.p = const_cast<uint32_t*>(cols);
	t.pb.p = const_cast<float*>(b);
	t.px.p = x;
	uint32_t* bad = reinterpret_cast<uint32_t*>(base + kOffFlag);

	hipLaunchKernelGGL(k_static_clear, dim3(1), dim3(1), 0, t.stream, bad);
	B3G_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_static_check, dim3((n + 255) / 256), dim3(256), 0, t.stream, n, nnz, rowptr, cols, bad);
	B3G_HIP(hipGetLastError());
	StaticSetup su;
	su.n = n; su.nnz = nnz; su.nTiles = t.nTiles; su.nRowBlk = t.nRowBlk; su.maxIt = max_iterations;
	su.eps2 = tolerance * tolerance;
	su.rowptr = rowptr; su.cols = cols; su.A = vals; su.bad = bad;
	su.tiles = dptr<uint32_t>(t.tiles); su.tileInst = dptr<uint32_t>(t.tileInst); su.tileBlk = dptr<uint32_t>(t.tileBlk);
	su.rowBlk = dptr<uint32_t>(t.rowBlk); su.blkInst = dptr<uint32_t>(t.blkInst);
	su.range = dptr<uint32_t>(t.instRange); su.st = dptr<SolverState>(t.state); su.invM = dptr<V3>(t.invM);
	const uint32_t setupThreads = std::max(n, t.nTiles + 1);
	hipLaunchKernelGGL(k_static_setup, dim3((setupThreads + 255) / 256), dim3(256), 0, t.stream, su);
	B3G_HIP(hipGetLastError());
	s = enqueue_solve_cg(t, max_iterations, tolerance);
	if (s != B3G_OK) return s;
	if (iterations) {
		hipLaunchKernelGGL(k_static_iterations, dim3(1), dim3(1), 0, t.stream, dptr<SolverState>(t.state), bad, iterations);
		B3G_HIP(hipGetLastError());
	}
	return B3G_OK;
}
