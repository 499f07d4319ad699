// Internal state of a b3g body (host bookkeeping + device arrays).
#pragma once

#include "b3g.h"
#include "b3g_math.h"

#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace b3g {

enum : uint8_t { kStatic = 0, kKinematic = 1, kDynamic = 2 };

// Contribution codes of the per-slot lists: a staged block index (< 2^30),
// kTetCode | (i + 4 j) << kTetBlockShift | e for block (i, j) of tetrahedron e
// computed during assembly, or
// kUVCode | variant << kUVShift | e for a stretch (variants 0..17: 9 part +
// 3 i + j) or shear (18 + 3 i + j) block computed during assembly from the
// element's record.
constexpr uint32_t kTetCode = 0x80000000u;
constexpr uint32_t kTetBlockShift = 27;
constexpr uint32_t kTetMaxElements = 1u << kTetBlockShift;   // tetrahedra per body or batch
constexpr uint32_t kUVCode = 0x40000000u;
constexpr uint32_t kUVShift = 25;
constexpr uint32_t kUVMaxElements = 1u << kUVShift;   // stretch / shear forces per body or batch
constexpr uint32_t kShearVariant = 18;

// Thread-local error text for b3g_last_error().
void set_error(const std::string& msg);

#define B3G_HIP(expr)                                                                                  \
	do {                                                                                               \
		hipError_t e_ = (expr);                                                                        \
		if (e_ != hipSuccess) {                                                                        \
			::b3g::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                       \
			return B3G_HIP_ERROR;                                                                      \
		}                                                                                              \
	} while (0)

inline uint32_t plane_stride(uint32_t nnz) { return (nnz + 3u) & ~3u; }

// A device allocation owned by the body.
struct DBuf {
	void* p = nullptr;
	size_t bytes = 0;
};

template <class T> inline T* dptr(const DBuf& b) { return static_cast<T*>(b.p); }

// Per-type element tables (device), list order within the type = reverse creation.
struct TetTable {
	uint32_t m = 0;
	DBuf p;          // 4 planes of uint32 (particle list index)
	DBuf rest;       // 12 planes: rest positions x1..x4
	DBuf rec;        // 16 floats per element: gradients, volume, Lame parameters (b3g_tet.h)
	DBuf q;          // float4 per element: warm-start quaternion (x y z s)
	DBuf kd;         // 1 plane
	DBuf param;      // 3 planes (E, nu, kd) for the set-up kernel
	bool anyDamping = false;
	uint32_t f = 0, fd = 0;   // force staging bases (vertex i: base + i*m + e)
};

struct TriTable {
	uint32_t m = 0;
	DBuf p;          // 3 planes
	DBuf x2;         // 6 planes: rest 2D coordinates
	DBuf invS;       // 4 planes
	DBuf K0;         // 36 planes: 9 Mat22 blocks
	DBuf kd;
	bool anyDamping = false;
	uint32_t jx = 0, jv = 0, f = 0, fd = 0;
};

struct UVTable {     // stretch (20 coefficient planes) and shear (16)
	uint32_t m = 0;
	DBuf p;          // 3 planes
	DBuf c;          // coefficient planes
	uint32_t ncoef = 0;
	uint32_t f = 0;  // force staging base; per element local slots as described in b3g_setup
	// float4 records the assembly rebuilds the Jacobian blocks from.
	// stretch, 3 per part (u, v): (n_w, |w|) written per linearisation,
	//   (dw/dx, alpha), (ks, kd, b, 0) from set-up;
	// shear, 4: (w_u, alpha), (w_v, ks) written per linearisation,
	//   (dw_u/dx, kd), (dw_v/dx, 0) from set-up.
	DBuf rec;
	bool anyKs = false, anyKd = false;
};

struct PairTable {   // springs (2 particles, 3 coef) and mouse (4 particles, 6 coef)
	uint32_t m = 0;
	uint32_t np = 0;
	DBuf p;
	DBuf c;
	uint32_t ncoef = 0;
	uint32_t jx = 0, jv = 0, f = 0;
	bool anyKs = false, anyKd = false;
};

// A world fixture (device copy is a flat array of these).
struct WorldShape {
	uint32_t type;
	float radius, friction;
	V3 a, b, c;              // sphere center / capsule centers / triangle vertices / box extents
	uint32_t hasE[3];
	V3 e[3];
	Xf xf;
	AABB grid;               // sdf voxel-grid aabb
	uint32_t W, H, D;
	uint32_t invert;
	uint64_t voxOffset;      // float offset of this grid in the voxel buffer
	AABB aabb;               // b3Shape::ComputeAABB of the fixture (static)
	// mesh (scale in a): the mesh's vertices start at meshVert, its triangles
	// (6 words each) at meshTri; its tree's leaves and internal nodes are
	// positions [meshNode, meshNode + meshCount) of the mesh-tree arrays
	uint32_t meshVert, meshTri, meshNode, meshCount;
};

// Per-instance, per-iteration scalars of one PCG solve (sparse_solver.cpp:57-103):
// entry k holds delta_k = r_k.s_k, alpha_k, beta_k and whether iteration k runs.
struct CGLog {
	float delta, alpha, beta;
	uint32_t active;
};

// Device-side scalar state of the Newton and PCG loops.
struct SolverState {
	// PCG (sparse_solver.cpp:57-103); per-iteration scalars are in the CGLog
	float delta0, eps2_sub;
	uint32_t it, maxIt;
	// Newton (sparse_force_solver.cpp:141-225)
	float error, error0, eps2;
	uint32_t newtonIt, newtonActive, newtonMax;
	// running statistics (force_solver.cpp:30-34)
	uint32_t statNewton, statMinSub, statMaxSub;
	uint32_t contactCount;
	// reduction mode of the dot products (B3G_REDUCE_*): FP64 tree (default / fused / separate) or sequential FP32 in row order
	uint32_t seq;
	// iterative refinement (precision form): FP64 residual norms of the
	// Newton iteration's linear system, gate and counters of the sweeps
	double res0, res;
	float refineTol2;
	uint32_t refineOn, refineActive, sweeps, maxSweeps, itSum, statSweeps;
	// B3G_FAULT_* bits (sticky): a non-positive Jacobi diagonal, a non-finite state
	uint32_t faults;
};


// The launch schedule of the PCG (b3g_schedule): SpMV tile shape, vector
// row-block size and where the FP64 dot products are summed. Any schedule
// gives the reference's per-row and per-block arithmetic; the tiles and row
// blocks only group the FP64 partial sums of the dot products.
struct Schedule {
	uint32_t tileBlocks = 1024;    // blocks per SpMV tile (at most 1024)
	uint32_t tileRows = 256;       // rows per SpMV tile (at most 256)
	uint32_t rowBlockRows = 1024;  // rows per vector-kernel block (at most 1024)
	int fused = -1;                // FP64 mode: -1 by instance size, 0 separate, 1 fused
	uint32_t maxGroups = 0;        // workgroups per PCG launch at most (0: one per tile or row block)
	// persistent PCG: instances of at most persistTiles SpMV tiles (0: none)
	// run all their iterations in one launch, persistGroups workgroups
	// (0: one per such instance), beside the launch sequence of the others
	uint32_t persistTiles = 0;
	uint32_t persistGroups = 0;
};

// A body holds a batch of one or more independent instances (bodies of the
// reference). Their particles, elements, blocks and contact pairs are
// concatenated; every instance keeps its own list orders, its own Newton /
// PCG state and convergence, and the reductions never mix instances.
// Instance k (creation order) occupies the list-order rows
// [rowBegin[k], rowEnd[k]); instances appear in reverse order along the rows
// because list order reverses creation order.
struct Body {
	int device = 0;
	hipStream_t stream = nullptr;
	bool built = false;

	// ---------------------------------------------------------------- sizes
	uint32_t n = 0;          // particles (all instances)
	uint32_t nnz = 0;        // 3x3 blocks of pA
	uint32_t aStride = 0;    // plane stride of pA's nine value planes and of cols: nnz rounded up to 4
	uint32_t nInst = 1;      // instances
	uint32_t nSpheres = 0, nWorld = 0, nPairs = 0;
	DBuf stampCounter;                            // uint32 on the device: FindNewContacts calls so far
	std::vector<uint32_t> hostRowInst;            // instance of each row (empty = all instance 0)
	std::vector<uint32_t> hostRowBegin, hostRowEnd;
	std::vector<V3> hostGravity;                  // per instance
	DBuf gravity;                                 // V3 per instance
	DBuf rowInst;                                 // uint32 per row
	DBuf instRange;                               // 6 uint32 per instance: tiles, row blocks, rows (begin, end)

	// ---------------------------------------------------------------- particles (list order)
	DBuf type, mass, invMass, damping;            // uint8, float x3
	DBuf x, v, force, translation;                // V3
	DBuf x0, v0, fe, y, w;                        // V3 (per-step work)
	DBuf sphereOf;                                // int32 per particle: sphere list index or -1
	DBuf sphereRadius, sphereFriction;            // per sphere (list order)
	DBuf sphereParticle;                          // uint32 per sphere
	std::vector<uint32_t> hostSphereParticle;
	std::vector<uint32_t> hostSphereCreation;     // creation index of each sphere fixture (list order)
	std::vector<float> hostMass;                  // list order
	std::vector<uint8_t> hostType;

	// ---------------------------------------------------------------- elements
	TetTable tet;
	TriTable tri;
	UVTable stretch, shear;
	PairTable spring, mouse;
	uint32_t nJStage = 0, nFStage = 0;
	DBuf jstage;     // M33 contributions (element-major per local block)
	DBuf fstage;     // V3 contributions

	// ---------------------------------------------------------------- symbolic pattern (pA CSR, list order)
	DBuf rowptr, cols, slotRow, diagSlot;         // uint32
	DBuf jxPtr, jxIdx, jvPtr, jvIdx;              // per-slot contribution lists (stage indices)
	DBuf bPtr, bSlot;                             // per-row dfdx slots in first-touch order (b's SpMV)
	DBuf fPtr, fIdx;                              // per-row force contributions
	std::vector<uint32_t> hostTiles, hostTileInst;   // tiles and their instances, host copies
	DBuf tiles;                                   // uint32 row boundaries of SpMV / row-assembly tiles
	uint32_t nTiles = 0;
	uint32_t maxInstPartials = 0;   // largest tile or row-block count of one instance
	DBuf tileInst;                                // uint32 instance per tile
	DBuf tileBlk;                                 // uint32 block boundaries of the tiles (rowptr at tiles)
	DBuf tileGate;                                // uint32 per tile: 0 once its instance is done with the current PCG solve
	DBuf blkGate;                                 // uint32 per row block: the same for the vector update
	DBuf rowBlk;                                  // uint32 boundaries of <= 256-row blocks (vector kernels)
	DBuf blkInst;
	uint32_t nRowBlk = 0;
	// Column-sorted storage of pA (sortedA): within each kTileCap-block chunk of
	// every SpMV tile the blocks are stored in column order, so a wavefront's
	// gathers of the vector at the columns touch neighbouring rows. code[k] =
	// (column - tileBase[tile]) << 10 | (k's CSR position within the chunk);
	// the SpMV writes each product to its CSR position, so every row still
	// sums in CSR order. slotPos maps a CSR slot to its storage position,
	// storeSlot the reverse. Without sortedA, storage position = CSR slot.
	bool sortedA = false;
	DBuf code, tileBase, slotPos, storeSlot;     // uint32
	DBuf slotTiles;                               // uint32 slot boundaries of slot-assembly tiles
	DBuf slotTileInst;
	uint32_t nSlotTiles = 0;
	bool anyJv = false;                           // the body has dfdv contributions beyond mass damping
	std::vector<uint32_t> hostRowptr, hostCols;

	// ---------------------------------------------------------------- linear system / PCG
	DBuf A;          // 9 planes x nnz (pA blocks, column-major element planes)
	DBuf tB;         // V3 per slot: (h dfdx_ij) w_j
	DBuf pb, invM;   // V3 per row
	DBuf px, r, s, q, d0, d1, fi;   // V3 per row (px = py warm start)
	DBuf partial;    // double per tile / row block: residual, vector-update and Newton reductions
	DBuf partialQ;   // double per tile: the d.q reduction of the SpMV kernel
	uint32_t partialCount = 0;
	DBuf cgLog;      // CGLog per instance and PCG iteration
	uint32_t logStride = 0;   // entries per instance
	// precision form: FP64 iterative refinement around an inner PCG on FP32
	// or BF16-off-diagonal blocks
	int solver = 0;           // B3G_SOLVER_*
	uint32_t innerIt = 10;
	float innerTol = 1e-2f;
	DBuf x64;        // double3 per row: refined solution
	DBuf dx, rhsF;   // V3 per row: inner correction, inner right-hand side
	DBuf partialR;   // double per tile: FP64 residual norm
	// BF16 blocks in pair planes: word m of plane c holds entry c of the
	// blocks at storage positions 2m and 2m + 1 of a tile (low, high half);
	// each tile starts a new pair (pairBase: its first pair), so every chunk
	// of kTileCap blocks starts on a word; diagonal slots zero
	DBuf AbP, pairBase;
	uint32_t pairStride = 0;   // words per plane (pairs, rounded up to 64)
	DBuf diagF;      // 9 planes x rows: FP32 diagonal blocks of the BF16 operator
	DBuf tet64;      // 12 planes x tets: FP64 invE (9), volume, lambda, mu
	DBuf mass64;     // double per row: lumped masses in FP64
	std::vector<double> hostMass64;
	DBuf fstage64;   // D3 per force staging slot (tetrahedra only): FP64 forces
	DBuf tB64;       // double3 per block: (h dfdx_ij) w_j in FP64
	DBuf pb64;       // double3 per row: FP64 right-hand side
	DBuf terms;      // float per row: per-row dot terms (sequential reduction mode)
	int reduction = 0;
	DBuf state;      // SolverState per instance
	DBuf faultWord;  // uint32: B3G_FAULT_* bits raised by any instance since the last b3g_synchronize

	// ---------------------------------------------------------------- contacts (sphere x world pairs of one instance)
	DBuf worlds;     // WorldShape[nWorld]
	DBuf voxels;     // float
	// mesh world fixtures: vertices (V3), triangles (6 words), and per tree
	// position: leaf triangle (local index), leaf box, internal node children
	// and box (b3g_bvh.h layout, one segment per mesh)
	DBuf meshV, meshT, meshLeaf, meshLeafBox, meshChild, meshNodeBox;
	DBuf pairStart;                               // uint32 per sphere + 1: its pairs, one per world fixture of its instance
	DBuf pairSphere, pairWorld;                   // uint32 per pair
	std::vector<uint32_t> hostPairStart, hostPairSphere, hostPairWorld;
	DBuf cActive, cStamp, cFriction;              // uint32 per pair
	DBuf cNormalForce;                            // float per pair
	DBuf cManifold;                               // 2 V3 per pair (point, normal)
	DBuf cHit;                                    // uint32 per pair (this linearisation)
	DBuf cF;                                      // V3 per pair
	DBuf cK;                                      // M33 per pair
	std::vector<WorldShape> hostWorlds;

	// ---------------------------------------------------------------- captured step (hipGraph)
	bool useGraph = false;
	hipGraphExec_t graphExec = nullptr;
	std::vector<double> graphKey;                 // the step parameters and modes the graph was captured with

	// ---------------------------------------------------------------- ray casts (triangle fixtures)
	uint32_t nRayTri = 0;
	std::vector<uint32_t> hostRayTri;             // per fixture: 3 particle list indices, creation index;
	                                              // ordered by instance, then creation index
	std::vector<uint32_t> hostRayInstBegin;       // nInst + 1 offsets into hostRayTri
	uint64_t xVersion = 0;                        // advanced by everything that moves particles
	uint64_t rtVersion = ~0ull;                   // xVersion the device tree was built at
	size_t rtSortBytes = 0;
	DBuf rtTri, rtInstBegin, rtBox, rtKey, rtKey2, rtVal, rtVal2, rtSortTemp, rtLeafTri, rtLeafBox, rtChild, rtNodeBox,
	    rtChunk1, rtChunk2, rtChunk3, rtInstBox, rtInstPart, rtRays, rtHits, rtFault;
	size_t rtRayCap = 0;

	// ---------------------------------------------------------------- action forces
	// m_f1.. of spring (2), stretch (3), shear (3) and mouse (4) forces:
	// accumulated over a step's Newton iterations when enabled; V3 planes
	// [base + i m + e] per type, element e in creation order
	bool trackActions = false;
	DBuf act;
	uint32_t actSpring = 0, actStretch = 0, actShear = 0, actMouse = 0, actCount = 0;

	// ---------------------------------------------------------------- early exit of the PCG
	// With earlyExit the PCG's iterations are enqueued in chunks; after each
	// chunk a one-workgroup kernel writes into pinned host memory whether any
	// instance's last iteration ran, and the host stops enqueuing once none
	// did (the reference's loop exit, sparse_solver.cpp:63-73). The host waits
	// on the chunk before the last one enqueued, so the device never idles.
	// Off while a step is captured into a graph (a fixed launch sequence).
	bool earlyExit = false;
	bool capturing = false;
	uint32_t* hostFlag = nullptr;         // hipHostMalloc'd: per slot, (sequence << 1) | any active
	hipEvent_t flagEvent[2] = {nullptr, nullptr};
	uint32_t flagSeq = 0;
	uint32_t lastConverged = 0;           // largest iteration count of the previous solve (first-chunk size)

	// ---------------------------------------------------------------- launch schedule
	// sched: the PCG's tile shape, row-block size and FP64 sum place
	// (b3g_set_schedule / b3g_autotune), which build_tiles lays out; lastStep:
	// the parameters of the last b3g_step, which the autotuner's timings repeat
	Schedule sched;
	struct {
		bool valid = false;
		float dt = 0.0f, tol = 0.0f, stol = 0.0f;
		uint32_t it = 0, sub = 0;
	} lastStep;
	struct StepWorker* worker = nullptr;  // the thread that enqueues eager steps (b3g_api.hip)

	// ---------------------------------------------------------------- persistent PCG (Schedule::persistTiles)
	// persist: uint32 per instance, 1 for the instances the persistent kernel
	// solves (the launch sequence skips them after the solve's start); plist:
	// those instances, largest first; stream2 runs the persistent kernel,
	// forked from and joined to the body's stream by events
	DBuf persist, plist;
	// the persistent kernel's r.s partials (tiles, then row blocks): its
	// instances run out of step with one another and with the launch sequence,
	// whose partial array holds tile and row-block partials by turns
	DBuf partialP;
	uint32_t nPersist = 0;
	hipStream_t stream2 = nullptr;
	hipEvent_t forkEvent = nullptr, joinEvent = nullptr;

	// ---------------------------------------------------------------- diagnostics
	bool captureFi = false;

	std::vector<DBuf*> owned;
};

// Allocation helpers (b3g_setup.hip)
b3g_status dalloc(Body& b, DBuf& buf, size_t bytes);
// Allocates the action-force planes when they are tracked (after a build).
b3g_status ensure_actions(Body& b);
// b3BuildSDF on the device (b3g_sdf.hip): the grid, then the voxels.
b3g_status sdf_grid(const float* vertices, uint32_t nv, const float cell[3], float extension, AABB& box, uint32_t dims[3]);
b3g_status build_sdf(int device, const float* vertices, uint32_t nv, const uint32_t* triangles, uint32_t nt,
                     uint32_t stride, const float cell[3], float extension, float* voxels, uint64_t capacity);
// b3Body::RayCastSingle for count rays against instance inst (b3g_raycast.hip).
b3g_status ray_cast(Body& b, uint32_t inst, uint32_t count, const float* p1, const float* p2, int64_t* index,
                    float* fraction, float* normal);
b3g_status ray_cast_device(Body& b, uint32_t inst, uint32_t count, const float* p1, const float* p2, int64_t* index,
                    float* fraction, float* normal);
b3g_status dupload(Body& b, DBuf& buf, const void* host, size_t bytes);
void dfree_all(Body& b);

// Host build (b3g_setup.hip): a batch of `count` instances.
b3g_status build_body(Body& b, const b3g_body_desc* descs, uint32_t count);

// Symbolic pattern on the host only (b3g_setup.hip), for diagnostics.
b3g_status host_pattern(const b3g_body_desc& d, std::vector<uint32_t>& rowptr, std::vector<uint32_t>& cols);

// Kernel launchers (b3g_kernels.hip)
b3g_status launch_tet_setup(Body& b);
b3g_status enqueue_step(Body& b, float dt, uint32_t iters, uint32_t subIters, float tol, float subTol,
                        bool assembleOnly);
b3g_status prepare_step(Body& b, uint32_t subIters);
void drop_graph(Body& b);
b3g_status enqueue_update_contacts(Body& b);
b3g_status enqueue_find_contacts(Body& b);
b3g_status enqueue_spmv(Body& b, const DBuf& xin, DBuf& yout);
b3g_status time_pcg_spmv(Body& b, uint32_t launches, double* mean_ms);
b3g_status time_pcg_spmv_cold(Body& b, uint32_t launches, size_t evictBytes, double* mean_ms);
b3g_status enqueue_solve_cg(Body& b, uint32_t maxIt, float eps);
// sorted: also the column-sorted storage of pA (bodies assembled on the device;
// operator-level matrices given in CSR storage keep it)
b3g_status build_tiles(Body& b, bool sorted = true, bool tiles = true);
// Times candidate schedules and keeps the fastest (b3g_autotune): `timer`
// measures the body under its current schedule (ms); without one, a PCG
// iteration on the body's system (time_pcg_iteration). *tunedMs /
// *defaultMs: the measure of the kept schedule and of the default.
b3g_status autotune(Body& b, const std::function<b3g_status(double*)>& timer, double* tunedMs, double* defaultMs);
// One PCG iteration's kernels (SpMV, then the vector update, and the finalize
// launches of the separate reduction) `reps` times, forced on the solver's
// scratch vectors whatever its state; mean ms per iteration.
b3g_status time_pcg_iteration(Body& b, uint32_t reps, double* ms);
b3g_status ensure_partials(Body& b, uint32_t blocks);
b3g_status ensure_cg_log(Body& b, uint32_t iterations);
// Pinned flag memory and events of the PCG's early exit (allocated once, before any capture).
b3g_status ensure_exit_flags(Body& b);
void free_exit_flags(Body& b);
// The persistent PCG's second stream and its events.
void free_persistent(Body& b);

// Operator-level backward Euler (b3g_be.hip): arguments of its kernels.
// Matrices of the force model are in list order (rows head to tail), one
// M33 per block; slots of pA carry their row, dfdx / dfdv entry (-1: none).
struct BEArgs {
	uint32_t n = 0, nnz = 0;
	float h = 0.0f, hh = 0.0f;
	uint32_t useV = 0;                // h != 0: the (h dfdx) w term of b exists
	const uint32_t* rowptr = nullptr; // pA (CSR slots)
	const uint32_t* cols = nullptr;
	const uint32_t* slotRow = nullptr;
	const uint32_t* slotPos = nullptr;  // storage position of a CSR slot (column-sorted storage), or null
	const uint32_t* diagSlot = nullptr;
	const int32_t* slotX = nullptr;
	const int32_t* slotV = nullptr;
	const uint32_t* rowptrX = nullptr; // dfdx
	const uint32_t* colsX = nullptr;
	const M33* valX = nullptr;
	const M33* valV = nullptr;
	const M33* M = nullptr;
	const M33* S = nullptr;
	const V3 *x0 = nullptr, *v0 = nullptr, *fe = nullptr, *fi = nullptr, *y = nullptr, *z = nullptr;
	V3 *x = nullptr, *v = nullptr, *w = nullptr, *px = nullptr, *pb = nullptr, *invM = nullptr;
	M33* Araw = nullptr;
	float* A = nullptr;
	uint32_t stride = 0;
	SolverState* st = nullptr;
};
// Solver state, py = 0 (once per solve); one Newton iteration after the force
// model's output is on the device: linearisation, PCG, update, error.
b3g_status enqueue_be_begin(Body& b, const BEArgs& a, uint32_t maxIt, float tol, uint32_t maxSub, float subTol,
                            uint32_t minSub, uint32_t maxSubStat);
b3g_status enqueue_be_iteration(Body& b, const BEArgs& a, uint32_t maxSub, float subTol);

} // namespace b3g
