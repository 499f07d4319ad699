/*
 * b3g -- MI355X implicit step of bounce_softbody, C ABI.
 *
 * One b3g_body holds a soft body resident in HBM: particles, force elements,
 * fixtures, world fixtures and contact pairs. b3g_step() performs
 * b3Body::Step (dynamics/body.cpp:552-604) on the GPU: contact update,
 * backward-Euler linearisation, element force/Jacobian assembly into the
 * block-sparse 3x3 system, S-filtered Jacobi PCG, friction and new-contact
 * search. Results equal the reference's FP32 results operation for operation,
 * except the PCG dot products and the Newton error norm, which are reduced
 * in FP64 instead of sequentially in FP32.
 *
 * Conventions
 *   - Arrays passed in are host arrays in *creation order* (the order in
 *     which a program calls b3Body::Create*). The library applies the
 *     reference's list order (reverse creation, body.cpp:67-75) internally.
 *   - Every function returns a b3g_status; b3g_last_error() describes the
 *     last failure of the calling thread.
 *   - A body owns one HIP stream. b3g_step() is asynchronous; functions that
 *     return data synchronise that stream. A body must not be used from two
 *     threads at once.
 *   - No torch types: plain pointers and sizes only.
 */
#ifndef B3G_H
#define B3G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum b3g_status {
	B3G_OK = 0,
	B3G_INVALID_ARGUMENT = 1,
	B3G_HIP_ERROR = 2,
	B3G_OUT_OF_MEMORY = 3,
	B3G_DEGENERATE_ELEMENT = 4,   /* zero-volume tetrahedron / zero-area triangle (replaces B3_ASSERT(det != 0)) */
	B3G_UNSUPPORTED = 5,
	B3G_NOT_COMMITTED = 6,
	B3G_NON_POSITIVE_DIAGONAL = 7, /* a Jacobi diagonal entry of pA was not > 0 (replaces B3_ASSERT, sparse_solver.cpp:45-52) */
	B3G_NON_FINITE = 8             /* a step produced a non-finite position or velocity */
} b3g_status;

/* Per-instance fault bits (b3g_step_stats::faults), sticky until the body is rebuilt or a state is loaded. */
enum { B3G_FAULT_DIAGONAL = 1, B3G_FAULT_NON_FINITE = 2 };

typedef struct b3g_body_s* b3g_body;

/* Force type codes (b3ForceType, dynamics/forces/force.h:30-39). */
enum { B3G_FORCE_STRETCH = 1, B3G_FORCE_SHEAR = 2, B3G_FORCE_SPRING = 3, B3G_FORCE_MOUSE = 4,
       B3G_FORCE_TRI_ELEMENT = 5, B3G_FORCE_TET_ELEMENT = 6 };

/* World shape codes (b3Shape::Type, collision/shapes/shape.h:42-51). Every shape type is supported; a body
 * (each instance of a batch) takes up to 4096 world fixtures, and sphere fixtures x world fixtures must stay
 * below 2^31 contact pairs (every sphere pairs with every world fixture, contact_manager.cpp:65-82). */
enum { B3G_SHAPE_SPHERE = 0, B3G_SHAPE_CAPSULE = 1, B3G_SHAPE_TRIANGLE = 2, B3G_SHAPE_BOX = 3,
       B3G_SHAPE_MESH = 4, B3G_SHAPE_SDF = 5 };

#define B3G_WORLD_PARAMS 32

/*
 * Body description, creation order. Mirrors the reference's definitions:
 *   particles    b3ParticleDef (particle.h:39-68): type, position, velocity, damping
 *   spheres      b3SphereFixtureDef (sphere_fixture.h:28-37) + b3FixtureDef radius/friction
 *   trifix       b3TriangleFixtureDef (triangle_fixture.h:31-45): p1..p3, rest v1..v3, density, radius, friction
 *   tetfix       b3TetrahedronFixtureDef (tetrahedron_fixture.h:28-43)
 *   forces       b3Body::CreateForce calls in order: (type, index into the type's table)
 *   tet          b3TetrahedronElementForceDef: p[4], rest v[12], {youngModulus, poissonRatio, stiffnessDamping}
 *   tri          b3TriangleElementForceDef: p[3], rest v[9], {Ex, Ey, Es, nu_xy, nu_yx, stiffnessDamping}
 *   stretch      b3StretchForceDef: p[3], {u1 v1 u2 v2 u3 v3 ks_u kd_u b_u ks_v kd_v b_v}
 *   shear        b3ShearForceDef: p[3], {u1 v1 u2 v2 u3 v3 ks kd}
 *   spring       b3SpringForceDef: p[2], {length ks kd}
 *   mouse        b3MouseForceDef: p[4], {w2 w3 w4 length ks kd}
 *   world        b3WorldFixtureDef (world_fixture.h:29-42): shape type, B3G_WORLD_PARAMS floats, friction
 *                param[0] = shape radius;
 *                sphere [1..3] center; capsule [1..6] centers;
 *                triangle [1..9] vertices, [10..12] hasE flags, [13..21] edge-neighbour vertices;
 *                box [1..3] extents, [4..6] translation, [7..10] quaternion (x y z s);
 *                sdf [1..6] voxel-grid aabb, [7..9] voxel counts W H D, [10..12] translation,
 *                    [13..16] quaternion, [17] invert; voxels[w] holds W*H*D values (index i + jW + kWH);
 *                mesh (b3MeshShape, mesh_shape.h:30-50) [1..3] scale, [4..6] translation, [7..10] quaternion;
 *                    the b3Mesh (mesh.h:52-87) in w_mesh_vertices / w_mesh_triangles / w_mesh_counts.
 *                    The library builds the mesh's tree (b3Mesh::BuildTree); wing vertices
 *                    (b3Mesh::BuildAdjacency) are part of the triangles as given.
 */
typedef struct b3g_body_desc {
	uint32_t n_particles;
	const uint8_t* p_type;        /* 0 static, 1 kinematic, 2 dynamic (particle.h:31-36) */
	const float* p_pos;
	const float* p_vel;
	const float* p_damping;

	uint32_t n_spheres;
	const uint32_t* s_particle;
	const float* s_radius;
	const float* s_friction;

	uint32_t n_trifix;
	const uint32_t* tf_p;
	const float* tf_v;
	const float* tf_density;
	const float* tf_radius;
	const float* tf_friction;

	uint32_t n_tetfix;
	const uint32_t* ttf_p;
	const float* ttf_v;
	const float* ttf_density;
	const float* ttf_radius;
	const float* ttf_friction;

	uint32_t n_forces;
	const uint32_t* f_type;
	const uint32_t* f_index;

	uint32_t n_tet;
	const uint32_t* tet_p;
	const float* tet_v;
	const float* tet_param;

	uint32_t n_tri;
	const uint32_t* tri_p;
	const float* tri_v;
	const float* tri_param;

	uint32_t n_stretch;
	const uint32_t* st_p;
	const float* st_param;

	uint32_t n_shear;
	const uint32_t* sh_p;
	const float* sh_param;

	uint32_t n_spring;
	const uint32_t* sp_p;
	const float* sp_param;

	uint32_t n_mouse;
	const uint32_t* mo_p;
	const float* mo_param;

	uint32_t n_world;
	const uint32_t* w_type;
	const float* w_param;
	const float* w_friction;
	const float* const* w_voxels;

	float gravity[3];             /* b3Body::SetGravity (body.h) */

	/* Mesh world fixtures, one entry per world fixture (NULL / 0 for other shapes):
	 * w_mesh_vertices[w]: 3 floats per vertex; w_mesh_triangles[w]: 6 per triangle, b3Triangle's
	 * v1 v2 v3 u1 u2 u3 (u = 0xFFFFFFFF, B3_NULL_VERTEX: open edge); w_mesh_counts[2 w], [2 w + 1]:
	 * vertex and triangle counts. Fixtures passing the same arrays share one device copy. */
	const float* const* w_mesh_vertices;
	const uint32_t* const* w_mesh_triangles;
	const uint32_t* w_mesh_counts;
} b3g_body_desc;

typedef struct b3g_step_stats {
	uint32_t newton_iterations;   /* b3_forceSolverIterations (force_solver.cpp:30) */
	uint32_t min_sub_iterations;  /* b3_forceSolverMinSubIterations (running minimum) */
	uint32_t max_sub_iterations;  /* b3_forceSolverMaxSubIterations (running maximum) */
	uint32_t contact_count;       /* live sphere-shape contacts (b3ContactManager::m_contactCount) */
	uint32_t refinement_sweeps;   /* precision form: FP64 residual sweeps of the last Newton iteration */
	float relative_residual;      /* precision form: last FP64 ||pb - pA x|| / ||pb - pA x0|| */
	uint32_t faults;              /* B3G_FAULT_* bits this instance has raised */
} b3g_step_stats;

const char* b3g_status_string(b3g_status s);
const char* b3g_last_error(void);

/* Number of visible devices. */
b3g_status b3g_device_count(int* count);

/* Creates an empty body on `device`. Replaces b3Body::b3Body (body.cpp:31-55). */
b3g_status b3g_body_create(b3g_body* out, int device);

/* Frees all device memory of the body. Replaces b3Body::~b3Body (body.cpp:57-60). */
b3g_status b3g_body_destroy(b3g_body body);

/*
 * Builds the body from a description: particles, fixtures (mass lumping of
 * b3Body::ResetMass, body.cpp:394-466), forces (element rest data of
 * ResetElementData, tetrahedron_element_force.cpp:219-265 and
 * triangle_element_force.cpp:208-279), world fixtures, and the symbolic
 * block pattern of the linear system. Replaces the sequence of
 * b3Body::CreateParticle/CreateSphere/CreateTriangle/CreateTetrahedron/
 * CreateForce/CreateFixture/SetGravity calls (body.cpp:62-354) in O(N).
 * Any previous contents are discarded.
 */
b3g_status b3g_body_build(b3g_body body, const b3g_body_desc* desc);

/*
 * Batched form (no counterpart in the reference, which steps one b3Body at a
 * time): builds `count` independent bodies ("instances") into one handle.
 * b3g_step then steps all of them with one launch sequence: every kernel
 * covers all instances, each instance keeps its own Newton / PCG state,
 * convergence and contact list, and no reduction mixes instances, so each
 * instance's trajectory is the one b3g_body_build + b3g_step would give it
 * alone. Per-particle and per-element arrays of the other entry points are
 * the instances' arrays concatenated in instance order (each in creation
 * order); b3g_get_stats and b3g_get_energy write one entry per instance.
 */
b3g_status b3g_body_build_batch(b3g_body body, const b3g_body_desc* descs, uint32_t count);

/* Number of instances of a built body (1 after b3g_body_build). */
b3g_status b3g_get_instance_count(b3g_body body, uint32_t* count);

/*
 * Topology edits: the body after b3Body::CreateParticle / CreateFixture /
 * CreateForce / Destroy* calls (body.cpp:62-354) between steps. The caller
 * passes the edited description (surviving objects in their creation order,
 * new ones after them, as the reference's lists hold them) and, per object of
 * the new description, the creation index of the object it continues in the
 * previous description, or -1 for a new object (any map may be NULL: nothing
 * of that kind carries over). Carried: particle positions and velocities
 * (from the device; new particles take the description's), tetrahedron
 * warm-start quaternions (tetrahedron_element_force.h:154), and the contacts
 * whose sphere and world fixture both carry over, with their manifold, normal
 * force and friction flag, in their list order. Everything else is rebuilt as
 * b3g_body_build_batch does; solver, reduction, graph and action-force
 * settings are kept. A single body passes count = 1. A map entry at or past
 * the previous description's count is rejected (B3G_INVALID_ARGUMENT) before
 * anything changes.
 */
typedef struct b3g_carry {
	const int64_t* particle_prev;   /* n_particles entries */
	const int64_t* tet_prev;        /* n_tet entries */
	const int64_t* sphere_prev;     /* n_spheres entries */
	const int64_t* world_prev;      /* n_world entries */
} b3g_carry;

b3g_status b3g_body_rebuild(b3g_body body, const b3g_body_desc* descs, uint32_t count, const b3g_carry* carry);

/*
 * Action forces: m_f1.. of b3SpringForce (GetActionForce / GetReactionForce,
 * spring_force.h:155-163), b3StretchForce and b3ShearForce (GetActionForce1..3)
 * and b3MouseForce (GetActionForce1..4): the force each applied to its
 * particles, summed over the Newton iterations of the last step. Off by
 * default (it costs one pass over those forces per iteration). get writes,
 * for the forces of `force_type` in creation order, 2 (spring), 3 (stretch,
 * shear) or 4 (mouse) vectors of 3 floats each, at most `capacity` floats.
 */
b3g_status b3g_set_action_forces(b3g_body body, int enabled);
b3g_status b3g_get_action_forces(b3g_body body, uint32_t force_type, float* out, uint64_t capacity);

/*
 * Replaces b3Body::RayCastSingle(output, p1, p2) (dynamics/body.cpp:497-521,
 * declared body.h:156) for `count` rays p1[i] -> p2[i] (3 floats each) against
 * the triangle fixtures of instance `instance` (0 for a single body), at the
 * current positions. Per ray: index = creation index of the triangle fixture
 * hit (the fixture b3BodyRayCastSingleOutput::triangle points to; in a batch,
 * an index into the instances' concatenated triangle-fixture arrays) or -1,
 * fraction and normal as b3TriangleShape::RayCast computes them
 * (triangle_shape.cpp:314-382); `normal` may be NULL. Among equal fractions
 * the lowest creation index wins (the reference takes whichever its tree
 * reports first). The reference keeps its triangle tree in step with the
 * particles after every b3Body::Step (body.cpp:584-600); here the step does no
 * tree work and the query builds the tree when particles have moved.
 */
b3g_status b3g_ray_cast(b3g_body body, uint32_t instance, uint32_t count, const float* p1, const float* p2,
                        int64_t* index, float* fraction, float* normal);

/*
 * b3g_ray_cast with every array in device memory (hipMalloc'd, on the body's
 * device): same arguments, results and ties, no host copies; the call blocks
 * until the results are written. An array that is not device memory is
 * rejected (B3G_INVALID_ARGUMENT). No counterpart in the reference, whose
 * rays live on the host; it serves callers that keep rays on the GPU.
 */
b3g_status b3g_ray_cast_device(b3g_body body, uint32_t instance, uint32_t count, const float* p1, const float* p2,
                               int64_t* index, float* fraction, float* normal);

/*
 * The voxel grid b3BuildSDF(sdf, mesh, cellSize, aabbVolumeExtension) creates
 * (collision/geometry/sdf_builder.cpp:155-174, declared sdf.h:53): the mesh's
 * box grown by `aabb_extension`, box[0..2] lower and box[3..5] upper bound,
 * dims = ceil(size / cell_size) + 1 voxels per axis. Host only.
 */
b3g_status b3g_sdf_grid(const float* vertices, uint32_t vertex_count, const float cell_size[3], float aabb_extension,
                        float box[6], uint32_t dims[3]);

/*
 * Replaces b3BuildSDF: the signed distance of every voxel of that grid to the
 * mesh (vertices: 3 floats each; triangles: `triangle_stride` >= 3 words each,
 * the vertex indices first, so b3Triangle arrays pass with stride 6), negative
 * inside, into `voxels` (W * H * D floats at i + j W + k W H, the layout of
 * b3VoxelGrid and of the descriptor's SDF voxels). The values are the
 * reference's bit for bit, including its inside test (5 parity rays,
 * sdf_builder.cpp:26-78). Computed on `device`; blocks until done.
 */
b3g_status b3g_build_sdf(int device, const float* vertices, uint32_t vertex_count, const uint32_t* triangles,
                         uint32_t triangle_count, uint32_t triangle_stride, const float cell_size[3],
                         float aabb_extension, float* voxels, uint64_t capacity);

/*
 * One time step. Replaces b3Body::Step(dt, forceIterations, forceSubIterations)
 * (body.h:153, body.cpp:552-604). tolerance / sub_tolerance are the Newton and
 * PCG tolerances of b3SolveBEInput (sparse_force_solver.h:55-61); a value <= 0
 * selects the reference default B3_EPSILON. Asynchronous.
 */
b3g_status b3g_step(b3g_body body, float dt, uint32_t iterations, uint32_t sub_iterations, float tolerance,
                    float sub_tolerance);

/* Positions / velocities, creation order (b3Particle::GetPosition / GetVelocity, particle.h:220-237). */
b3g_status b3g_get_state(b3g_body body, float* pos, float* vel);

/* Overwrites positions and/or velocities (NULL keeps); b3Particle::SetPosition / SetVelocity (particle.h:213-232). */
b3g_status b3g_set_state(b3g_body body, const float* pos, const float* vel);

/* Accumulates external forces / translations for the next step (b3Particle::ApplyForce / ApplyTranslation, particle.h:249-275). NULL skips. */
b3g_status b3g_apply_external(b3g_body body, const float* force, const float* translation);

/* Lumped masses (b3Particle::GetMass), creation order. */
b3g_status b3g_get_masses(b3g_body body, float* mass);

/* Solver statistics (the reference's b3_forceSolver* globals) and contact count, one entry per instance. */
b3g_status b3g_get_stats(b3g_body body, b3g_step_stats* stats);

/* Kinetic energy 0.5 * sum m v.v in list order, FP32 sequential (b3Body::GetEnergy, body.cpp:384-392), one value per instance. */
b3g_status b3g_get_energy(b3g_body body, float* energy);

/*
 * Contact list in list order (b3ContactManager::m_contactList): creation
 * index of the sphere fixture's particle, creation index of the world
 * fixture and the accumulated normal force. *count receives the number of
 * contacts; at most `capacity` entries are written. A batch lists each
 * instance's contacts in turn, instances in creation order.
 */
b3g_status b3g_get_contacts(b3g_body body, uint32_t* count, uint32_t* particle, uint32_t* world, float* normal_force,
                            uint32_t capacity);

/* Warm-start quaternions (x y z s) of the tetrahedron elements, creation order (tetrahedron_element_force.h:154). */
b3g_status b3g_get_tet_quats(b3g_body body, float* quats);

/*
 * Diagnostics: runs the contact update and the first backward-Euler
 * linearisation of a step (sparse_force_solver.cpp:146-174) on the current
 * state without advancing it, and keeps pA, pb and the internal force fi.
 * b3g_get_system copies them out: rows in list order with the columns of each
 * row in the reference's traversal order (head to tail), blocks column-major.
 * Tet warm starts are updated as ApplyForces does in the reference.
 */
b3g_status b3g_assemble(b3g_body body, float dt, uint32_t* nnz);
b3g_status b3g_get_system(b3g_body body, uint32_t* rowptr, uint32_t* cols, float* vals, float* pb, float* fi);

/*
 * Reduction of the PCG dot products and of the Newton error norm.
 *   B3G_REDUCE_FP64_SEPARATE: deterministic FP64 tree over the device; a
 *   small kernel per reduction forms alpha / beta / convergence.
 *   B3G_REDUCE_FP64_FUSED: the same sums in the same order, formed instead at
 *   the start of the kernel that consumes them: two launches fewer per PCG
 *   iteration, but every consuming workgroup re-reads the instance's partials
 *   (one per SpMV tile or row block), so the extra work grows with the square
 *   of the instance's size.
 *   B3G_REDUCE_FP64 (default): the fused form when every instance has at most
 *   6144 partials per reduction (about 480k rows), the separate form
 *   otherwise. The three FP64 modes give identical results bit for bit.
 *   B3G_REDUCE_SEQUENTIAL: FP32 terms added one by one in row order, as
 *   b3Dot / b3LengthSquared do (dense_vec3.h:151-175). With it a step is the
 *   reference's step operation for operation; it is O(rows) latency per
 *   reduction and meant for validation, not throughput.
 */
enum { B3G_REDUCE_FP64 = 0, B3G_REDUCE_SEQUENTIAL = 1, B3G_REDUCE_FP64_FUSED = 2, B3G_REDUCE_FP64_SEPARATE = 3 };
b3g_status b3g_set_reduction(b3g_body body, int mode);

/*
 * Checkpoint / resume of everything a step reads besides the description:
 * positions, velocities, pending external forces and translations, tet
 * warm-start quaternions, the contact pairs (with their frozen manifolds,
 * normal forces and list order) and the solver statistics. With blob == NULL,
 * *bytes receives the size. Loading requires a body built from the same
 * description. No reference counterpart (the reference has no resume).
 */
b3g_status b3g_save_state(b3g_body body, void* blob, uint64_t* bytes);
b3g_status b3g_load_state(b3g_body body, const void* blob, uint64_t bytes);

/*
 * Precision form of the linear solve (no counterpart in the reference).
 *   B3G_SOLVER_PCG (default): the reference's Jacobi PCG on the FP32 pA
 *   (sparse_solver.cpp:28-109).
 *   B3G_SOLVER_REFINED_FP32 / B3G_SOLVER_REFINED_BF16: iterative refinement.
 *   Each sweep forms the residual pb - pA x in FP64 (FP32 blocks, FP64 x and
 *   sums), then runs the reference's PCG for at most `inner_iterations` at
 *   relative tolerance `inner_tolerance` on pA dx = residual, with pA's
 *   off-diagonal blocks in FP32 or in BF16 (FP32 accumulation; diagonal
 *   blocks FP32), and adds dx to x in FP64. In these modes b3g_step's
 *   sub_tolerance is the target of ||residual|| relative to the first sweep's
 *   and sub_iterations caps the inner iterations of a Newton iteration
 *   (sweeps = ceil(sub_iterations / inner_iterations)).
 */
enum { B3G_SOLVER_PCG = 0, B3G_SOLVER_REFINED_FP32 = 1, B3G_SOLVER_REFINED_BF16 = 2 };
b3g_status b3g_set_solver(b3g_body body, int mode, uint32_t inner_iterations, float inner_tolerance);

/*
 * Captured steps (no counterpart in the reference): with enabled != 0,
 * b3g_step records its launch sequence into a HIP graph once per set of step
 * parameters and solver modes and replays it afterwards (one launch per
 * step). Results are identical to uncaptured steps.
 */
b3g_status b3g_set_graph(b3g_body body, int enabled);

/*
 * Waits for all work queued on the body. The reference asserts on a
 * non-positive Jacobi diagonal (sparse_solver.cpp:45-52) and has no check for
 * a diverging state; here the steps go on as the reference's release build
 * does, the device records the fault per instance (b3g_step_stats::faults),
 * and b3g_synchronize returns B3G_NON_FINITE or B3G_NON_POSITIVE_DIAGONAL
 * once for the steps since the previous b3g_synchronize that raised one
 * (b3g_last_error() names the first instance).
 */
b3g_status b3g_synchronize(b3g_body body);

/*
 * Roofline pass (no counterpart in the reference): `launches` back-to-back
 * launches of the PCG's block-sparse SpMV kernel -- one PCG iteration's full
 * work, q = pA d with d = s + beta d_old, on the body's current pA and its PCG
 * scratch vectors -- between one HIP event pair on the body's stream.
 * *mean_ms = elapsed / launches; *bytes_per_launch = the algorithmic bytes of
 * one launch (40 per 3x3 block + 4 per row offset + 24 per row for the vector
 * in/out; with the BF16 operator of B3G_SOLVER_REFINED_BF16, 22 per block +
 * 36 per row for the FP32 diagonal instead of 40 per block). Leaves
 * positions, velocities and the solver state unchanged.
 */
b3g_status b3g_time_spmv(b3g_body body, uint32_t launches, double* mean_ms, double* bytes_per_launch);

/*
 * The roofline pass from HBM (no counterpart in the reference): as
 * b3g_time_spmv, but before every SpMV launch a kernel streams reads over
 * `evict_bytes` (at least 512 MiB) of device memory, so the matrix and the
 * vectors are no longer in the L2 or the 256 MB memory-side cache (MALL)
 * and the launch reads them from HBM (reads leave no dirty lines whose
 * write-back would overlap the SpMV). The same number of eviction launches
 * alone is timed too; *mean_ms = (both - evictions alone) / launches, which
 * still includes one launch gap per SpMV. The eviction buffer is allocated
 * for the call and freed after it.
 */
b3g_status b3g_time_spmv_cold(b3g_body body, uint32_t launches, uint64_t evict_bytes, double* mean_ms,
                              double* bytes_per_launch);

/*
 * Early exit of the PCG (no counterpart in the reference's API; it is the
 * reference's own loop exit at sparse_solver.cpp:63-73): with enabled != 0
 * (the default) an uncaptured b3g_step enqueues the PCG iterations in chunks
 * and stops enqueuing once no instance's last iteration ran. Iterations past
 * convergence do nothing either way, so the results are identical; only
 * their launches are saved. The step's launch sequence is then enqueued by
 * a thread of the body, and b3g_step returns at once; every other function
 * of the body first waits until that step is fully enqueued. Captured steps
 * (b3g_set_graph) always record the fixed sequence.
 */
b3g_status b3g_set_early_exit(b3g_body body, int enabled);

/*
 * The PCG's launch schedule (no counterpart in the reference, which runs one
 * thread): the SpMV's tiles (at most tile_blocks 3x3 blocks and tile_rows
 * rows each, at most 1024 and 256), the rows per workgroup of the vector
 * kernels (row_block_rows, at most 1024) and where the FP64 dot products are
 * summed (fused: 1 in the consuming kernels, 0 in a kernel of their own, -1
 * by instance size as B3G_REDUCE_FP64 does). Every schedule computes each row
 * and block as the reference does; tiles and row blocks only group the FP64
 * partial sums of the dot products, whose rounding to FP32 then agrees in
 * all but rare ties. max_groups caps the workgroups of a PCG launch (0: one
 * per tile or row block); under a cap of at most half the tiles (row blocks)
 * each workgroup takes every G-th one and skips, 64 at a time, those of
 * converged instances. persist_tiles > 0 runs the PCG iterations of every
 * instance of at most that many SpMV tiles in one persistent launch beside
 * the launch sequence of the others: one workgroup works through a whole
 * solve of an instance, largest instances first, on persist_groups
 * workgroups (0: one per such instance); its tiles, row blocks and sums are
 * the launch sequence's, so the results are bit-identical. It serves the
 * plain solver in the FP64 reduction modes. Default: 1024, 256, 1024, -1, 0,
 * 0, 0.
 */
typedef struct b3g_schedule {
	uint32_t tile_blocks;
	uint32_t tile_rows;
	uint32_t row_block_rows;
	int32_t fused;
	uint32_t max_groups;
	uint32_t persist_tiles;
	uint32_t persist_groups;
} b3g_schedule;

b3g_status b3g_get_schedule(b3g_body body, b3g_schedule* schedule);
b3g_status b3g_set_schedule(b3g_body body, const b3g_schedule* schedule);

/*
 * Autotuner (no counterpart in the reference): times candidate schedules --
 * five SpMV tile shapes, three row-block sizes, both places of the FP64 sums,
 * then four workgroup caps for the best of those, then (a batch, after a
 * step) the persistent PCG for instances of at most 4, 16 or 64 tiles -- and
 * keeps the fastest; a candidate that beats the default is timed
 * against it twice more and kept only if it is faster by more than 1% both
 * times. After a
 * b3g_step each timing runs whole steps with that step's parameters from the
 * body's current state, which is saved first and loaded again at the end;
 * before any step, PCG iterations on the body's system. The assembled
 * system, the step statistics and the action forces of the last step are
 * kept. With cache_path != NULL the choice is looked up first, and stored
 * after a search, in that text file, keyed by a signature of the system's
 * shape (rows, blocks, rows per instance), the reduction mode and the
 * device; a hit sets the cached schedule without timing. *chosen receives
 * the schedule kept; *tuned_ms and *default_ms (either may be NULL) the
 * measured milliseconds per step (per PCG iteration before any step) with it
 * and with the default (those recorded in the cache after a hit). A step
 * that faults during a timing ends the search with its status.
 */
b3g_status b3g_autotune(b3g_body body, const char* cache_path, b3g_schedule* chosen, double* tuned_ms,
                        double* default_ms);

/*
 * Host-only diagnostics: the block pattern of pA for a description (list-order
 * CSR, columns of each row in the reference's traversal order). Needs no
 * device. *nnz receives the block count; cols is written when capacity >= nnz.
 */
b3g_status b3g_pattern(const b3g_body_desc* desc, uint32_t* nnz, uint32_t* rowptr, uint32_t* cols, uint32_t capacity);

/* Size of the linear system: rows (particles) and 3x3 blocks of pA. */
b3g_status b3g_get_system_size(b3g_body body, uint32_t* rows, uint32_t* blocks);

/*
 * Operator-level replacement of b3SparseSolveCG (sparse/sparse_solver.h:47,
 * sparse_solver.cpp:28-109): Jacobi-preconditioned CG on a block-sparse
 * matrix given as list-order CSR (rows head to tail, 3x3 blocks column-major,
 * host arrays). x holds the initial guess on entry and the solution on exit.
 */
b3g_status b3g_solve_cg(int device, uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals,
                        const float* b, float* x, uint32_t max_iterations, float tolerance, uint32_t* iterations);

/*
 * b3g_solve_cg with every array in device memory and no allocation: no
 * counterpart in the reference (whose b3SparseSolveCG allocates its
 * temporaries per call, sparse_solver.cpp:38-56). It is for callers that
 * cannot allocate device memory -- code being captured into a HIP graph,
 * tight serving loops, harnesses that forbid it. The solve is the same PCG
 * on the same kernels (Jacobi preconditioner, FP64-reduced dot products,
 * true residual every 50th iteration); its scratch is a fixed device array
 * of the library (about 5.05 MB per GPU, held until the process ends), which
 * limits it to B3G_CG_STATIC_MAX_ROWS rows and B3G_CG_STATIC_MAX_ITERATIONS
 * iterations (larger requests: B3G_INVALID_ARGUMENT).
 *   rowptr (n + 1) / cols (nnz): list-order CSR as for b3g_solve_cg;
 *   vals: the blocks as nine planes, entry c (column-major) of block k at
 *         vals[c * nnz + k] -- the layout the library keeps pA in;
 *   b, x: 3 n floats; x holds the initial guess on entry, the solution after;
 *   iterations (optional, device uint32): the iteration count, or
 *         B3G_CG_STATIC_INVALID when rowptr / cols are malformed (not
 *         non-decreasing from 0 to nnz, or a column >= n), in which case no
 *         PCG kernel reads the matrix and x is left unchanged.
 * Asynchronous on `stream` (a hipStream_t; NULL: the device's null stream);
 * nothing is synchronised, allocated or copied to the host. Every call on a
 * device uses that device's one scratch array, which keeps the last call's
 * contents: calls that overlap on one device (other streams or threads)
 * race and must be ordered by the caller.
 */
#define B3G_CG_STATIC_MAX_ROWS 65536u
#define B3G_CG_STATIC_MAX_ITERATIONS 4095u
#define B3G_CG_STATIC_INVALID 0xFFFFFFFFu
b3g_status b3g_solve_cg_static(int device, void* stream, uint32_t n, uint32_t nnz, const uint32_t* rowptr,
                               const uint32_t* cols, const float* vals, const float* b, float* x,
                               uint32_t max_iterations, float tolerance, uint32_t* iterations);

/* y = A x for a list-order CSR matrix (b3Mul, sparse/sparse_mat33.h:494-512), host arrays. */
b3g_status b3g_spmv(int device, uint32_t n, const uint32_t* rowptr, const uint32_t* cols, const float* vals,
                    const float* x, float* y);

/*
 * Operator-level replacement of b3SparseSolveBE (sparse/sparse_force_solver.h:97,
 * sparse_force_solver.cpp:104-232) behind its plugin, b3SparseForceModel
 * (sparse_force_solver.h:44-49). The caller's force model runs on the host
 * once per Newton iteration, as the reference calls it; everything after it
 * runs on the device: A = M - h dfdv - h^2 dfdx and the right-hand side
 * with the reference's list orders (every sparse copy reverses a row,
 * sparse_mat33.h:210-230), the pre-filter pA = S A S^T + I - S,
 * pb = S (b - A z) for any 3x3 diagonal blocks M and S, the Jacobi PCG
 * (sparse_solver.cpp:28-109, early exit included) warm-started across Newton
 * iterations, and the Newton update and error. With B3G_REDUCE_SEQUENTIAL the
 * results are the reference's bit for bit; the default FP64 reduction
 * differs from it as b3g_step's does.
 */

/* A b3SparseMat33 being filled by a force model (host memory, the reference's list semantics). */
typedef struct b3g_matrix_s* b3g_matrix;

/*
 * (*m)(rows[k], cols[k]) += blocks[9 k .. 9 k + 8] for k = 0 .. count - 1 in
 * order: b3SparseMat33::operator()(i, j) (search the row from its head,
 * insert a zero block at the head when absent; sparse_mat33.h:268-288), then
 * b3Mat33::operator+= on the column-major block. b3g_matrix_sub does -=.
 * Indices must be below the system's row count (B3G_INVALID_ARGUMENT).
 */
b3g_status b3g_matrix_add(b3g_matrix m, uint32_t count, const uint32_t* rows, const uint32_t* cols,
                          const float* blocks);
b3g_status b3g_matrix_sub(b3g_matrix m, uint32_t count, const uint32_t* rows, const uint32_t* cols,
                          const float* blocks);

/*
 * Replaces row `row` of m with `count` blocks, head to tail (the order in which
 * a b3SparseMat33 row lists them from its head, sparse_mat33.h:31-48), values
 * taken as given; columns must be distinct and below the row count. For an
 * integration that lets an unchanged force model fill the reference's own
 * b3SparseMat33 and then hands its rows over.
 */
b3g_status b3g_matrix_set_row(b3g_matrix m, uint32_t row, uint32_t count, const uint32_t* cols, const float* blocks);

/* b3SparseForceSolverData (sparse_force_solver.h:31-39): what the force model reads and fills. */
typedef struct b3g_force_data {
	float h, inv_h;
	uint32_t n;
	const float* x;     /* 3 n: the Newton iterate's positions */
	const float* v;     /* 3 n: velocities */
	float* f;           /* 3 n, zero on entry: the forces */
	b3g_matrix dfdx;    /* empty on entry */
	b3g_matrix dfdv;    /* empty on entry */
} b3g_force_data;

/* b3SparseForceModel::ApplyForces. A status other than B3G_OK ends the solve with that status. */
typedef b3g_status (*b3g_force_model)(void* user, const b3g_force_data* data);

/* b3SolveBEInput (sparse_force_solver.h:52-81); b3g_be_input_defaults sets its constructor's values. */
typedef struct b3g_be_input {
	float h, inv_h;
	b3g_force_model force_model;
	void* user;
	uint32_t n;                    /* dofCount */
	const float* x0;               /* 3 n */
	const float* v0;               /* 3 n */
	const float* fe;               /* 3 n: external forces over the step */
	const float* M;                /* 9 n: diagonal 3x3 blocks of the mass matrix, column-major */
	const float* y;                /* 3 n: translation */
	const float* S;                /* 9 n: diagonal 3x3 blocks of the constraint filter */
	const float* z;                /* 3 n: desired sub-solution (NULL: zero) */
	uint32_t max_iterations;       /* Newton */
	float tolerance;
	uint32_t max_sub_iterations;   /* PCG */
	float sub_tolerance;
} b3g_be_input;

/* b3SolveBEOutput (sparse_force_solver.h:84-93). min/max_sub_iterations are in/out, as in the reference:
 * the solve folds its PCG counts into the values they hold on entry. */
typedef struct b3g_be_output {
	float* x;                      /* 3 n */
	float* v;                      /* 3 n */
	uint32_t iterations;
	float error;
	uint32_t min_sub_iterations;
	uint32_t max_sub_iterations;
} b3g_be_output;

/* A device-side solver context: buffers and the symbolic pattern, kept between calls while the force model's
 * pattern does not change. One HIP stream; not to be used from two threads at once. */
typedef struct b3g_be_s* b3g_be;

b3g_status b3g_be_create(b3g_be* out, int device);
b3g_status b3g_be_destroy(b3g_be be);
/* B3G_REDUCE_* as b3g_set_reduction; B3G_REDUCE_SEQUENTIAL gives the reference's results bit for bit. */
b3g_status b3g_be_set_reduction(b3g_be be, int mode);
/* The b3SolveBEInput constructor (sparse_force_solver.h:56-62): 1 Newton iteration, 20 PCG iterations,
 * both tolerances FLT_EPSILON; every other field zero. */
b3g_status b3g_be_input_defaults(b3g_be_input* in);
/* Replaces b3SparseSolveBE(output, input). Blocks until the solve ends. */
b3g_status b3g_solve_be(b3g_be be, b3g_be_output* out, const b3g_be_input* in);

#ifdef __cplusplus
}
#endif

#endif
