/*
 * Flat, creation-ordered description of a soft body scene.
 *
 * TEST INFRASTRUCTURE. This header is shared by the oracle (cpu_ref) and by
 * the reference harness (ref_harness.cpp, built into oracle/_ref/). Both
 * consume the same arrays so that parity runs start from identical inputs.
 *
 * Objects are listed in the order a user would create them through the
 * reference API (b3Body::CreateParticle / CreateForce / CreateSphere /
 * CreateTriangle / CreateTetrahedron / CreateFixture, body.cpp:62-354).
 * The reference pushes every object at the head of its list, so its
 * iteration ("list") order is the reverse of this creation order
 * (body.cpp:67-75, 114-122).
 */
#ifndef B3S_SCENE_H
#define B3S_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Force type codes: same values as b3ForceType (force.h:30-39). */
enum {
	B3S_FORCE_STRETCH = 1,
	B3S_FORCE_SHEAR = 2,
	B3S_FORCE_SPRING = 3,
	B3S_FORCE_MOUSE = 4,
	B3S_FORCE_TRI_ELEMENT = 5,
	B3S_FORCE_TET_ELEMENT = 6
};

/* World shape codes: same values as b3Shape::Type (shape.h:42-51). */
enum {
	B3S_SHAPE_SPHERE = 0,
	B3S_SHAPE_CAPSULE = 1,
	B3S_SHAPE_TRIANGLE = 2,
	B3S_SHAPE_BOX = 3,
	B3S_SHAPE_MESH = 4,
	B3S_SHAPE_SDF = 5
};

#define B3S_WORLD_PARAMS 32

/*
 * World fixture parameter block (B3S_WORLD_PARAMS floats), [0] is always the
 * shape radius (b3Shape::m_radius):
 *   sphere   [1..3] center
 *   capsule  [1..3] center1, [4..6] center2
 *   triangle [1..9] v1 v2 v3, [10..12] hasE1..hasE3 (0/1), [13..21] e1 e2 e3
 *   box      [1..3] extents, [4..6] translation, [7..10] quaternion (x y z s)
 *   sdf      [1..3] aabb lower, [4..6] aabb upper, [7..9] voxel counts W H D,
 *            [10..12] translation, [13..16] quaternion (x y z s), [17] invert
 *   mesh     [1..3] scale, [4..6] translation, [7..10] quaternion (x y z s);
 *            vertices / triangles (v1 v2 v3 u1 u2 u3) in w_mesh_*
 */

typedef struct b3s_scene {
	uint32_t n_particles;
	const uint8_t* p_type;     /* 0 static, 1 kinematic, 2 dynamic (particle.h:31-36) */
	const float* p_pos;        /* 3 per particle */
	const float* p_vel;        /* 3 per particle */
	const float* p_damping;    /* 1 per particle */

	uint32_t n_spheres;        /* sphere fixtures (contact proxies) */
	const uint32_t* s_particle;
	const float* s_radius;
	const float* s_friction;

	uint32_t n_trifix;         /* triangle fixtures (mass + proxies) */
	const uint32_t* tf_p;      /* 3 per fixture */
	const float* tf_v;         /* 9 per fixture: rest vertices */
	const float* tf_density;
	const float* tf_radius;
	const float* tf_friction;

	uint32_t n_tetfix;         /* tetrahedron fixtures (mass) */
	const uint32_t* ttf_p;     /* 4 per fixture */
	const float* ttf_v;        /* 12 per fixture */
	const float* ttf_density;
	const float* ttf_radius;
	const float* ttf_friction;

	uint32_t n_forces;         /* forces in creation order */
	const uint32_t* f_type;    /* B3S_FORCE_* */
	const uint32_t* f_index;   /* index into the per-type table */

	uint32_t n_tet;            /* b3TetrahedronElementForceDef */
	const uint32_t* tet_p;     /* 4 */
	const float* tet_v;        /* 12: rest v1..v4 */
	const float* tet_param;    /* 3: youngModulus poissonRatio stiffnessDamping */

	uint32_t n_tri;            /* b3TriangleElementForceDef */
	const uint32_t* tri_p;     /* 3 */
	const float* tri_v;        /* 9 */
	const float* tri_param;    /* 6: Ex Ey Es nu_xy nu_yx stiffnessDamping */

	uint32_t n_stretch;        /* b3StretchForceDef */
	const uint32_t* st_p;      /* 3 */
	const float* st_param;     /* 12: u1 v1 u2 v2 u3 v3 ks_u kd_u b_u ks_v kd_v b_v */

	uint32_t n_shear;          /* b3ShearForceDef */
	const uint32_t* sh_p;      /* 3 */
	const float* sh_param;     /* 8: u1 v1 u2 v2 u3 v3 ks kd */

	uint32_t n_spring;         /* b3SpringForceDef */
	const uint32_t* sp_p;      /* 2 */
	const float* sp_param;     /* 3: length ks kd */

	uint32_t n_mouse;          /* b3MouseForceDef */
	const uint32_t* mo_p;      /* 4 */
	const float* mo_param;     /* 6: w2 w3 w4 length ks kd */

	uint32_t n_world;          /* world fixtures, creation order */
	const uint32_t* w_type;    /* B3S_SHAPE_* */
	const float* w_param;      /* B3S_WORLD_PARAMS per fixture */
	const float* w_friction;
	const float* const* w_voxels; /* per fixture: W*H*D voxel values for SDF shapes, else NULL */

	float gravity[3];

	/* mesh world fixtures, per fixture (NULL / 0 for other shapes) */
	const float* const* w_mesh_vertices;      /* 3 per vertex */
	const uint32_t* const* w_mesh_triangles;  /* 6 per triangle: v1 v2 v3 u1 u2 u3 (0xFFFFFFFF: open edge) */
	const uint32_t* w_mesh_counts;            /* 2 per fixture: vertices, triangles */
} b3s_scene;

#ifdef __cplusplus
}
#endif

#endif
