"""Creation-ordered description of a soft body (host side).

:class:`SceneBuilder` records objects in the order a program would create
them through the reference API (``b3Body::CreateParticle``,
``CreateSphere``, ``CreateTriangle``, ``CreateTetrahedron``, ``CreateForce``,
``CreateFixture``; dynamics/body.cpp:62-354) with the parameters of the
reference's ``*Def`` structs. :meth:`SceneBuilder.arrays` returns flat numpy
arrays that feed the device library (:mod:`bounce_softbody_amd.body`).

The reference pushes every object at the head of its list, so it iterates
objects in reverse creation order (body.cpp:67-75, 114-122); the device
library receives creation order and applies the same reversal.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2                    # particle.h:31-36

FORCE_STRETCH, FORCE_SHEAR, FORCE_SPRING, FORCE_MOUSE, FORCE_TRI_ELEMENT, FORCE_TET_ELEMENT = 1, 2, 3, 4, 5, 6  # force.h:30-39

SHAPE_SPHERE, SHAPE_CAPSULE, SHAPE_TRIANGLE, SHAPE_BOX, SHAPE_MESH, SHAPE_SDF = 0, 1, 2, 3, 4, 5  # shape.h:42-51

WORLD_PARAMS = 32


def uv_initialize(A: np.ndarray, B: np.ndarray, C: np.ndarray) -> np.ndarray:
    """``b3StretchForceDef::Initialize`` (stretch_force.cpp:36-60), vectorised in float32.

    Returns (m, 6) = u1 v1 u2 v2 u3 v3. Each operation is an IEEE float32 op in
    the reference's order, so the result is bit-identical.
    """
    A, B, C = (np.asarray(x, f32).reshape(-1, 3) for x in (A, B, C))
    AB = (B - A).astype(f32)
    AC = (C - A).astype(f32)
    u2 = np.sqrt(((AB[:, 0] * AB[:, 0] + AB[:, 1] * AB[:, 1]) + AB[:, 2] * AB[:, 2]).astype(f32)).astype(f32)
    nAB = (AB / u2[:, None]).astype(f32)
    cx = (AB[:, 1] * AC[:, 2] - AB[:, 2] * AC[:, 1]).astype(f32)
    cy = (AB[:, 2] * AC[:, 0] - AB[:, 0] * AC[:, 2]).astype(f32)
    cz = (AB[:, 0] * AC[:, 1] - AB[:, 1] * AC[:, 0]).astype(f32)
    a2 = np.sqrt(((cx * cx + cy * cy) + cz * cz).astype(f32)).astype(f32)
    u3 = ((AC[:, 0] * nAB[:, 0] + AC[:, 1] * nAB[:, 1]) + AC[:, 2] * nAB[:, 2]).astype(f32)
    v3 = (a2 / u2).astype(f32)
    z = np.zeros_like(u2)
    return np.stack([z, z, u2, z, u3, v3], axis=1).astype(f32)


class SceneBuilder:
    """Accumulates a body description in creation order."""

    def __init__(self):
        self._p_type, self._p_pos, self._p_vel, self._p_damp = [], [], [], []
        self._s = []            # (particle, radius, friction) arrays
        self._tf = []           # (p3, v9, density, radius, friction)
        self._ttf = []          # (p4, v12, density, radius, friction)
        self._forces = []       # (type array, index array)
        self._tables = {t: [] for t in (FORCE_STRETCH, FORCE_SHEAR, FORCE_SPRING, FORCE_MOUSE,
                                         FORCE_TRI_ELEMENT, FORCE_TET_ELEMENT)}
        self._counts = {t: 0 for t in self._tables}
        self._world = []        # (type, params[32], friction, voxels|None)
        self.n_particles = 0
        self.gravity = np.zeros(3, dtype=f32)

    # ------------------------------------------------------------ particles
    def add_particles(self, positions, *, types=None, velocities=None, damping=0.0) -> np.ndarray:
        pos = np.asarray(positions, f32).reshape(-1, 3)
        n = len(pos)
        self._p_pos.append(pos)
        self._p_type.append(np.full(n, DYNAMIC, np.uint8) if types is None else np.asarray(types, np.uint8))
        self._p_vel.append(np.zeros((n, 3), f32) if velocities is None else np.asarray(velocities, f32).reshape(-1, 3))
        self._p_damp.append(np.broadcast_to(np.asarray(damping, f32), (n,)).copy())
        idx = np.arange(self.n_particles, self.n_particles + n)
        self.n_particles += n
        return idx

    def set_particle_types(self, indices, ptype: int):
        """``b3Particle::SetType`` applied after creation (particle.cpp:58-87)."""
        t = np.concatenate(self._p_type)
        v = np.concatenate(self._p_vel)
        t[np.asarray(indices)] = ptype
        if ptype == STATIC:
            v[np.asarray(indices)] = 0.0
        self._p_type, self._p_vel = [t], [v]

    # ------------------------------------------------------------ fixtures
    def add_spheres(self, particles, radius, friction):
        p = np.asarray(particles, np.int64).reshape(-1)
        self._s.append((p, np.broadcast_to(f32(radius), p.shape).astype(f32),
                        np.broadcast_to(f32(friction), p.shape).astype(f32)))

    def add_triangle_fixtures(self, p3, rest9, *, density, radius, friction):
        p3 = np.asarray(p3, np.int64).reshape(-1, 3)
        m = len(p3)
        self._tf.append((p3, np.asarray(rest9, f32).reshape(m, 9), np.full(m, density, f32),
                         np.full(m, radius, f32), np.full(m, friction, f32)))

    def add_tet_fixtures(self, p4, rest12, *, density, radius, friction):
        p4 = np.asarray(p4, np.int64).reshape(-1, 4)
        m = len(p4)
        self._ttf.append((p4, np.asarray(rest12, f32).reshape(m, 12), np.broadcast_to(np.asarray(density, f32), (m,)).copy(),
                          np.full(m, radius, f32), np.full(m, friction, f32)))

    # ------------------------------------------------------------ forces
    def _add_force_rows(self, ftype: int, rows: dict) -> np.ndarray:
        m = len(next(iter(rows.values())))
        start = self._counts[ftype]
        self._tables[ftype].append(rows)
        self._counts[ftype] += m
        return np.arange(start, start + m)

    def add_tet_elements(self, p4, rest12, *, young, poisson, stiffness_damping=0.0):
        """``b3TetrahedronElementForceDef`` (tetrahedron_element_force.h:28-72)."""
        p4 = np.asarray(p4, np.int64).reshape(-1, 4)
        m = len(p4)
        prm = np.stack([np.broadcast_to(np.asarray(young, f32), (m,)), np.broadcast_to(np.asarray(poisson, f32), (m,)),
                        np.broadcast_to(np.asarray(stiffness_damping, f32), (m,))], 1).astype(f32)
        idx = self._add_force_rows(FORCE_TET_ELEMENT, {"p": p4, "v": np.asarray(rest12, f32).reshape(m, 12), "q": prm})
        self._forces.append((np.full(m, FORCE_TET_ELEMENT, np.uint32), idx))
        return idx

    def add_tets_interleaved(self, p4, rest12, *, density, radius, friction, young, poisson, stiffness_damping):
        self.add_tet_fixtures(p4, rest12, density=density, radius=radius, friction=friction)
        self.add_tet_elements(p4, rest12, young=young, poisson=poisson, stiffness_damping=stiffness_damping)

    def add_tri_elements(self, p3, rest9, *, ex, ey, es, nuxy, nuyx, stiffness_damping=0.0):
        """``b3TriangleElementForceDef`` (triangle_element_force.h)."""
        p3 = np.asarray(p3, np.int64).reshape(-1, 3)
        m = len(p3)
        prm = np.tile(np.array([ex, ey, es, nuxy, nuyx, stiffness_damping], f32), (m, 1))
        idx = self._add_force_rows(FORCE_TRI_ELEMENT, {"p": p3, "v": np.asarray(rest9, f32).reshape(m, 9), "q": prm})
        self._forces.append((np.full(m, FORCE_TRI_ELEMENT, np.uint32), idx))
        return idx

    def add_stretch(self, p3, uv6, *, ks_u, kd_u=0.0, b_u=1.0, ks_v, kd_v=0.0, b_v=1.0):
        """``b3StretchForceDef`` (stretch_force.h); ``uv6`` from :func:`uv_initialize`."""
        p3 = np.asarray(p3, np.int64).reshape(-1, 3)
        m = len(p3)
        prm = np.concatenate([np.asarray(uv6, f32).reshape(m, 6),
                              np.tile(np.array([ks_u, kd_u, b_u, ks_v, kd_v, b_v], f32), (m, 1))], 1)
        idx = self._add_force_rows(FORCE_STRETCH, {"p": p3, "q": prm})
        self._forces.append((np.full(m, FORCE_STRETCH, np.uint32), idx))
        return idx

    def add_shear(self, p3, uv6, *, ks, kd=0.0):
        """``b3ShearForceDef`` (shear_force.h)."""
        p3 = np.asarray(p3, np.int64).reshape(-1, 3)
        m = len(p3)
        prm = np.concatenate([np.asarray(uv6, f32).reshape(m, 6), np.tile(np.array([ks, kd], f32), (m, 1))], 1)
        idx = self._add_force_rows(FORCE_SHEAR, {"p": p3, "q": prm})
        self._forces.append((np.full(m, FORCE_SHEAR, np.uint32), idx))
        return idx

    def add_springs(self, p2, *, length, ks, kd=0.0):
        """``b3SpringForceDef`` (spring_force.h)."""
        p2 = np.asarray(p2, np.int64).reshape(-1, 2)
        m = len(p2)
        prm = np.stack([np.broadcast_to(np.asarray(length, f32), (m,)), np.full(m, ks, f32), np.full(m, kd, f32)], 1)
        idx = self._add_force_rows(FORCE_SPRING, {"p": p2, "q": prm.astype(f32)})
        self._forces.append((np.full(m, FORCE_SPRING, np.uint32), idx))
        return idx

    def add_mouse(self, p4, *, w2, w3, w4, length, ks, kd=0.0):
        """``b3MouseForceDef`` (mouse_force.h)."""
        p4 = np.asarray(p4, np.int64).reshape(-1, 4)
        m = len(p4)
        prm = np.tile(np.array([w2, w3, w4, length, ks, kd], f32), (m, 1))
        idx = self._add_force_rows(FORCE_MOUSE, {"p": p4, "q": prm})
        self._forces.append((np.full(m, FORCE_MOUSE, np.uint32), idx))
        return idx

    def add_cloth_triangles(self, p3, rest9, *, density, radius, friction, create_elements, element, stretch, shear):
        """Per triangle: fixture, then element or stretch (+ optional shear) force (uniform_body.cpp:51-119)."""
        p3 = np.asarray(p3, np.int64).reshape(-1, 3)
        rest9 = np.asarray(rest9, f32).reshape(-1, 9)
        m = len(p3)
        self.add_triangle_fixtures(p3, rest9, density=density, radius=radius, friction=friction)
        if create_elements:
            ex, ey, es, nxy, nyx, kd = element
            self.add_tri_elements(p3, rest9, ex=ex, ey=ey, es=es, nuxy=nxy, nuyx=nyx, stiffness_damping=kd)
            return
        uv = uv_initialize(rest9[:, 0:3], rest9[:, 3:6], rest9[:, 6:9])
        ks_u, kd_u, b_u, ks_v, kd_v, b_v = stretch
        if shear[0] <= 0 and shear[1] <= 0:
            self.add_stretch(p3, uv, ks_u=ks_u, kd_u=kd_u, b_u=b_u, ks_v=ks_v, kd_v=kd_v, b_v=b_v)
            return
        # Interleave stretch_k, shear_k per triangle.
        prm_st = np.concatenate([uv, np.tile(np.array([ks_u, kd_u, b_u, ks_v, kd_v, b_v], f32), (m, 1))], 1)
        prm_sh = np.concatenate([uv, np.tile(np.array([shear[0], shear[1]], f32), (m, 1))], 1)
        ist = self._add_force_rows(FORCE_STRETCH, {"p": p3, "q": prm_st})
        ish = self._add_force_rows(FORCE_SHEAR, {"p": p3, "q": prm_sh})
        types = np.empty(2 * m, np.uint32)
        types[0::2], types[1::2] = FORCE_STRETCH, FORCE_SHEAR
        idx = np.empty(2 * m, np.int64)
        idx[0::2], idx[1::2] = ist, ish
        self._forces.append((types, idx))

    # ------------------------------------------------------------ world
    def _add_world(self, stype, params, friction, voxels=None, mesh=None):
        q = np.zeros(WORLD_PARAMS, f32)
        q[: len(params)] = np.asarray(params, f32)
        self._world.append((stype, q, f32(friction), None if voxels is None else np.ascontiguousarray(voxels, f32).reshape(-1),
                            mesh))
        return len(self._world) - 1

    def add_world_sphere(self, center, radius, friction=0.5):
        return self._add_world(SHAPE_SPHERE, [radius, *center], friction)

    def add_world_capsule(self, c1, c2, radius, friction=0.5):
        return self._add_world(SHAPE_CAPSULE, [radius, *c1, *c2], friction)

    def add_world_triangle(self, v1, v2, v3, radius=0.0, friction=0.5, e=(None, None, None)):
        prm = [radius, *v1, *v2, *v3] + [0.0 if x is None else 1.0 for x in e]
        for x in e:
            prm += list(x) if x is not None else [0.0, 0.0, 0.0]
        return self._add_world(SHAPE_TRIANGLE, prm, friction)

    def add_world_box(self, extents, translation=(0, 0, 0), rotation=(0, 0, 0, 1), radius=0.0, friction=0.5):
        return self._add_world(SHAPE_BOX, [radius, *extents, *translation, *rotation], friction)

    def add_world_sdf(self, voxels, lower, upper, translation=(0, 0, 0), rotation=(0, 0, 0, 1), radius=0.0,
                      invert=False, friction=0.5):
        """``voxels`` has shape (D, H, W): index i + j*W + k*W*H (index_3d.h:41-46)."""
        vox = np.asarray(voxels, f32)
        D, H, W = vox.shape
        prm = [radius, *lower, *upper, W, H, D, *translation, *rotation, 1.0 if invert else 0.0]
        return self._add_world(SHAPE_SDF, prm, friction, vox)

    def add_world_mesh(self, vertices, triangles, scale=(1, 1, 1), translation=(0, 0, 0), rotation=(0, 0, 0, 1),
                       radius=0.0, friction=0.5, adjacency=True):
        """``b3MeshShape`` over a ``b3Mesh`` (mesh.h:52-87): vertices (n, 3); triangles (m, 3) vertex indices, or
        (m, 6) with the edge wing vertices u1 u2 u3 (``NULL_VERTEX``: open edge). With (m, 3) triangles the wings
        come from :func:`mesh_adjacency` when ``adjacency`` (``b3Mesh::BuildAdjacency``), else all edges are open.
        The arrays are passed by reference, so fixtures given the same arrays share one device copy."""
        v = np.ascontiguousarray(vertices, f32).reshape(-1, 3)
        t = np.asarray(triangles, np.uint32)
        if t.shape[1] == 3:
            wings = mesh_adjacency(t) if adjacency else np.full(t.shape, NULL_VERTEX, np.uint32)
            t = np.concatenate([t, wings], axis=1)
        t = np.ascontiguousarray(t, np.uint32)
        return self._add_world(SHAPE_MESH, [radius, *scale, *translation, *rotation], friction, mesh=(v, t))

    # ------------------------------------------------------------ output
    def arrays(self) -> dict:
        cat = lambda xs, shape, dt: np.ascontiguousarray(np.concatenate(xs), dt) if xs else np.zeros(shape, dt)
        out = {
            "p_type": cat(self._p_type, (0,), np.uint8),
            "p_pos": cat(self._p_pos, (0, 3), f32).reshape(-1),
            "p_vel": cat(self._p_vel, (0, 3), f32).reshape(-1),
            "p_damping": cat(self._p_damp, (0,), f32),
            "s_particle": cat([s[0] for s in self._s], (0,), np.uint32),
            "s_radius": cat([s[1] for s in self._s], (0,), f32),
            "s_friction": cat([s[2] for s in self._s], (0,), f32),
            "tf_p": cat([t[0] for t in self._tf], (0, 3), np.uint32).reshape(-1),
            "tf_v": cat([t[1] for t in self._tf], (0, 9), f32).reshape(-1),
            "tf_density": cat([t[2] for t in self._tf], (0,), f32),
            "tf_radius": cat([t[3] for t in self._tf], (0,), f32),
            "tf_friction": cat([t[4] for t in self._tf], (0,), f32),
            "ttf_p": cat([t[0] for t in self._ttf], (0, 4), np.uint32).reshape(-1),
            "ttf_v": cat([t[1] for t in self._ttf], (0, 12), f32).reshape(-1),
            "ttf_density": cat([t[2] for t in self._ttf], (0,), f32),
            "ttf_radius": cat([t[3] for t in self._ttf], (0,), f32),
            "ttf_friction": cat([t[4] for t in self._ttf], (0,), f32),
            "f_type": cat([f[0] for f in self._forces], (0,), np.uint32),
            "f_index": cat([f[1] for f in self._forces], (0,), np.uint32),
            "gravity": np.asarray(self.gravity, f32).copy(),
        }
        spec = {FORCE_TET_ELEMENT: ("tet", {"p": 4, "v": 12, "q": 3}), FORCE_TRI_ELEMENT: ("tri", {"p": 3, "v": 9, "q": 6}),
                FORCE_STRETCH: ("st", {"p": 3, "q": 12}), FORCE_SHEAR: ("sh", {"p": 3, "q": 8}),
                FORCE_SPRING: ("sp", {"p": 2, "q": 3}), FORCE_MOUSE: ("mo", {"p": 4, "q": 6})}
        for ftype, (name, cols) in spec.items():
            rows = self._tables[ftype]
            for key, width in cols.items():
                dt = np.uint32 if key == "p" else f32
                suffix = {"p": "_p", "v": "_v", "q": "_param"}[key]
                out[name + suffix] = cat([r[key] for r in rows], (0, width), dt).reshape(-1)
        out["w_type"] = np.array([w[0] for w in self._world], np.uint32)
        out["w_param"] = (np.stack([w[1] for w in self._world]) if self._world else np.zeros((0, WORLD_PARAMS), f32)).reshape(-1)
        out["w_friction"] = np.array([w[2] for w in self._world], f32)
        out["w_voxels"] = [w[3] for w in self._world]
        out["w_mesh_vertices"] = [None if w[4] is None else w[4][0].reshape(-1) for w in self._world]
        out["w_mesh_triangles"] = [None if w[4] is None else w[4][1].reshape(-1) for w in self._world]
        out["w_mesh_counts"] = np.array([[0, 0] if w[4] is None else [len(w[4][0]), len(w[4][1])] for w in self._world],
                                        np.uint32).reshape(-1)
        return out


NULL_VERTEX = 0xFFFFFFFF


def mesh_adjacency(triangles) -> np.ndarray:
    """Wing vertex of each triangle edge (mesh.cpp:41-98, b3Mesh::BuildAdjacency): for edge j (vertex j -> j + 1)
    of triangle i1, the third vertex of the first triangle i2 != i1 (in index order, then edge order) that has
    the reversed edge; NULL_VERTEX when there is none."""
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    first = {}
    for i2 in range(len(t)):
        for j2 in range(3):
            key = (int(t[i2, j2]), int(t[i2, (j2 + 1) % 3]))
            first.setdefault(key, []).append((i2, int(t[i2, (j2 + 2) % 3])))
    out = np.full(t.shape, NULL_VERTEX, np.uint32)
    for i1 in range(len(t)):
        for j1 in range(3):
            for i2, n2 in first.get((int(t[i1, (j1 + 1) % 3]), int(t[i1, j1])), []):
                if i2 != i1:
                    out[i1, j1] = n2
                    break
    return out


def edit_arrays(arrays: dict, edits):
    """A description after topology edits, as the reference's body would hold it after the same calls
    (body.cpp:62-354), and the maps from the new objects to the ones they continue.

    Edit records, applied in order (creation indices as in ``arrays``; a new object takes the next index):
      ("destroy_force", force_type, index_in_type)     b3Body::DestroyForce
      ("create_particle", type, position, velocity, damping)
      ("create_sphere", particle, radius, friction)
      ("create_spring", p1, p2, length, ks, kd)
    Returns (new arrays, carry) with carry = {"particle_prev", "tet_prev", "sphere_prev", "world_prev"}: per new
    object its creation index in ``arrays``, or -1 for an object the edits created."""
    a = {k: (v.copy() if isinstance(v, np.ndarray) else list(v) if isinstance(v, list) else v) for k, v in arrays.items()}
    n0, t0, s0, w0 = len(a["p_type"]), len(a["tet_param"]) // 3, len(a["s_particle"]), len(a["w_type"])
    particle_prev = list(range(n0))
    tet_prev = list(range(t0))
    widths = {FORCE_TET_ELEMENT: ("tet", {"_p": 4, "_v": 12, "_param": 3}), FORCE_TRI_ELEMENT: ("tri", {"_p": 3, "_v": 9, "_param": 6}),
              FORCE_STRETCH: ("st", {"_p": 3, "_param": 12}), FORCE_SHEAR: ("sh", {"_p": 3, "_param": 8}),
              FORCE_SPRING: ("sp", {"_p": 2, "_param": 3}), FORCE_MOUSE: ("mo", {"_p": 4, "_param": 6})}
    for e in edits:
        kind = e[0]
        if kind == "destroy_force":
            ftype, k = e[1], e[2]
            name, cols = widths[ftype]
            for suffix, w in cols.items():
                arr = a[name + suffix].reshape(-1, w)
                a[name + suffix] = np.ascontiguousarray(np.delete(arr, k, axis=0).reshape(-1))
            keep = ~((a["f_type"] == ftype) & (a["f_index"] == k))
            a["f_type"], a["f_index"] = a["f_type"][keep], a["f_index"][keep].copy()
            later = (a["f_type"] == ftype) & (a["f_index"] > k)
            a["f_index"][later] -= 1
            if ftype == FORCE_TET_ELEMENT:
                del tet_prev[k]
        elif kind == "create_particle":
            a["p_type"] = np.append(a["p_type"], np.uint8(e[1]))
            a["p_pos"] = np.append(a["p_pos"], np.asarray(e[2], f32))
            a["p_vel"] = np.append(a["p_vel"], np.asarray(e[3], f32))
            a["p_damping"] = np.append(a["p_damping"], f32(e[4]))
            particle_prev.append(-1)
        elif kind == "create_sphere":
            a["s_particle"] = np.append(a["s_particle"], np.uint32(e[1]))
            a["s_radius"] = np.append(a["s_radius"], f32(e[2]))
            a["s_friction"] = np.append(a["s_friction"], f32(e[3]))
        elif kind == "create_spring":
            a["sp_p"] = np.append(a["sp_p"], np.uint32([e[1], e[2]]))
            a["sp_param"] = np.append(a["sp_param"], f32([e[3], e[4], e[5]]))
            a["f_type"] = np.append(a["f_type"], np.uint32(FORCE_SPRING))
            a["f_index"] = np.append(a["f_index"], np.uint32(len(a["sp_param"]) // 3 - 1))
        else:
            raise ValueError(kind)
    ns = len(a["s_particle"])
    carry = {"particle_prev": np.array(particle_prev, np.int64), "tet_prev": np.array(tet_prev, np.int64),
             "sphere_prev": np.array(list(range(s0)) + [-1] * (ns - s0), np.int64),
             "world_prev": np.arange(w0, dtype=np.int64)}
    return a, carry
