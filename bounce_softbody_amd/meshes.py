"""Grid meshes and the uniform-body builder of the reference testbed.

The connectivity and vertex numbering follow ``GridTetMesh`` and
``GridClothMesh`` (examples/testbed/framework/uniform_body.h:133-542), and
:func:`tet_body` / :func:`cloth_body` create objects in the order of
``UniformBody`` (examples/testbed/framework/uniform_body.cpp:23-232), so a body
built here is the body the testbed would build from the same mesh.

All positions are float32 and every arithmetic step is done in float32, as in
the reference (``b3Vec3 position(j, i, k); position -= center`` etc.).
"""

from __future__ import annotations

import numpy as np

from .scene import SceneBuilder, STATIC, DYNAMIC

f32 = np.float32


# ---------------------------------------------------------------- randomness
def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` uniforms in [0, 1) from the splitmix64 sequence (double precision)."""
    mask = (1 << 64) - 1
    out = np.empty(count, dtype=np.float64)
    state = seed & mask
    for i in range(count):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def splitmix64_np(seed: int, count: int) -> np.ndarray:
    """Vectorised splitmix64 (same values as :func:`splitmix64`)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        state = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# ---------------------------------------------------------------- meshes
class Mesh:
    """Vertices (float32, n x 3), triangles (t x 3) and tetrahedra (m x 4)."""

    def __init__(self, vertices, triangles, tetrahedra):
        self.vertices = np.ascontiguousarray(vertices, dtype=f32)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int64).reshape(-1, 3)
        self.tetrahedra = np.ascontiguousarray(tetrahedra, dtype=np.int64).reshape(-1, 4)

    @property
    def vertex_count(self) -> int:
        return len(self.vertices)

    def scale(self, s) -> "Mesh":           # uniform_body.h:58-69
        self.vertices = (self.vertices * np.asarray(s, dtype=f32)).astype(f32)
        return self

    def translate(self, t) -> "Mesh":       # uniform_body.h:71-77
        self.vertices = (self.vertices + np.asarray(t, dtype=f32)).astype(f32)
        return self

    def jitter(self, amplitude: float, seed: int) -> "Mesh":
        """Adds U(-amplitude, amplitude) per coordinate (splitmix64 stream from ``seed``)."""
        u = splitmix64_np(seed, self.vertices.size).reshape(self.vertices.shape)
        d = (-amplitude + 2.0 * amplitude * u).astype(f32)
        self.vertices = (self.vertices + d).astype(f32)
        return self


class GridTetMesh(Mesh):
    """``GridTetMesh<H, W, D>`` (uniform_body.h:220-542): 5 tetrahedra per cell."""

    def __init__(self, H: int, W: int, D: int):
        self.H, self.W, self.D = H, W, D
        i, j, k = np.meshgrid(np.arange(H + 1), np.arange(W + 1), np.arange(D + 1), indexing="ij")
        center = np.array([f32(0.5) * f32(W), f32(0.5) * f32(H), f32(0.5) * f32(D)], dtype=f32)
        pos = np.stack([j, i, k], axis=-1).reshape(-1, 3).astype(f32) - center
        idx = self.vertex(i, j, k).reshape(-1)
        verts = np.empty_like(pos)
        verts[idx] = pos
        tris = self._boundary(H, W, D)
        tets = self._tets(H, W, D)
        super().__init__(verts, tris, tets)

    def vertex(self, i, j, k):
        return k + (self.D + 1) * (j + (self.W + 1) * i)

    def _boundary(self, H, W, D):
        v = self.vertex
        out = []
        for i in range(H):               # x-y plane
            for j in range(W):
                v1, v2, v3, v4 = v(i, j, 0), v(i + 1, j, 0), v(i + 1, j + 1, 0), v(i, j + 1, 0)
                out += [(v1, v2, v3), (v3, v4, v1)]
                v1, v2, v3, v4 = v(i, j, D), v(i + 1, j, D), v(i + 1, j + 1, D), v(i, j + 1, D)
                out += [(v3, v2, v1), (v1, v4, v3)]
        for i in range(H):               # y-z plane
            for k in range(D):
                v1, v2, v3, v4 = v(i, 0, k), v(i, 0, k + 1), v(i + 1, 0, k + 1), v(i + 1, 0, k)
                out += [(v1, v2, v3), (v3, v4, v1)]
                v1, v2, v3, v4 = v(i, W, k), v(i, W, k + 1), v(i + 1, W, k + 1), v(i + 1, W, k)
                out += [(v3, v2, v1), (v1, v4, v3)]
        for j in range(W):               # x-z plane
            for k in range(D):
                v1, v2, v3, v4 = v(0, j, k), v(0, j + 1, k), v(0, j + 1, k + 1), v(0, j, k + 1)
                out += [(v1, v2, v3), (v3, v4, v1)]
                v1, v2, v3, v4 = v(H, j, k), v(H, j + 1, k), v(H, j + 1, k + 1), v(H, j, k + 1)
                out += [(v3, v2, v1), (v1, v4, v3)]
        return np.array(out, dtype=np.int64).reshape(-1, 3)

    def _tets(self, H, W, D):
        i, j, k = np.meshgrid(np.arange(H), np.arange(W), np.arange(D), indexing="ij")
        i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
        v = self.vertex
        v1, v2, v3, v4 = v(i, j, k), v(i, j, k + 1), v(i + 1, j, k + 1), v(i + 1, j, k)
        v5, v6, v7, v8 = v(i, j + 1, k), v(i, j + 1, k + 1), v(i + 1, j + 1, k + 1), v(i + 1, j + 1, k)
        odd = ((i + j + k) % 2 == 1)[:, None, None]
        a = np.stack([np.stack(t, -1) for t in ((v2, v6, v7, v5), (v5, v7, v4, v8), (v2, v4, v7, v3),
                                               (v2, v5, v4, v1), (v2, v7, v4, v5))], 1)
        b = np.stack([np.stack(t, -1) for t in ((v6, v1, v3, v2), (v6, v8, v1, v5), (v6, v3, v8, v7),
                                               (v1, v8, v3, v4), (v6, v1, v8, v3))], 1)
        return np.where(odd, a, b).reshape(-1, 4)


class GridClothMesh(Mesh):
    """``GridClothMesh<H, W>`` (uniform_body.h:133-218): (H+1) x (W+1) vertices."""

    def __init__(self, H: int, W: int):
        self.H, self.W = H, W
        i, j = np.meshgrid(np.arange(H + 1), np.arange(W + 1), indexing="ij")
        center = np.array([f32(0.5) * f32(W), f32(0), f32(0.5) * f32(H)], dtype=f32)
        pos = np.stack([j, np.zeros_like(j), i], -1).reshape(-1, 3).astype(f32) - center
        ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
        ii, jj = ii.reshape(-1), jj.reshape(-1)
        v1, v2, v3, v4 = self.vertex(ii, jj), self.vertex(ii + 1, jj), self.vertex(ii + 1, jj + 1), self.vertex(ii, jj + 1)
        tris = np.stack([np.stack([v1, v2, v3], -1), np.stack([v3, v4, v1], -1)], 1).reshape(-1, 3)
        super().__init__(pos, tris, np.zeros((0, 4), dtype=np.int64))

    def vertex(self, i, j):
        return i * (self.W + 1) + j


# ---------------------------------------------------------------- bodies
def tet_body(mesh: Mesh, *, density=0.1, radius=0.02, friction=0.3, young=1000.0, poisson=0.3,
             stiffness_damping=0.0, damping=0.0, builder: SceneBuilder | None = None) -> SceneBuilder:
    """``UniformBody(const TetDef&)`` (uniform_body.cpp:122-232).

    Defaults are ``TetDef``'s (uniform_body.h:99-111). ``damping`` is the
    particle mass damping (``b3ParticleDef::damping``).
    """
    b = builder or SceneBuilder()
    base = b.n_particles
    n = mesh.vertex_count
    b.add_particles(mesh.vertices, types=np.full(n, DYNAMIC, np.uint8), damping=damping)
    b.add_spheres(base + np.arange(n), radius, friction)
    tri = base + mesh.triangles
    b.add_triangle_fixtures(tri, mesh.vertices[mesh.triangles].reshape(-1, 9), density=0.0, radius=radius,
                            friction=friction)
    tet = base + mesh.tetrahedra
    rest = mesh.vertices[mesh.tetrahedra].reshape(-1, 12)
    # Fixture then element, per tetrahedron (uniform_body.cpp:180-231).
    b.add_tets_interleaved(tet, rest, density=density, radius=radius, friction=friction,
                           young=young, poisson=poisson, stiffness_damping=stiffness_damping)
    return b


def cloth_body(mesh: Mesh, *, density=0.1, radius=0.0, friction=0.3, stretching_stiffness=1000.0,
               stretch_damping=0.0, shear_stiffness=0.0, shear_damping=0.0, create_elements=False,
               element_young=500.0, element_shear=500.0, element_poisson=0.3, element_damping=0.0,
               damping=0.0, builder: SceneBuilder | None = None) -> SceneBuilder:
    """``UniformBody(const ClothDef&)`` (uniform_body.cpp:23-120), defaults of ``ClothDef``.

    ``shear_stiffness > 0`` additionally creates one ``b3ShearForce`` per
    triangle right after its stretch force (the testbed's cloth scenes pair
    stretch with shear); 0 reproduces ``UniformBody`` exactly.
    """
    b = builder or SceneBuilder()
    base = b.n_particles
    n = mesh.vertex_count
    b.add_particles(mesh.vertices, types=np.full(n, DYNAMIC, np.uint8), damping=damping)
    b.add_spheres(base + np.arange(n), radius, friction)
    tri = base + mesh.triangles
    rest = mesh.vertices[mesh.triangles].reshape(-1, 9)
    b.add_cloth_triangles(tri, rest, density=density, radius=radius, friction=friction,
                          create_elements=create_elements, element=(element_young, element_young, element_shear,
                                                                    element_poisson, element_poisson, element_damping),
                          stretch=(stretching_stiffness, stretch_damping, 1.0, stretching_stiffness,
                                   stretch_damping, 1.0),
                          shear=(shear_stiffness, shear_damping))
    return b
