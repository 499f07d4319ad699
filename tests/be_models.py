"""Force models and problems for the operator-level backward-Euler tests (b3SparseSolveBE with a
b3SparseForceModel, sparse_force_solver.h:44-97).

A model here is a callable ``model(x, v) -> (f, ops)``: the forces (n, 3) float32 and the matrix operations
in the order a force model would make them, ``(matrix, sign, rows, cols, blocks)`` with matrix ``"x"`` (dfdx)
or ``"v"`` (dfdv), sign +1 (``A(i, j) += K``) or -1 (``-=``) and column-major blocks (k, 9). The same model
output feeds the device, the oracle and the reference build, so the three solvers see identical force-model
calls whenever their iterates agree.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32
EYE = np.eye(3, dtype=f32)


class SpringNetwork:
    """Springs with rest lengths (stiffness k, Jacobian -k[(1 - L/l)(I - n n^T) + n n^T]), damping springs
    along the current direction (dfdv -c n n^T) and per-particle damping (dfdv -c I). Elements contribute in
    a seeded order; diagonal blocks go in with +=, off-diagonal ones with -=."""

    def __init__(self, pairs, k, L, damp_pairs=None, damp_k=None, damp_idx=None, damp_c=None, tension_only=False):
        self.tension_only = tension_only   # Jacobian blocks only for stretched springs (the pattern follows x)
        self.pairs = np.asarray(pairs, np.uint32)
        self.k = np.asarray(k, f32)
        self.L = np.asarray(L, f32)
        self.dpairs = np.zeros((0, 2), np.uint32) if damp_pairs is None else np.asarray(damp_pairs, np.uint32)
        self.dk = np.zeros(0, f32) if damp_k is None else np.asarray(damp_k, f32)
        self.didx = np.zeros(0, np.uint32) if damp_idx is None else np.asarray(damp_idx, np.uint32)
        self.dc = np.zeros(0, f32) if damp_c is None else np.asarray(damp_c, f32)
        self.calls = 0

    @staticmethod
    def _dir(x, i, j):
        d = (x[i] - x[j]).astype(f32)
        ln = np.sqrt((d * d).sum(axis=1, dtype=f32)).astype(f32)
        return d / ln[:, None], ln

    def __call__(self, x, v):
        self.calls += 1
        x = np.asarray(x, f32)
        v = np.asarray(v, f32)
        n = len(x)
        f = np.zeros((n, 3), f32)
        ops = []
        i, j = self.pairs[:, 0], self.pairs[:, 1]
        if len(i) and self.tension_only:
            _, ln = self._dir(x, i, j)
            keep = ln > self.L
            i, j = i[keep], j[keep]
            kk, LL = self.k[keep], self.L[keep]
        else:
            kk, LL = self.k, self.L
        if len(i):
            nrm, ln = self._dir(x, i, j)
            fs = ((-kk * (ln - LL))[:, None] * nrm).astype(f32)
            np.add.at(f, i, fs)
            np.add.at(f, j, -fs)
            nn = (nrm[:, :, None] * nrm[:, None, :]).astype(f32)
            a = (f32(1) - LL / ln).astype(f32)
            K = (-kk[:, None, None] * (a[:, None, None] * (EYE - nn) + nn)).astype(f32)
            Kc = np.ascontiguousarray(K.transpose(0, 2, 1)).reshape(-1, 9)   # column-major
            ops.append(("x", +1, np.stack([i, j], 1).reshape(-1), np.stack([i, j], 1).reshape(-1),
                        np.repeat(Kc, 2, axis=0)))
            ops.append(("x", -1, np.stack([i, j], 1).reshape(-1), np.stack([j, i], 1).reshape(-1),
                        np.repeat(Kc, 2, axis=0)))
        if len(self.dpairs):
            di, dj = self.dpairs[:, 0], self.dpairs[:, 1]
            nrm, _ = self._dir(x, di, dj)
            nn = (nrm[:, :, None] * nrm[:, None, :]).astype(f32)
            C = (-self.dk[:, None, None] * nn).astype(f32)
            dv = (v[di] - v[dj]).astype(f32)
            fd = np.einsum("kab,kb->ka", C, dv).astype(f32)
            np.add.at(f, di, fd)
            np.add.at(f, dj, -fd)
            Cc = np.ascontiguousarray(C.transpose(0, 2, 1)).reshape(-1, 9)
            ops.append(("v", +1, np.stack([di, dj], 1).reshape(-1), np.stack([di, dj], 1).reshape(-1),
                        np.repeat(Cc, 2, axis=0)))
            ops.append(("v", -1, np.stack([di, dj], 1).reshape(-1), np.stack([dj, di], 1).reshape(-1),
                        np.repeat(Cc, 2, axis=0)))
        if len(self.didx):
            f[self.didx] -= (self.dc[:, None] * v[self.didx]).astype(f32)
            blocks = (-self.dc[:, None] * EYE.reshape(1, 9)).astype(f32)
            ops.append(("v", +1, self.didx, self.didx, blocks))
        return f, ops


def problem(shape=(6, 5, 4), seed=0, h=1.0 / 60.0, general=True, damping=True, tension_only=False):
    """A jittered grid of particles joined by edge and diagonal springs, elements in a seeded order. general:
    mass blocks with off-diagonal terms, constraint filters of three kinds (I, 0, a plane projection), a
    translation y and a desired sub-solution z; otherwise M = m I, S in {0, I}, y = z = 0 as b3ForceSolver
    passes them (force_solver.cpp:103-139)."""
    rng = np.random.default_rng(seed)
    a, b, c = shape
    grid = np.stack(np.meshgrid(np.arange(a), np.arange(b), np.arange(c), indexing="ij"), -1).reshape(-1, 3)
    n = len(grid)
    rest = (grid * 0.1 + rng.uniform(-0.01, 0.01, (n, 3))).astype(f32)
    idx = np.arange(n).reshape(a, b, c)
    pairs = []
    for off in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)):
        s0 = idx[: a - off[0], : b - off[1], : c - off[2]].reshape(-1)
        s1 = idx[off[0]:, off[1]:, off[2]:].reshape(-1)
        pairs.append(np.stack([s0, s1], 1))
    pairs = np.concatenate(pairs)
    pairs = pairs[rng.permutation(len(pairs))]
    flip = rng.random(len(pairs)) < 0.5
    pairs[flip] = pairs[flip][:, ::-1]
    L = np.linalg.norm(rest[pairs[:, 0]] - rest[pairs[:, 1]], axis=1).astype(f32)
    k = rng.uniform(50.0, 500.0, len(pairs)).astype(f32)
    kw = {}
    if damping:
        dsel = rng.permutation(len(pairs))[: len(pairs) // 3]
        kw = dict(damp_pairs=pairs[dsel], damp_k=rng.uniform(0.1, 2.0, len(dsel)).astype(f32),
                  damp_idx=rng.permutation(n)[: n // 4].astype(np.uint32),
                  damp_c=rng.uniform(0.05, 0.5, n // 4).astype(f32))
    model = SpringNetwork(pairs, k, L, tension_only=tension_only, **kw)
    m = rng.uniform(0.05, 0.2, n).astype(f32)
    x0 = (rest + rng.uniform(-0.02, 0.02, (n, 3))).astype(f32)
    v0 = rng.uniform(-0.5, 0.5, (n, 3)).astype(f32)
    fe = (m[:, None] * np.array([0.0, -9.8, 0.0], f32)).astype(f32)
    M = (m[:, None, None] * EYE).astype(f32)
    S = np.tile(EYE, (n, 1, 1)).astype(f32)
    kind = rng.random(n)
    S[kind < 0.1] = 0.0
    y = np.zeros((n, 3), f32)
    z = np.zeros((n, 3), f32)
    if general:
        P = rng.normal(size=(n, 3, 3)).astype(f32)
        M = (M + (0.02 * m)[:, None, None] * (P + P.transpose(0, 2, 1))).astype(f32)
        plane = (kind >= 0.1) & (kind < 0.25)
        nrm = rng.normal(size=(n, 3)).astype(f32)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True).astype(f32)
        proj = (EYE - nrm[:, :, None] * nrm[:, None, :]).astype(f32)
        S[plane] = proj[plane]
        y = rng.uniform(-1e-3, 1e-3, (n, 3)).astype(f32)
        fixed = kind < 0.25
        z[fixed] = rng.uniform(-0.05, 0.05, (fixed.sum(), 3)).astype(f32)
    col = lambda A: np.ascontiguousarray(A.transpose(0, 2, 1)).reshape(-1, 9)   # noqa: E731 (column-major)
    return dict(model=model, h=f32(h), x0=x0, v0=v0, fe=fe, M=col(M), y=y, S=col(S), z=z if general else None)
