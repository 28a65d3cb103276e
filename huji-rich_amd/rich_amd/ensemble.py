"""Batched ensembles: many independent meshes as one disjoint union.

`union(meshes)` concatenates meshes into one Mesh: the real cells of every
instance first (instance by instance), then every instance's ghosts, with
face, CSR and ghost ids renumbered.  `cell_off` gives each instance's cell
range.  The GPU context then advances all instances together, each with its
own time step (RichGPU.set_instances).

`sedov_ensemble` draws the ensemble of Sedov blasts: cell counts uniform in
[n_min^2, n_max^2] on square jittered lattices in [-1, 1]^2 with rigid walls,
blast energy log-uniform in [e_min, e_max] deposited as uniform thermal
energy within `dep` cell spacings of the origin, ambient density 1 and
pressure p_amb."""
from dataclasses import dataclass

import numpy as np

from . import mesh as rmesh


@dataclass
class Union:
    mesh: rmesh.Mesh
    cell_off: np.ndarray   # (K+1,)
    face_off: np.ndarray   # (K+1,)
    ghost_off: np.ndarray  # (K+1,)

    def split(self, cells):
        """(4, N) union array -> list of per-instance (4, N_k) arrays."""
        return [cells[:, a:b] for a, b in zip(self.cell_off[:-1], self.cell_off[1:])]


def union(meshes):
    K = len(meshes)
    bc = meshes[0].bc
    if any(m.bc != bc for m in meshes):
        raise ValueError("all instances need the same boundary kind")
    cn = np.array([m.n_cells for m in meshes])
    fn = np.array([m.n_faces for m in meshes])
    gn = np.array([m.n_ghosts for m in meshes])
    coff = np.concatenate([[0], np.cumsum(cn)]).astype(np.int32)
    foff = np.concatenate([[0], np.cumsum(fn)]).astype(np.int64)
    goff = np.concatenate([[0], np.cumsum(gn)]).astype(np.int64)
    N, F, G = int(coff[-1]), int(foff[-1]), int(goff[-1])
    ff = np.empty(F, np.int32)
    fs = np.empty(F, np.int32)
    pts = {k: np.empty(N + G) for k in ("pt_x", "pt_y", "cm_x", "cm_y")}
    csr_off = np.empty(N + 1, np.int32)
    csr_face = []
    ghost_orig = np.empty(max(G, 0), np.int32)
    for k, m in enumerate(meshes):
        n = m.n_cells

        def remap(ids):
            ids = np.asarray(ids, np.int64)
            return np.where(ids < n, ids + coff[k], ids - n + N + goff[k]).astype(np.int32)

        ff[foff[k]:foff[k + 1]] = remap(m.face_first)
        fs[foff[k]:foff[k + 1]] = remap(m.face_second)
        for key in pts:
            a = getattr(m, key)
            pts[key][coff[k]:coff[k + 1]] = a[:n]
            pts[key][N + goff[k]:N + goff[k + 1]] = a[n:]
        ghost_orig[goff[k]:goff[k + 1]] = m.ghost_orig + coff[k]
        csr_face.append(m.csr_face.astype(np.int64) + foff[k])
    # CSR offsets: concatenate with running nnz
    nnz = np.array([m.csr_off[-1] for m in meshes], np.int64)
    noff = np.concatenate([[0], np.cumsum(nnz)])
    for k, m in enumerate(meshes):
        csr_off[coff[k]:coff[k + 1]] = m.csr_off[:-1] + noff[k]
    csr_off[N] = noff[-1]
    cat = lambda key: np.concatenate([getattr(m, key) for m in meshes])  # noqa: E731
    um = rmesh.Mesh(
        n_cells=N, n_faces=F, n_ghosts=G, bc=bc, face_first=ff, face_second=fs,
        v1x=cat("v1x"), v1y=cat("v1y"), v2x=cat("v2x"), v2y=cat("v2y"),
        pt_x=pts["pt_x"], pt_y=pts["pt_y"], cm_x=pts["cm_x"], cm_y=pts["cm_y"],
        volume=cat("volume"), tess_volume=cat("tess_volume"), csr_off=csr_off,
        csr_face=np.concatenate(csr_face).astype(np.int32), ghost_orig=ghost_orig)
    return Union(um, coff, foff, goff)


def sedov_cells(m, energy, dep_radius, gamma=5.0 / 3.0, p_amb=1e-5):
    N = m.n_cells
    r = np.hypot(m.pt_x[:N], m.pt_y[:N])
    hot = r < dep_radius
    if not hot.any():
        hot = r <= r.min()
    c = np.zeros((4, N))
    c[0] = 1.0
    c[1] = p_amb
    c[1, hot] = (gamma - 1.0) * energy / m.volume[hot].sum()
    return c


def sedov_ensemble(K=256, n_min=64, n_max=256, e_min=0.1, e_max=10.0, dep=3.0, seed=2024,
                   gamma=5.0 / 3.0, p_amb=1e-5):
    """(meshes, cells per instance, parameters) of K seeded Sedov blasts."""
    rng = np.random.default_rng(seed)
    counts = rng.uniform(n_min * n_min, n_max * n_max, K)
    ns = np.clip(np.rint(np.sqrt(counts)).astype(int), n_min, n_max)
    energies = np.exp(rng.uniform(np.log(e_min), np.log(e_max), K))
    meshes, cells = [], []
    for k in range(K):
        n = int(ns[k])
        m = rmesh.lattice(n, n, (-1.0, 1.0, -1.0, 1.0), jitter=1e-3, seed=int(seed + 7 * k + 1),
                          bc=rmesh.RIGID)
        meshes.append(m)
        cells.append(sedov_cells(m, energies[k], dep * 2.0 / n, gamma, p_amb))
    return meshes, cells, {"n": ns, "energy": energies}
