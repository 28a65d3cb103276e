"""rg_set_mesh rejects a mesh whose topology the kernels cannot trust, with
RG_E_ARG and before anything is uploaded: side ids out of range or equal,
faces between two ghosts, a CSR that is not ascending or lists a face the
cell is not on, a face between two real cells missing from one of their
lists, a bad CSR offset or ghost original.  The intact mesh is accepted."""
import types

import numpy as np
import pytest

import cases
from rich_amd.gpu import RichGPU, RichGPUError

pytestmark = pytest.mark.gpu

RG_E_ARG = -1
FIELDS = ("face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x", "cm_y",
          "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig")


def _copy(m):
    c = types.SimpleNamespace(n_cells=m.n_cells, n_faces=m.n_faces, n_ghosts=m.n_ghosts, bc=m.bc)
    for k in FIELDS:
        setattr(c, k, np.array(getattr(m, k), copy=True))
    return c


def _real_real_entry(m):
    """(cell, CSR position) of a face between two real cells."""
    N = m.n_cells
    for i in range(N):
        for k in range(m.csr_off[i], m.csr_off[i + 1]):
            f = m.csr_face[k]
            if m.face_first[f] < N and m.face_second[f] < N:
                return i, k
    raise AssertionError("no interior face")


def _drop_entry(m, i, k):
    m.csr_face = np.delete(m.csr_face, k)
    m.csr_off = m.csr_off.copy()
    m.csr_off[i + 1:] -= 1


def _corruptions(m):
    N, G = m.n_cells, m.n_ghosts
    out = []

    c = _copy(m)
    c.face_first[3] = N + G
    out.append(("side out of range", c))

    c = _copy(m)
    c.face_second[5] = c.face_first[5]
    out.append(("both sides equal", c))

    if G >= 2:
        c = _copy(m)
        c.face_first[0], c.face_second[0] = N, N + 1
        out.append(("two ghost sides", c))

    c = _copy(m)
    i = next(i for i in range(N) if m.csr_off[i + 1] - m.csr_off[i] >= 2)
    k = m.csr_off[i]
    c.csr_face[k], c.csr_face[k + 1] = c.csr_face[k + 1], c.csr_face[k]
    out.append(("CSR not ascending", c))

    c = _copy(m)
    i, k = _real_real_entry(m)
    others = np.nonzero((m.face_first != i) & (m.face_second != i))[0]
    c.csr_face[k] = others[others > (c.csr_face[k - 1] if k > m.csr_off[i] else -1)][0]
    out.append(("face not on the cell", c))

    c = _copy(m)
    i, k = _real_real_entry(m)
    _drop_entry(c, i, k)
    out.append(("face missing from one list", c))

    c = _copy(m)
    c.csr_off[0] = 1
    out.append(("first offset not zero", c))

    c = _copy(m)
    c.csr_off[N // 2] = c.csr_off[N // 2 + 1] + 1
    out.append(("offsets decreasing", c))

    if G:
        c = _copy(m)
        c.ghost_orig[0] = N
        out.append(("ghost original out of range", c))
    return out


@pytest.mark.parametrize("name", [n for n, _ in cases.small_meshes()])
def test_set_mesh_rejects_bad_topology(name):
    m = dict(cases.small_meshes())[name]
    g = RichGPU(_copy(m), order=2)  # the intact copy is accepted
    g.close()
    for what, bad in _corruptions(m):
        with pytest.raises(RichGPUError) as e:
            RichGPU(bad, order=2)
        assert e.value.status == RG_E_ARG, what


def test_rejected_swap_leaves_the_context_intact():
    """rg_replace_mesh with a corrupted mesh fails before touching the
    device: the context steps on exactly as one that never saw it."""
    m = dict(cases.small_meshes())["rigid"]
    cells = cases.rough_cells(m, seed=5)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    runs = []
    for swap in (False, True):
        g = RichGPU(_copy(m), order=2)
        g.set_cells(cells)
        g.init_extensives()
        g.advance(3, integrator=2)
        if swap:
            for what, bad in _corruptions(m):
                with pytest.raises(RichGPUError) as e:
                    g.replace_mesh(bad)
                assert e.value.status == RG_E_ARG, what
        g.advance(3, integrator=2)
        runs.append((g.get_cells(), g.get_extensives()))
        g.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
