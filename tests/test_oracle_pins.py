"""Pins the oracle's ghost generators, shock switch and both slope limiters
against restatements written here in plain Python from the reference's text
(RigidWallGenerator.cpp, PeriodicGhostGenerator.cpp, GhostPointGenerator.cpp,
LinearGaussImproved.cpp), on inputs built so that each plausible misreading
of those rules changes the result.

Every restatement takes its rules from a `Rules` record.  Besides checking
the oracle against the reference's rules, each test evaluates the same
restatement with one rule misread and asserts that the result changes on
these inputs: that is the proof that the test would catch the oracle making
that mistake.  (tests/test_oracle_mutations.py plants the same mistakes in
the oracle's C source and checks that this file fails.)

Hand-evaluated outcomes are asserted exactly where the rule fixes them:
a slope whose every face change lies inside the 1e-9 guard keeps its naive
value bit for bit, a rigid ghost face state equals the ghost cell.

The naive (unlimited) slopes the limiters start from are the oracle's own
`naive` output, which test_oracle_scheme.py pins separately (exact for
linear fields).
"""
import math
from dataclasses import dataclass, replace

import numpy as np
import pytest

import cases
import pyoracle as po
from rich_amd import mesh as rmesh

G = 5.0 / 3.0


# ------------------------------------------------------------ helpers --
def smax(a, b):
    """std::max(a, b)"""
    return b if a < b else a


def smin(a, b):
    """std::min(a, b)"""
    return b if b < a else a


def orig(m, i):
    """Tessellation::GetOriginalIndex"""
    return i if i < m.n_cells else int(m.ghost_orig[i - m.n_cells])


@dataclass(frozen=True)
class Rules:
    """The reference's rules; each misreading below changes one field."""
    rigid_reflect: bool = True        # RigidWallGenerator.cpp:13
    rigid_source: str = "first_face"  # flat_map insert keeps the first face (:32)
    gate: float = 0.05                # slope_limit, LinearGaussImproved.cpp:232
    vel_abs_guard: bool = True        # slope_limit velocity guards abs(), :253-265
    skip_mirror: bool = True          # slope_limit skips mirror faces, :194-195
    shocked_gate: float = 0.1         # shocked_slope_limit, :344
    shocked_abs_guard: bool = False   # shocked guards have no abs() on the right, :346-364
    theta: float = None               # diffusecoeff (None: the run's parameter)
    div_test: bool = True             # is_shock cond1, :153-154
    pressure_test: bool = True        # is_shock cond2, :155


REF = Rules()


# ------------------------------------------------------- ghost cells --
def ref_ghosts(m, cells, rules=REF):
    """Ghost values [4][N+G] (only ghosts that touch a face are filled).

    GetOuterEdgesIndeces (GhostPointGenerator.cpp:18-39) lists the faces in
    ascending order with the side that is a ghost (first if first >= N,
    else second if second >= N).  RigidWallGenerator (RigidWallGenerator.cpp:
    17-37) inserts for each listed face, into a flat_map whose insert keeps
    an existing key, the cell on the face's other side with its velocity
    reversed along n = (r_ghost - r_real)/|r_ghost - r_real|:
    v -= 2 (v.n) n (:5-14).  PeriodicGhostGenerator copies the original."""
    N, G = m.n_cells, m.n_ghosts
    out = np.zeros((4, N + G))
    out[:, :N] = cells
    seen = set()
    for f in range(m.n_faces):
        a, b = int(m.face_first[f]), int(m.face_second[f])
        if a >= N:
            side, g, real = 1, a, b
        elif b >= N:
            side, g, real = 2, b, a
        else:
            continue
        if m.bc == rmesh.PERIODIC:
            out[:, g] = cells[:, orig(m, g)]
            continue
        if rules.rigid_source == "first_face":
            if g in seen:
                continue
        elif rules.rigid_source == "orig":
            real = orig(m, g)
        # "last_face": every listed face overwrites
        seen.add(g)
        d, p, vx, vy = (float(cells[k, real]) for k in range(4))
        if rules.rigid_reflect:
            if side == 1:
                nx = float(m.pt_x[a]) - float(m.pt_x[b])
                ny = float(m.pt_y[a]) - float(m.pt_y[b])
            else:
                nx = float(m.pt_x[b]) - float(m.pt_x[a])
                ny = float(m.pt_y[b]) - float(m.pt_y[a])
            an = math.sqrt(nx * nx + ny * ny)
            nx, ny = nx / an, ny / an
            s2 = 2 * (vx * nx + vy * ny)
            vx, vy = vx - s2 * nx, vy - s2 * ny
        out[:, g] = (d, p, vx, vy)
    return out


def test_rigid_ghosts_reflect_about_the_ghost_real_axis():
    """RigidWallGenerator on meshes whose wall cells move towards and along
    the walls: ghosts equal the restatement exactly; dropping the reflection
    or taking the ghost's original instead of its first face's real side
    changes them."""
    checked = 0
    for name, m in cases.small_meshes() + [("merged", cases.merged_ghost_mesh()[0])]:
        if m.bc != rmesh.RIGID:
            continue
        cells = cases.rough_cells(m, seed=4)
        cells[0] = np.abs(cells[0]) + 0.2
        cells[1] = np.abs(cells[1]) + 0.2
        cells[2] += 0.7  # a wall-normal velocity at every wall
        cells[3] -= 0.4
        got = po.Oracle(m).ghost_cells(cells)
        want = ref_ghosts(m, cells)
        touched = np.unique(np.concatenate([m.face_first[m.face_first >= m.n_cells],
                                            m.face_second[m.face_second >= m.n_cells]]))
        np.testing.assert_array_equal(got[:, touched], want[:, touched], err_msg=name)
        for mis in (replace(REF, rigid_reflect=False), replace(REF, rigid_source="orig"),
                    replace(REF, rigid_source="last_face")):
            if name != "merged" and mis.rigid_source != "first_face":
                continue  # one real neighbour per ghost: the readings coincide
            other = ref_ghosts(m, cells, mis)
            assert not np.array_equal(other[:, touched], want[:, touched]), (name, mis)
        checked += len(touched)
    assert checked > 100


def test_merged_mesh_ghost_takes_its_first_face():
    """Hand-evaluated: a ghost that borders two real cells copies the real
    side of its lower-numbered face, reflected about that side's axis."""
    m, merged = cases.merged_ghost_mesh()
    N = m.n_cells
    cells = cases.rough_cells(m, seed=8)
    cells[0] = np.abs(cells[0]) + 0.2
    cells[1] = np.abs(cells[1]) + 0.2
    got = po.Oracle(m).ghost_cells(cells)
    for g, lo, hi in merged:
        real = int(m.face_second[lo]) if int(m.face_first[lo]) == g else int(m.face_first[lo])
        assert real < N and real != orig(m, g)
        n = np.array([m.pt_x[g] - m.pt_x[real], m.pt_y[g] - m.pt_y[real]])
        n = n / np.hypot(*n)
        v = cells[2:, real]
        np.testing.assert_allclose(got[2:, g], v - 2 * (v @ n) * n, rtol=1e-14, atol=1e-15)
        np.testing.assert_array_equal(got[:2, g], cells[:2, real])


# ------------------------------------------------------------ limiters --
def cell_faces(m, i):
    fs = m.csr_face[m.csr_off[i]:m.csr_off[i + 1]]
    nb = [int(m.face_second[f]) if int(m.face_first[f]) == i else int(m.face_first[f])
          for f in fs]
    return [int(f) for f in fs], nb


def ref_limited_slope(m, prm, cng, naive, i, rules=REF):
    """calc_slope's limiting half (LinearGaussImproved.cpp:475-493) for cell
    i, from its naive slope: is_shock (:150-157), then slope_limit
    (:174-300) or shocked_slope_limit (:302-399).  Returns (slope[8],
    shocked)."""
    faces, nb = cell_faces(m, i)
    cell = [float(cng[k, i]) for k in range(4)]
    ncell = [[float(cng[k, j]) for k in range(4)] for j in nb]
    sx = [float(naive[k, i]) for k in range(4)]
    sy = [float(naive[4 + k, i]) for k in range(4)]
    cmx, cmy = float(m.cm_x[i]), float(m.cm_y[i])
    theta = prm["theta"] if rules.theta is None else rules.theta
    # is_shock with PressureRatio (:136-148)
    width = math.sqrt(float(m.tess_volume[i]) / math.pi)
    cs = math.sqrt(prm["gamma"] * cell[1] / cell[0])
    cond1 = (sx[2] + sy[3]) * width < (-prm["delta_v"]) * cs
    pr = 1.0
    for nc in ncell:
        pr = smin(pr, nc[1] / cell[1]) if cell[1] > nc[1] else smin(pr, cell[1] / nc[1])
    cond2 = pr < prm["delta_p"]
    shocked = (rules.div_test and cond1) or (rules.pressure_test and cond2)
    mirror = [orig(m, int(m.face_first[f])) == orig(m, int(m.face_second[f])) for f in faces]
    cmax, cmin = list(cell), list(cell)
    for nc, mir in zip(ncell, mirror):
        if not shocked and rules.skip_mirror and mir:
            continue
        for k in range(4):
            cmax[k] = smax(cmax[k], nc[k])
            cmin[k] = smin(cmin[k], nc[k])
    maxd = [cmax[k] - cell[k] for k in range(4)]
    mind = [cmin[k] - cell[k] for k in range(4)]
    psi = [1.0] * 4
    for f, nc in zip(faces, ncell):
        fx = (float(m.v1x[f]) + float(m.v2x[f])) * 0.5
        fy = (float(m.v1y[f]) + float(m.v2y[f])) * 0.5
        for k in range(4):
            cval = cell[k] + sx[k] * (fx - cmx)
            cval = cval + sy[k] * (fy - cmy)
            dphi = cval - cell[k]
            big = smax(abs(maxd[k]), abs(mind[k]))
            if not shocked:
                if abs(dphi) > rules.gate * big or cval * cell[k] < 0:
                    guard = abs(1e-9 * cell[k]) if (k >= 2 and rules.vel_abs_guard) else 1e-9 * cell[k]
                    if dphi > guard:
                        psi[k] = smin(psi[k], maxd[k] / dphi)
                    elif dphi < -guard:
                        psi[k] = smin(psi[k], mind[k] / dphi)
            else:
                if abs(dphi) > rules.shocked_gate * big or cval * cell[k] < 0:
                    guard = abs(1e-9 * cell[k]) if rules.shocked_abs_guard else 1e-9 * cell[k]
                    if abs(dphi) > guard:
                        psi[k] = smin(psi[k], smax(theta * (nc[k] - cell[k]) / dphi, 0.0))
    out = [sx[k] * psi[k] for k in range(4)] + [sy[k] * psi[k] for k in range(4)]
    return np.array(out), shocked


PRM = dict(gamma=G, delta_v=0.2, theta=0.5, delta_p=0.7)


def run_case(m, cells, rules_list, prm=PRM):
    """Oracle slopes against the restatement with the reference's rules;
    returns {misreading: number of cells whose slope it changes}."""
    o = po.Oracle(m, order=2, slf=True, delta_v=prm["delta_v"], theta=prm["theta"],
                  delta_p=prm["delta_p"])
    s, naive, _, _ = o.lgi(cells)
    cng = ref_ghosts(m, cells)
    N = m.n_cells
    want = np.zeros((8, N))
    shocked = np.zeros(N, bool)
    for i in range(N):
        want[:, i], shocked[i] = ref_limited_slope(m, prm, cng, naive, i)
    np.testing.assert_allclose(s, want, rtol=1e-13, atol=1e-300)
    changed = {}
    for name, rules in rules_list:
        n = 0
        for i in range(N):
            alt, _ = ref_limited_slope(m, prm, cng, naive, i, rules)
            n += not np.array_equal(alt, want[:, i])
        changed[name] = n
    return s, naive, shocked, changed


MISREAD_UNSHOCKED = [("gate 0.1 in slope_limit", replace(REF, gate=0.1)),
                     ("no mirror skip", replace(REF, skip_mirror=False)),
                     ("velocity guard without abs", replace(REF, vel_abs_guard=False))]
MISREAD_SHOCKED = [("shocked gate 0.05", replace(REF, shocked_gate=0.05)),
                   ("shocked guard with abs", replace(REF, shocked_abs_guard=True)),
                   ("theta 1", replace(REF, theta=1.0))]
MISREAD_SWITCH = [("no divergence test", replace(REF, div_test=False)),
                  ("no pressure test", replace(REF, pressure_test=False))]


def test_slope_limit_velocity_guard_keeps_tiny_changes():
    """slope_limit (LinearGaussImproved.cpp:253-265): a face change of vx
    or vy no larger than |1e-9 v| limits nothing, whatever v's sign, so
    both velocity slopes stay equal to the naive slopes bit for bit.
    Reading the guard without abs() makes every such vx face (v < 0)
    limit."""
    m = rmesh.lattice(12, 10, (0.0, 1.2, 0.0, 1.0), jitter=3e-2, seed=13, bc=rmesh.PERIODIC)
    cells = cases.tiny_velocity_cells(m, shocked=False)
    s, naive, shocked, changed = run_case(m, cells, MISREAD_UNSHOCKED)
    assert not shocked.any()
    for k in (2, 3, 6, 7):
        np.testing.assert_array_equal(s[k], naive[k])
        assert np.count_nonzero(naive[k]) > m.n_cells // 2
    assert changed["velocity guard without abs"] > m.n_cells // 4


def test_shocked_guard_has_no_abs():
    """shocked_slope_limit (LinearGaussImproved.cpp:358,364) guards with
    |dphi| > 1e-9 v, no abs() on the right: for vx < 0 every gated face
    limits however small its change, while vy > 0 keeps its naive slope bit
    for bit.  Adding the abs() leaves vx unlimited."""
    m = rmesh.lattice(12, 10, (0.0, 1.2, 0.0, 1.0), jitter=3e-2, seed=13, bc=rmesh.PERIODIC)
    cells = cases.tiny_velocity_cells(m, shocked=True)
    s, naive, shocked, changed = run_case(m, cells, MISREAD_SHOCKED)
    assert shocked.all()
    np.testing.assert_array_equal(s[3], naive[3])
    np.testing.assert_array_equal(s[7], naive[7])
    limited = (s[2] != naive[2]).sum()
    assert limited > m.n_cells // 4
    assert changed["shocked guard with abs"] >= limited


def test_limiters_on_rough_states():
    """Smooth waves, a jump and noise on rigid and periodic meshes:
    unlimited, limited and shocked cells, wall cells with mirror faces.
    The oracle equals the restatement on every cell, and each misreading of
    a gate, the mirror skip, theta or the shock switch changes some
    cell."""
    total = {}
    for name, m in cases.small_meshes():
        for seed in (3, 5):
            cells = cases.rough_cells(m, seed=seed)
            cells[0] = np.abs(cells[0]) + 0.2
            cells[1] = np.abs(cells[1]) + 0.2
            _, _, shocked, changed = run_case(
                m, cells, MISREAD_UNSHOCKED[:2] + MISREAD_SHOCKED[::2] + MISREAD_SWITCH)
            assert shocked.any() and not shocked.all(), name
            for k, v in changed.items():
                total[k] = total.get(k, 0) + v
    for k, v in total.items():
        assert v > 0, k


def test_divergence_switch_alone():
    """A converging flow at uniform pressure: is_shock's divergence test
    (cond1, LinearGaussImproved.cpp:153-154) alone switches the shocked
    limiter on; without it the cells would take slope_limit."""
    m = rmesh.lattice(16, 16, (-1.0, 1.0, -1.0, 1.0), jitter=2e-2, seed=3, bc=rmesh.RIGID)
    N = m.n_cells
    x, y = m.cm_x[:N], m.cm_y[:N]
    rng = np.random.default_rng(2)
    cells = np.vstack([1.0 + 0.2 * np.sin(3 * x + 1) + 0.05 * rng.random(N),
                       1.0 + 0.01 * rng.random(N), -3.0 * x + 0.05 * rng.random(N),
                       -3.0 * y + 0.05 * rng.random(N)])
    _, _, shocked, changed = run_case(m, cells, MISREAD_SWITCH)
    assert shocked.sum() > N // 2
    assert changed["no divergence test"] > 0


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
def test_ghost_face_states(bc):
    """LinearGaussImproved.cpp:597-634: a boundary face's ghost side is the
    ghost cell plus the ghost gradient at the ghost's CM -- zero for
    RigidWallGenerator (RigidWallGenerator.cpp:39-54: the face state equals
    the ghost cell), the original cell's limited slope for
    PeriodicGhostGenerator (PeriodicGhostGenerator.cpp:19-24)."""
    m = dict(cases.small_meshes())["rigid_flip" if bc == rmesh.RIGID else "periodic_flip"]
    cells = cases.rough_cells(m, seed=7)
    cells[0] = np.abs(cells[0]) + 0.2
    cells[1] = np.abs(cells[1]) + 0.2
    o = po.Oracle(m, order=2)
    s, _, L, R = o.lgi(cells)
    cng = ref_ghosts(m, cells)
    N = m.n_cells
    n = 0
    for f in range(m.n_faces):
        a, b = int(m.face_first[f]), int(m.face_second[f])
        if a < N and b < N:
            continue
        g, got = (a, L[:, f]) if a >= N else (b, R[:, f])
        if bc == rmesh.RIGID:
            np.testing.assert_array_equal(got, cng[:, g])
        else:
            j = orig(m, g)
            fx = (m.v1x[f] + m.v2x[f]) * 0.5
            fy = (m.v1y[f] + m.v2y[f]) * 0.5
            want = cng[:, g] + s[:4, j] * (fx - m.cm_x[g])
            want = want + s[4:, j] * (fy - m.cm_y[g])
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
            # a zero ghost gradient would give the ghost cell itself
            n += not np.allclose(got, cng[:, g], rtol=1e-12, atol=0)
        n += 1
    assert n > 20
