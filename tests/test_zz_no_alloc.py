"""rg_advance_noalloc (the entry point that allocates no device memory)
against rg_advance through the C ABI: two calls in a row on different
contexts, the second on a smaller mesh with fewer steps, then rg_advance
again on a context that used the no-alloc entry point, and the no-alloc
entry point again once rg_advance has built the fused stages' tiles; every
dt, cell and extensive bit-identical.  Sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest

import cases
from parity import assert_parity
from rich_amd import mesh as rmesh

pytestmark = pytest.mark.gpu


def _ctx(m, order=2, **kw):
    from rich_amd.gpu import RichGPU
    g = RichGPU(m, order=order, **kw)
    cells = cases.rough_cells(m, seed=11)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    g.set_cells(cells)
    g.init_extensives()
    return g


def _same(a, b, what):
    assert_parity(a, b, what=what)
    np.testing.assert_array_equal(a, b, err_msg=what)


def test_noalloc_equals_advance_twice_smaller_second():
    big = rmesh.lattice(64, 48, (0.0, 2.0, 0.0, 1.5), jitter=2e-2, seed=9, bc=rmesh.PERIODIC)
    small = rmesh.lattice(20, 16, (0.0, 1.0, 0.0, 0.8), jitter=2e-2, seed=3, bc=rmesh.RIGID)
    for m, integrator, order, steps in ((big, 2, 2, 40), (small, 1, 1, 12)):
        ref, new = _ctx(m, order), _ctx(m, order)
        d_ref = ref.advance(steps, integrator=integrator)
        d_new = new.advance_noalloc(steps, integrator=integrator)
        _same(d_new, d_ref, "dt")
        _same(new.get_cells(), ref.get_cells(), "cells")
        _same(new.get_extensives(), ref.get_extensives(), "extensives")
        # the context's own buffers are back in use afterwards
        _same(new.advance(5, integrator=integrator), ref.advance(5, integrator=integrator), "dt")
        _same(new.get_cells(), ref.get_cells(), "cells after rg_advance")
        # once rg_advance has built the fused stages' tiles, the no-alloc
        # entry point runs them too (odd and even step counts: the state
        # alternates between two buffers)
        for k in (7, 4):
            _same(new.advance_noalloc(k, integrator=integrator),
                  ref.advance(k, integrator=integrator), "dt (tiles built)")
            _same(new.get_cells(), ref.get_cells(), "cells (tiles built)")
            _same(new.get_extensives(), ref.get_extensives(), "extensives (tiles built)")
        ref.close()
        new.close()


def test_noalloc_refuses_what_does_not_fit():
    from rich_amd.gpu import RG_MIXED, RichGPUError
    m = rmesh.lattice(12, 10, (0.0, 1.0, 0.0, 1.0), jitter=1e-2, seed=1, bc=rmesh.RIGID)
    g = _ctx(m)
    with pytest.raises(RichGPUError) as e:
        g.advance_noalloc(65537, record_dt=False)
    assert e.value.status == -1 and "65536 steps" in e.value.text
    g.close()
    g = _ctx(m, precision=RG_MIXED)
    with pytest.raises(RichGPUError) as e:
        g.advance_noalloc(3)
    assert e.value.status == -1 and "FP64" in e.value.text
    g.close()
