"""cpm_total_energy_noalloc (the allocation-free entry point) against cpm_total_energy.

Both run the same reduction kernels; the no-alloc variant keeps its partial sums in the
library's static device array. Called through the C-ABI on two lattices in a row, the second
smaller than the first, the results must be bit-identical to cpm_total_energy's, and calling
again after the other context used the array must still give the same values.
"""
import ctypes as C

import pytest

from corepotts import init_scientific, load_library, workloads


def _noalloc(lib, ctx):
    out = C.c_double()
    rc = lib.cpm_total_energy_noalloc(ctx, C.byref(out))
    assert rc == 0, lib.cpm_last_error(ctx)
    return out.value


def _alloc(lib, ctx):
    out = C.c_double()
    rc = lib.cpm_total_energy(ctx, C.byref(out))
    assert rc == 0, lib.cpm_last_error(ctx)
    return out.value


@pytest.mark.gpu
def test_noalloc_energy_equals_allocating_entry_point():
    lib = load_library()
    big = init_scientific(*workloads.scaled_2d(1024))          # 1024^2, 10,404 cells
    small = init_scientific(*workloads.cubes_3d(32))           # 32^3, Float32 J
    big.step(3)
    small.step(2)
    e_big = _noalloc(lib, big._ctx)
    assert e_big == _alloc(lib, big._ctx)
    e_small = _noalloc(lib, small._ctx)                        # second call, smaller input
    assert e_small == _alloc(lib, small._ctx)
    assert e_small != e_big
    # the array now holds the small lattice's partial sums; the big one overwrites them all
    assert _noalloc(lib, big._ctx) == e_big
    big.step(1)
    assert _noalloc(lib, big._ctx) == _alloc(lib, big._ctx)
    big.close()
    small.close()
