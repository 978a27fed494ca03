"""felbm_export_tables_noalloc, the allocation-free form of felbm_export_tables
(tables staged in the library's static 32 MiB device workspace), through the
C ABI: its tables equal those of felbm_export_tables on two calls in a row,
the second on a smaller instance, in both precisions; an instance beyond the
workspace is refused with FELBM_EINVAL and later calls still work."""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import Batch, MultiPhaseGPU
from felbm_gpu_amd import workloads as W
from felbm_gpu_amd._lib import FelbmError

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_export_tables_noalloc_equals_export_tables(precision):
    big = W.spheres(40, 40, 36, 3.0, 5.0, 0.55, 8)
    small, _ = W.cylinders(48, 40, 1, 6, 4.0, 7.0, seed=5)
    b = Batch(precision=precision)
    prm = W.derive_params()
    gb = MultiPhaseGPU(batch=b).init(big, prm, {})
    gs = MultiPhaseGPU(batch=b).init(small, prm, {}, periodic=(1, 0, 1))
    assert gs.n < gb.n <= 46091
    _say("noalloc export %s: instances of %d and %d sites" % (precision, gb.n, gs.n))
    first = gb.export_tables(noalloc=True)           # two calls in a row,
    second = gs.export_tables(noalloc=True)          # the second smaller
    for g, got in ((gb, first), (gs, second)):
        ref = g.export_tables()
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (g.n, k)
    # again after the allocating form ran, in the other order
    assert all(np.array_equal(gs.export_tables(noalloc=True)[k], second[k]) for k in second)
    assert all(np.array_equal(gb.export_tables(noalloc=True)[k], first[k]) for k in first)
    # the step state is untouched by the exports
    gb.init_rest(W.slab_c(big, 0, 0, 20, 4.0), 1.0)
    h0 = gb.download_state()[0]
    gb.export_tables(noalloc=True)
    assert np.array_equal(gb.download_state()[0], h0)
    b.free()


def test_export_tables_noalloc_refuses_beyond_workspace():
    b = Batch(precision="f64")
    prm = W.derive_params()
    over = MultiPhaseGPU(batch=b).init(np.zeros((40, 40, 40), np.uint8), prm, {})   # 64,000 sites
    ok = MultiPhaseGPU(batch=b).init(W.spheres(20, 20, 20, 3.0, 5.0, 0.55, 3), prm, {})
    with pytest.raises(FelbmError) as ei:
        over.export_tables(noalloc=True)
    assert ei.value.code == -1 and "46091" in str(ei.value)
    ref = ok.export_tables()
    got = ok.export_tables(noalloc=True)
    assert all(np.array_equal(got[k], ref[k]) for k in ref)
    b.free()
