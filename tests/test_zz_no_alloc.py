"""iskra_rng_export_noalloc: the draw export staged in the library's fixed
per-GPU scratch array instead of per-call device allocations.  Its draws must
equal those of iskra_rng_export bit for bit, call after call, also when a
later call is smaller than an earlier one (the scratch keeps the previous
call's contents)."""
import numpy as np
import pytest

from iskra_amd import _lib as L


def _ids(rng, n):
    return rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.gpu
def test_rng_export_noalloc_equals_allocating_export(gpu_ready):
    rng = np.random.default_rng(7)
    ctx = L.Context(0)
    try:
        ctx.rng(0x1234_5678_9ABC)
        # first call: close to the scratch limit (ids 120 KB + draws 3.84 MB)
        n1, nb1 = 40000, 6
        assert (n1 * 4 + 255) // 256 * 256 + 16 * n1 * nb1 <= L.NOALLOC_SCRATCH_BYTES
        ids1 = _ids(rng, n1)
        want1 = ctx.rng_export(3, 17, ids1, nb1)
        got1 = ctx.rng_export_noalloc(3, 17, ids1, nb1)
        assert np.array_equal(got1, want1)
        # second call, smaller, other stream and step
        n2, nb2 = 999, 2
        ids2 = _ids(rng, n2)
        want2 = ctx.rng_export(1, 4096, ids2, nb2)
        got2 = ctx.rng_export_noalloc(1, 4096, ids2, nb2)
        assert np.array_equal(got2, want2)
        # the draws differ from the first call's for the shared-prefix slots,
        # so the second call did not return stale scratch contents
        assert not np.array_equal(got2[:, :nb2], got1[:n2, :nb2])
        # a request beyond the scratch is refused before anything runs ...
        with pytest.raises(L.IskraError) as ei:
            ctx.rng_export_noalloc(0, 0, _ids(rng, 70000), 4)
        assert ei.value.code == L.E_ARG
        # ... and the next call still works
        got3 = ctx.rng_export_noalloc(1, 4096, ids2, nb2)
        assert np.array_equal(got3, want2)
        # empty request
        assert ctx.rng_export_noalloc(0, 0, np.zeros(0, np.uint32), 3).size == 0
    finally:
        ctx.close()


@pytest.mark.gpu
def test_rng_export_noalloc_two_contexts_share_the_device_scratch(gpu_ready):
    rng = np.random.default_rng(11)
    a, b = L.Context(0), L.Context(0)
    try:
        a.rng(1)
        b.rng(2)
        ids = _ids(rng, 5000)
        ga = a.rng_export_noalloc(0, 5, ids, 3)
        gb = b.rng_export_noalloc(0, 5, ids, 3)
        assert np.array_equal(ga, a.rng_export(0, 5, ids, 3))
        assert np.array_equal(gb, b.rng_export(0, 5, ids, 3))
        assert not np.array_equal(ga, gb)
    finally:
        a.close()
        b.close()
