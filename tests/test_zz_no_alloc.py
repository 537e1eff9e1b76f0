"""The allocation-free fingerprint stage (emp_stage_fingerprints_noalloc) against the allocating
one (emp_stage_fingerprints), through the C ABI, on calls in a row that share the library's
device array: a large call, a smaller one with a byte outside ACGTacgt (status 1, where the
allocating entry point returns EMP_ERR_BASE), then a smaller clean one (status back to 0)."""
import numpy as np
import pytest

from empanada_amd import gen, hipmem, native

pytestmark = pytest.mark.gpu


def _noalloc(reads):
    n, L = reads.shape
    d_r = hipmem.DeviceBuffer.from_array(reads)
    d_k = hipmem.DeviceBuffer(16 * n)
    d_s = hipmem.DeviceBuffer(4)
    rc = native.lib().emp_stage_fingerprints_noalloc(d_r.ptr, n, L, d_k.ptr, d_s.ptr, None)
    assert rc == 0, rc
    hipmem.synchronize()
    keys = d_k.to_array(np.uint64, (2 * n,))
    status = int(d_s.to_array(np.uint32, (1,))[0])
    for b in (d_r, d_k, d_s):
        b.free()
    return keys, status


def test_noalloc_matches_the_allocating_entry_point_on_calls_in_a_row():
    g = gen.genome(10**6, 3)
    big = gen.reads(g, 1 << 20, 150, 0.5, 0.005, 4)  # the largest input the array takes
    bad = gen.reads(g, 30_000, 100, 0.5, 0.0, 5)
    bad[12_345, 17] = ord("N")
    small = gen.reads(g, 7_777, 101, 0.5, 0.005, 6)
    keys, status = _noalloc(big)
    assert status == 0
    assert np.array_equal(keys, native.stage_fingerprints(big))
    keys, status = _noalloc(bad)
    assert status == 1
    with pytest.raises(native.EmpError) as e:
        native.stage_fingerprints(bad)
    assert e.value.code == -5
    keys, status = _noalloc(small)
    assert status == 0
    assert np.array_equal(keys, native.stage_fingerprints(small))


def test_noalloc_rejects_inputs_over_its_limit():
    lib = native.lib()
    assert lib.emp_stage_fingerprints_noalloc(None, (1 << 20) + 1, 100, None, None, None) == -8
    assert lib.emp_stage_fingerprints_noalloc(None, 10, 0, None, None, None) == -4
    assert lib.emp_stage_fingerprints_noalloc(None, 10, 100, None, None, None) == -1
