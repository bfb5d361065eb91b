"""ha_free_bytes_noalloc_c, the statistics entry point that allocates nothing,
against ha_free_bytes_c, through the C ABI: the same free-byte count on two
calls in a row, the second on a smaller heap (fewer slabs to reduce), each
also equal to the slab bytes minus the blocks the test allocated."""
import ctypes

import numpy as np
import pytest

import halloc_amd as H

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def free_bytes_both():
    L = H.lib()
    a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
    assert L.ha_free_bytes_c(ctypes.byref(a)) == 0, L.ha_last_error()
    assert L.ha_free_bytes_noalloc_c(ctypes.byref(b)) == 0, L.ha_last_error()
    return int(a.value), int(b.value)


@pytest.mark.parametrize("memory,sb_sz_sh,nblocks,size", [(1 << 30, 22, 300000, 64), (64 << 20, 21, 20000, 48)])
def test_no_alloc_free_bytes_matches_the_allocating_call(memory, sb_sz_sh, nblocks, size):
    say(f"no-alloc free bytes: {memory >> 20} MiB heap, {nblocks} blocks of {size} B")
    H.Halloc.init(H.HallocOpts(memory=memory, sb_sz_sh=sb_sz_sh))
    try:
        arena = H.Halloc.geometry()["arena_bytes"]
        assert free_bytes_both() == (arena, arena)
        ptrs = H.Halloc.malloc_sizes(np.full(nblocks, size, dtype=np.uint32))
        assert (ptrs != 0).all()
        a, b = free_bytes_both()
        assert a == b == arena - nblocks * size
        H.Halloc.free_ptrs(ptrs[::2])
        a, b = free_bytes_both()
        assert a == b == arena - (nblocks - len(ptrs[::2])) * size
        H.Halloc.free_ptrs(ptrs[1::2])
        assert free_bytes_both() == (arena, arena)
    finally:
        H.Halloc.shutdown()


def test_no_alloc_free_bytes_needs_an_initialized_heap():
    v = ctypes.c_uint64(0)
    assert H.lib().ha_free_bytes_noalloc_c(ctypes.byref(v)) == -2   # HA_E_STATE
    assert H.lib().ha_free_bytes_noalloc_c(None) == -1             # HA_E_INVALID
