"""Host-side logic of the request server, on the CPU: the HBM arena
allocator (first fit, alignment, coalescing, exhaustion) and the admission
rule (arrival order, budget, batch cap).  Driven through
libbmi_arena_check.so, a g++ build of the same header the server uses."""
import ctypes
import random

import pytest

NONE = ctypes.c_size_t(-1).value


@pytest.fixture(scope="module")
def A(pkg):
    path = pkg.PKG_DIR / "libbmi_arena_check.so"
    if not path.exists():
        pytest.fail(f"{path} missing: run __graft_entry__.build()")
    L = ctypes.CDLL(str(path))
    sz, vp = ctypes.c_size_t, ctypes.c_void_p
    L.bmi_arena_check_new.restype = vp
    L.bmi_arena_check_new.argtypes = [sz]
    L.bmi_arena_check_delete.argtypes = [vp]
    L.bmi_arena_check_alloc.restype = sz
    L.bmi_arena_check_alloc.argtypes = [vp, sz, sz]
    L.bmi_arena_check_release.argtypes = [vp, sz]
    for f in ("free_bytes", "largest_free", "used_blocks"):
        getattr(L, f"bmi_arena_check_{f}").restype = sz
        getattr(L, f"bmi_arena_check_{f}").argtypes = [vp]
    L.bmi_arena_check_admit.argtypes = [ctypes.POINTER(sz), ctypes.c_int, sz, ctypes.c_int]
    return L


def test_alloc_alignment_and_exhaustion(A):
    a = A.bmi_arena_check_new(4096)
    offs = [A.bmi_arena_check_alloc(a, 100, 256) for _ in range(16)]
    assert offs == [i * 256 for i in range(16)]
    assert A.bmi_arena_check_alloc(a, 1, 256) == NONE
    assert A.bmi_arena_check_free_bytes(a) == 0
    A.bmi_arena_check_delete(a)


def test_release_coalesces_neighbours(A):
    a = A.bmi_arena_check_new(1 << 20)
    x = [A.bmi_arena_check_alloc(a, 1 << 16, 256) for _ in range(16)]
    assert A.bmi_arena_check_alloc(a, 1, 256) == NONE
    # free every other block: no 2-block hole exists
    for o in x[::2]:
        assert A.bmi_arena_check_release(a, o) == 1
    assert A.bmi_arena_check_largest_free(a) == 1 << 16
    assert A.bmi_arena_check_alloc(a, 2 << 16, 256) == NONE
    # free one neighbour: blocks 0,1,2 merge into one range
    assert A.bmi_arena_check_release(a, x[1]) == 1
    assert A.bmi_arena_check_largest_free(a) == 3 << 16
    assert A.bmi_arena_check_alloc(a, 3 << 16, 256) == 0
    assert A.bmi_arena_check_release(a, 12345) == 0  # not a block
    A.bmi_arena_check_delete(a)


def test_random_alloc_free_keeps_accounting(A):
    cap = 1 << 24
    a = A.bmi_arena_check_new(cap)
    rng = random.Random(5)
    live = {}
    for _ in range(4000):
        if live and rng.random() < 0.45:
            o = rng.choice(list(live))
            assert A.bmi_arena_check_release(a, o) == 1
            del live[o]
        else:
            n = rng.randint(1, 1 << 18)
            o = A.bmi_arena_check_alloc(a, n, 256)
            if o == NONE:
                continue
            assert o % 256 == 0 and o + n <= cap
            for p, m in live.items():  # no overlap with a live block
                assert o + n <= p or p + m <= o
            live[o] = (n + 255) & ~255
        assert A.bmi_arena_check_free_bytes(a) == cap - sum(live.values())
    for o in list(live):
        A.bmi_arena_check_release(a, o)
    assert A.bmi_arena_check_largest_free(a) == cap
    assert A.bmi_arena_check_used_blocks(a) == 0
    A.bmi_arena_check_delete(a)


def _admit(A, sizes, avail, cap):
    arr = (ctypes.c_size_t * len(sizes))(*sizes)
    return A.bmi_arena_check_admit(arr, len(sizes), avail, cap)


def test_admission_in_arrival_order(A):
    assert _admit(A, [10, 20, 30], 100, 8) == 3
    assert _admit(A, [10, 20, 30], 59, 8) == 2
    # a large request at the head blocks the small ones behind it (no
    # overtaking, so it cannot starve)
    assert _admit(A, [80, 1, 1], 50, 8) == 0
    assert _admit(A, [1] * 20, 100, 8) == 8
    assert _admit(A, [], 100, 8) == 0
