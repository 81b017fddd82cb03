"""The allocation-free range evaluation (fbh_index_evaluate_noalloc) against
fbh_index_evaluate through the C ABI: two calls in a row, the second on a
smaller index, under the literal, scan and heuristic plans; the literal
result in a caller-owned device buffer equals the decompressed words of
the allocating entry point's result, and the counts are equal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


class DevBuf:
    """A caller-owned device buffer from the HIP runtime the library itself
    links (hipMalloc / hipMemcpy / hipFree through ctypes)."""

    def __init__(self, words):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(4 * words)) == 0
        self.words = words

    def read(self, n):
        out = np.zeros(n, np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.p, self.C.c_size_t(4 * n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _literal_of(bv, groups, anb):
    bv.decompress()
    w = bv.write()
    lit = np.zeros(groups + 1, np.uint32)
    lit[:groups] = w[:groups]
    if anb:
        assert int(w[-1]) == anb
        lit[groups] = (int(w[-2]) << (31 - anb)) & 0xFFFFFFFF
    return lit


def test_noalloc_matches_evaluate(L):
    rng = np.random.default_rng(2027)
    sizes = (31 * 9000 + 13, 31 * 500 + 7)           # the second call on a smaller index
    idx = []
    for n in sizes:
        col = rng.normal(size=n).astype(np.float32)
        idx.append((L.Index.bin_equal_weight(col, 120), n))
    cap = sizes[0] // 31 + 8
    buf = DevBuf(cap)
    for rep in range(2):
        for ix, n in idx:
            groups, anb = n // 31, n % 31
            for plan in (1, 3, 0):
                lo, hi = sorted(rng.normal(size=2).tolist())
                print("noalloc case", rep, n, plan, lo, hi, flush=True)
                h, used = ix.evaluate_noalloc(lo, 2, hi, 1, plan, buf.p.value, cap)
                got = buf.read(groups + 1)
                want_h, want_used, bv = ix.evaluate(lo, 2, hi, 1, plan=used, want_bitvector=True)
                assert used in (1, 3) and (plan == 0 or used == plan)
                assert h == want_h
                assert np.array_equal(got, _literal_of(bv, groups, anb)), (rep, n, plan)
    buf.free()


def test_noalloc_rejects_before_running(L):
    col = np.arange(31 * 100 + 3, dtype=np.float32)
    ix = L.Index.bin_equal_weight(col, 10)
    buf = DevBuf(200)
    with pytest.raises(L.native.FastbitError):
        ix.evaluate_noalloc(5.0, 2, 900.0, 1, 1, buf.p.value, 50)      # 101 words needed
    with pytest.raises(L.native.FastbitError):
        ix.evaluate_noalloc(5.0, 2, 900.0, 1, 2, buf.p.value, 200)     # compressed plan
    h, used = ix.evaluate_noalloc(5.0, 2, 900.0, 1, 1, buf.p.value, 200)
    assert h == 895 and used == 1
    buf.free()
