"""The duplicate filter's entry point that allocates nothing
(pq_dedup_reads_noalloc: device pointers in and out, scratch in a fixed device
array) against the allocating entry point (pq_dedup_reads) and the oracle's
sequential filter, on two calls in a row, the second smaller (the fixed scratch
keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc

pytestmark = pytest.mark.gpu


def _records(rng, nrec, L, dup_frac):
    # records with duplicates and reverse-complement duplicates of earlier ones
    # and a few with N (ok = 0)
    base = rng.integers(1, 5, (nrec, L)).astype(np.uint8)
    for r in range(1, nrec):
        u = rng.random()
        if u < dup_frac:
            base[r] = base[rng.integers(0, r)]
        elif u < 1.5 * dup_frac:
            base[r] = (5 - base[rng.integers(0, r)][::-1]).astype(np.uint8)
    ok = (rng.random(nrec) > 0.01).astype(np.uint8)
    return base.reshape(-1), ok


class _Dev:
    """Device buffer of the library's device (pq_dev_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.ptr = ctx, C.c_void_p()
        ctx._check(ctx._L.pq_dev_alloc(ctx._h, max(1, nbytes), C.byref(self.ptr)), "dev_alloc")

    def put(self, a):
        a = np.ascontiguousarray(a)
        self.ctx._check(self.ctx._L.pq_dev_upload(self.ctx._h, self.ptr, a.ctypes.data, a.nbytes),
                        "dev_upload")

    def get(self, n):
        out = np.zeros(n, np.uint8)
        self.ctx._check(self.ctx._L.pq_dev_download(self.ctx._h, out.ctypes.data, self.ptr, n),
                        "dev_download")
        return out

    def free(self):
        self.ctx._L.pq_dev_free(self.ptr)


def _noalloc(ctx, dev, codes, ok, L, bits, nf, width):
    nrec = len(ok)
    dev["codes"].put(codes)
    dev["ok"].put(ok)
    passes = C.c_uint32()
    ctx._check(ctx._L.pq_dedup_reads_noalloc(ctx._h, dev["codes"].ptr, dev["ok"].ptr, nrec, L,
                                             bits, nf, width, dev["keep"].ptr, C.byref(passes)),
               "dedup_reads_noalloc")
    return dev["keep"].get(nrec), passes.value


@pytest.mark.parametrize("width", [32, 64])
def test_noalloc_dedup_matches_allocating_entry_point(width):
    rng = np.random.default_rng(width)
    L, nf = 60, 2
    with pq.Context(0) as ctx:
        big, small = 60000, 7000
        dev = {"codes": _Dev(ctx, big * L), "ok": _Dev(ctx, big), "keep": _Dev(ctx, big)}
        try:
            for nrec, bloom_m in ((big, 36), (small, 3)):
                codes, ok = _records(rng, nrec, L, 0.2)
                bits = nrec * bloom_m
                got, passes = _noalloc(ctx, dev, codes, ok, L, bits, nf, width)
                want, wpasses = ctx.dedup_reads(codes, ok, L, bits, nf, width)
                assert np.array_equal(got, want), (nrec, bloom_m)
                assert passes == wpasses
                assert np.array_equal(got, orc.dedup(codes, ok, L, bits, nf, width))
                assert 0 < got.sum() < nrec
            # beyond the fixed scratch: refused before any device work
            with pytest.raises(pq.PqError, match="fixed scratch"):
                _noalloc(ctx, dev, codes, ok, L, pq.NOALLOC_SCRATCH_BYTES, nf, width)
        finally:
            for d in dev.values():
                d.free()
