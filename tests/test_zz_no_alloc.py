"""tribit_count_noalloc: the count that allocates nothing (its task lists,
counters and result live in the library's per-GPU scratch array), called
through the C ABI and checked against tribit_count with the same options:
equal exact counts and equal task splits, on two calls in a row, the second
on a smaller graph (the array keeps the first call's contents), for every
schedule it accepts. With an allocation cap of one byte tribit_count fails
and tribit_count_noalloc still counts."""
import ctypes

import pytest

import tribit_amd

pytestmark = pytest.mark.gpu

SCHEDS = ["wedge", "directed", "plain"]


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def noalloc(ctx, g, sched, flags=0):
    o = tribit_amd.CountOpts()
    o.sched = tribit_amd.SCHED[sched]
    o.repeats = 1
    o.flags = flags
    ci = tribit_amd.CountInfo()
    st = ctx.lib.tribit_count_noalloc(ctx.h, g.h, ctypes.byref(o), ctypes.byref(ci))
    return st, ci


@pytest.mark.parametrize("sched", SCHEDS)
def test_noalloc_matches_count(gpu_ctx, sched):
    assert gpu_ctx.lib.tribit_noalloc_scratch_bytes() == 128 << 20
    with gpu_ctx.gen_rmat(18, 16, seed=11) as big, gpu_ctx.gen_rmat(13, 12, seed=1) as small:
        pairs = ((big, "rmat s18"), (small, "rmat s13 (smaller, second call)"))
        wants = [g.count(sched=sched) for g, _ in pairs]
        gots = []
        for g, name in pairs:  # the two no-allocation calls back to back
            note(f"noalloc {sched} {name}")
            gots.append(noalloc(gpu_ctx, g, sched))
        for (st, got), want in zip(gots, wants):
            assert st == 0, gpu_ctx.lib.tribit_status_str(st)
            assert got.triangles == want.triangles
            assert got.tasks == want.tasks
            nv = got.nvariants
            assert nv == len(want.ms_by_variant) and got.ms_count > 0
            assert list(got.tasks_by_variant)[:nv] == list(want.tasks_by_variant)[:nv]
        # forced multi-task waves: the per-task counters come from the same array
        st, got = noalloc(gpu_ctx, big, sched, tribit_amd.FORCE_MULTI)
        assert st == 0 and got.triangles == big.count(sched=sched).triangles


def test_noalloc_under_allocation_cap(gpu_ctx):
    with gpu_ctx.gen_rmat(16, 16, seed=3) as g:
        want = g.count(sched="wedge").triangles
        st, _ = noalloc(gpu_ctx, g, "directed")
        assert st == 6  # the directed layout is built on first use: a count that allocates builds it
        assert g.count(sched="directed").triangles == want
        gpu_ctx.set_alloc_cap(1)  # every pool allocation through the context now fails
        try:
            with pytest.raises(tribit_amd.TribitError):
                g.count(sched="wedge")
            st, got = noalloc(gpu_ctx, g, "wedge")
            assert st == 0 and got.triangles == want
            st, got = noalloc(gpu_ctx, g, "directed")
            assert st == 0 and got.triangles == want
        finally:
            gpu_ctx.set_alloc_cap(0)


def test_noalloc_rejects_what_would_allocate(gpu_ctx):
    with gpu_ctx.gen_rmat(12, 8, seed=2) as g:
        st, _ = noalloc(gpu_ctx, g, "learned")
        assert st == 6  # TRIBIT_ERR_ARG
        st, _ = noalloc(gpu_ctx, g, "wedge", tribit_amd.FORCE_TILES)
        assert st == 6
