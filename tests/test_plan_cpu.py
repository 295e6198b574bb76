"""CPU checks of the device work plan: the plan executed by plain loops
(tests/plan_emulator.cpp) must equal the FP64 oracle H*v on the same files."""
import numpy as np

from oracle.refcpu import RefCPU
from tests.emulator import emu_hv, plan_stats


def test_plan_emulation_matches_refcpu(models):
    rng = np.random.default_rng(5)
    for name, mf in models.items():
        for J2 in mf.J2s:
            args = mf.args(J2)
            ref = RefCPU(*args, fp64=True)
            v = rng.standard_normal(ref.N)
            w = emu_hv(args, v)
            w_ref = ref.apply(v)
            err = np.abs(w - w_ref).max() / max(np.abs(w_ref).max(), 1e-300)
            assert err < 1e-12, (name, J2, err)


def test_plan_stats_sd44(models):
    st = plan_stats(models["sd44"].args(8))
    assert st["items2"] > 0 and st["items"] == 0  # every sd block takes the staged path
    assert st["lds_buf"] <= 5120


def test_plan_merge_equals_single_pass(models):
    """Instances planned separately and merged give bit-identical H*v, and the
    merge deduplicates operator blocks shared between instances."""
    from tests.emulator import emu_hv_batch
    args = [models["sd44"].args(0), models["sd44"].args(4), models["sd22"].args(2), models["sd44"].args(8)]
    n = sum(RefCPU(*a, fp64=True).N for a in args)
    v = np.random.default_rng(9).standard_normal(n)
    w1, s1 = emu_hv_batch(args, v, merge=False)
    w2, s2 = emu_hv_batch(args, v, merge=True)
    assert np.array_equal(w1, w2)
    assert (s1 == s2).all()
    off = 0
    for a in args:
        r = RefCPU(*a, fp64=True)
        ref = r.apply(v[off:off + r.N])
        assert np.abs(w2[off:off + r.N] - ref).max() <= 1e-12 * np.abs(ref).max()
        off += r.N


def test_stacked_plan_matches_refcpu(models):
    """The stacked two-stage plan (shared X = V Bhat^T, then W += A sum r X)
    gives the oracle's H*v, for several chunk capacities."""
    from tests.emulator import emu_hv_stacked
    rng = np.random.default_rng(11)
    for name, J2s in (("sd22", [0, 4]), ("sd44", [0, 8]), ("sd23", [3])):
        for J2 in J2s:
            args = models[name].args(J2)
            r = RefCPU(*args, fp64=True)
            v = rng.standard_normal(r.N)
            ref = r.apply(v)
            for cap, work in ((512, 64), (2048, 512), (8192, 4096)):
                w = emu_hv_stacked(args, v, cap, work)
                assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), (name, J2, cap)


def test_mstk_plan_matches_refcpu(models):
    """The matrix-core stacked plan, executed lane by lane as its kernels run
    it (16 x 16 x 4 steps with their operand and result lane layouts, X stored
    transposed, four wavefronts per stage-2 tile), gives the oracle's H*v and
    stays inside every array, for several task sizes, a truncated basis and
    blocks up to the small-path limit of 40 (inner sizes of 10 K steps)."""
    from tests.emulator import emu_hv_mstk
    rng = np.random.default_rng(17)
    cases = [("sd22", 0), ("sd22", 4), ("sd44", 0), ("sd44", 8), ("sd66", 4), ("sd23", 3), ("psd33t", 2)]
    for name, J2 in cases:
        args = models[name].args(J2)
        r = RefCPU(*args, fp64=True)
        v = rng.standard_normal(r.N)
        ref = r.apply(v)
        for units1, terms2 in ((1, 1), (4, 16), (64, 5)):
            w, st = emu_hv_mstk(args, v, units1, terms2)
            assert st[0] > 0 and st[1] > 0
            assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), (name, J2, units1, terms2)


def test_mxf_plan_matches_refcpu(models):
    """The fused matrix-core plan, executed lane by lane as its kernel runs it
    (X_u tiles on 16 x 16 x 4 steps, Y_s = sum r X_u in the result layout fed
    back as the next product's K steps, four wavefronts per task), gives the
    oracle's H*v and stays inside every array, on blocks up to 33 (sd 6p6n:
    several tiles per side and ten K steps), a truncated basis and an odd J,
    with the like-particle part stored first or folded into the tasks (two
    more slabs per block: V I against H_a, V H_b^T against I).  The tasks are
    reached through the kernel's workgroup map (workgroup b: the (b / 8)-th
    task of XCD b % 8's range), which must run every task exactly once; in
    the deterministic variant every used slot's partial tile is summed into
    its W tile exactly once."""
    from tests.emulator import emu_hv_mxf
    rng = np.random.default_rng(23)
    cases = [("sd22", 0), ("sd22", 4), ("sd44", 0), ("sd44", 8), ("sd66", 4), ("sd23", 3), ("psd33t", 2)]
    for name, J2 in cases:
        args = models[name].args(J2)
        r = RefCPU(*args, fp64=True)
        v = rng.standard_normal(r.N)
        ref = r.apply(v)
        for steps, slots, like, det in ((0, 4, 0, 0), (1, 4, 1, 0), (3, 2, 1, 0), (1, 2, 0, 0), (2, 4, 1, 1),
                                        (1, 2, 0, 1)):
            w, st = emu_hv_mxf(args, v, steps, slots, like, det)
            assert st[0] > 0 and st[1] > 0
            # folded where the plan allows it: not on the truncated basis
            assert (st[3] >= 100) == (like == 1 and name != "psd33t"), (name, st)
            assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), (name, J2, steps, slots, like, det)


def test_fused_plan_matches_refcpu(models):
    """The fused column plan (per task: rounds of V_R / Bhat^T / A staging,
    X = V Bhat^T, Y = sum r X, W += A Y into the column's outputs; then the
    reduce over the column partials) gives the oracle's H*v, with LDS caps and
    task sizes that force many rounds and many tasks per column, on a
    truncated basis, and beside large-block items."""
    from tests.emulator import emu_hv_fused
    rng = np.random.default_rng(13)
    cases = [("sd22", 0), ("sd22", 4), ("sd44", 0), ("sd44", 8), ("sd66", 4), ("sd23", 3), ("psd33t", 2)]
    for name, J2 in cases:
        args = models[name].args(J2)
        r = RefCPU(*args, fp64=True)
        v = rng.standard_normal(r.N)
        ref = r.apply(v)
        for cap, work in ((256, 2000), (1024, 20000), (3072, 250000), (16384, 10 ** 8)):
            w, st = emu_hv_fused(args, v, cap, work)
            assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), (name, J2, cap, work)
            assert st[0] > 0 and st[1] >= st[0]


def test_fused_plan_beside_large_blocks(models, monkeypatch):
    from tests.emulator import emu_hv_fused
    monkeypatch.setenv("NSX_BIG_MIN", "20")
    args = models["sd44"].args(4)
    r = RefCPU(*args, fp64=True)
    v = np.random.default_rng(2).standard_normal(r.N)
    ref = r.apply(v)
    w, st = emu_hv_fused(args, v, 2048, 50000)
    assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max()
    assert st[0] > 0
