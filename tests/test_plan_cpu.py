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
