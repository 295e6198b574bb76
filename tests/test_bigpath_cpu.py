"""The large-block GEMM plan (X = V_R Bhat^T tiles, then W tiles over chunks of
a-terms, both restricted to the K steps where the operators are not zero),
executed on the CPU from the plan arrays and compared with the oracle.  Small
models reach it by lowering the block side above which it runs; J > 0 gives
a-terms with several X references, the separable operator other sparsity."""
import numpy as np

from oracle.refcpu import RefCPU
from tests.emulator import emu_hv_big


def test_big_plan_matches_oracle(models, dops):
    rng = np.random.default_rng(23)
    for name, J2, big_min, dop in (("sd44", 0, 8, None), ("sd44", 8, 8, None), ("sd44", 4, 16, None),
                                   ("sd23", 3, 4, None), ("sd44", 4, 8, "cut"), ("psd33t", 2, 4, None)):
        args = models[name].args(J2)
        dopfile = dops[name][dop] if dop else None
        ref = RefCPU(*args, fp64=True, dopfile=dopfile)
        v = rng.standard_normal(ref.N)
        w_ref = ref.apply(v)
        w, (t1, t2, nx, nt) = emu_hv_big(args, v, big_min, dopfile)
        assert t1 > 0 and t2 > 0, (name, J2, big_min)
        err = np.abs(w - w_ref).max() / np.abs(w_ref).max()
        assert err < 1e-12, (name, J2, big_min, dop, err)
        if J2 > 0 and dop is None:
            assert nx > nt  # some a-terms read several X products
