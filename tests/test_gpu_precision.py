"""Reduced-precision block contractions (FP32, BF16 hi/lo split) against the
FP64 oracle; the Lanczos around them stays FP64."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

# relative H*v error bounds (operand rounding: FP32 2^-24, BF16 hi/lo pair ~2^-16)
HV_TOL = {"fp32": 2e-6, "bf16x3": 2e-4}
# eigenvalue error bounds (MeV) of a converged Lanczos with the reduced-precision H*v:
# the 1e-4 MeV target of the mixed-precision solver for BF16 hi/lo
EV_TOL = {"fp32": 1e-5, "bf16x3": 1e-4}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_lowp_hv_error(models, prec):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(17)
    worst = 0.0
    for name in ("sd22", "sd44", "sd66"):
        mf = models[name]
        for J2 in mf.J2s[:3]:
            print(f"case {name} 2J={J2} {prec}", flush=True)
            args = mf.args(J2)
            ref = RefCPU(*args, fp64=True)
            h = Hamiltonian(*args, precision=prec)
            v = rng.standard_normal(h.N)
            w_ref = ref.apply(v)
            err = np.abs(h.multiply(v) - w_ref).max() / np.abs(w_ref).max()
            worst = max(worst, err)
            assert err < HV_TOL[prec], (name, J2, err)
    print(f"{prec}: worst relative H*v error {worst:.3e}")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_lowp_lanczos_eigenvalues(models, prec):
    from nushellx_amd.hv import Hamiltonian
    for name, J2 in (("sd44", 0), ("sd44", 8)):
        args = models[name].args(J2)
        H = RefCPU(*args, fp64=True).dense()
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:5]
        r = Hamiltonian(*args, precision=prec).lanczos(nev=5, tol=1e-6)
        err = np.abs(r["evals"] - ev)
        print(f"{prec} {name} 2J={J2}: eigenvalue errors {err}")
        assert err.max() < EV_TOL[prec], (name, J2, err)
