"""pf shell, 4 protons + 4 neutrons, J = 0 (N = 41,355): block sides up to 99,
so the direct (non-staged) pn kernel runs beside the staged one.  H*v against
the FP64 oracle, and the device Lanczos against ARPACK (scipy eigsh) driven by
the oracle's H*v."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu


def test_pf44_hv_matches_oracle(pf44):
    from nushellx_amd.hv import Hamiltonian
    args = pf44.args(0)
    ref = RefCPU(*args, fp64=True)
    assert ref.N == 41355
    rng = np.random.default_rng(5)
    v = rng.standard_normal(ref.N)
    w_ref = ref.apply(v)
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)):
        print(f"case pf44 J=0 {prec}", flush=True)
        h = Hamiltonian(*args, precision=prec)
        assert h.N == ref.N
        err = np.abs(h.multiply(v) - w_ref).max() / np.abs(w_ref).max()
        print(f"{prec}: relative error {err:.2e}", flush=True)
        assert err < tol, (prec, err)
        h.close()


@pytest.mark.parametrize("group", ["16", "0", "1,4"])
def test_pf44_bf16_panel_tiles(pf44, monkeypatch, group):
    """BF16 hi/lo large blocks on the panel tiles (operators and V_R as
    step-major hi/lo units, tiles in XCD runs of the given lengths) against the
    tiles on the Bf16x2 stores: the same products in the same order per tile,
    so the two agree to the FP32 accumulation's rounding (the FP64 atomic adds
    may come in another order); both against the FP64 oracle."""
    from nushellx_amd.hv import Hamiltonian
    args = pf44.args(0)
    ref = RefCPU(*args, fp64=True)
    v = np.random.default_rng(8).standard_normal(ref.N)
    w_ref = ref.apply(v)
    out = {}
    for panels in ("1", "0"):
        print(f"case pf44 bf16 panels={panels} group={group}", flush=True)
        monkeypatch.setenv("NSX_BF_PANELS", panels)
        monkeypatch.setenv("NSX_BIG_XCD_GROUP", group)
        h = Hamiltonian(*args, precision="bf16x3")
        out[panels] = h.multiply(v)
        h.close()
        err = np.abs(out[panels] - w_ref).max() / np.abs(w_ref).max()
        assert err < 2e-4, (panels, err)
    assert np.abs(out["1"] - out["0"]).max() <= 1e-9 * np.abs(w_ref).max()


def test_pf44_autotuned_large_block_tiles(pf44):
    """The autotuner also picks the large-block tile shape (stage-2 K chunk;
    stage-1 rows in BF16 hi/lo); the tuned plan still equals the oracle."""
    from nushellx_amd.hv import Hamiltonian
    args = pf44.args(0)
    ref = RefCPU(*args, fp64=True)
    v = np.random.default_rng(6).standard_normal(ref.N)
    w_ref = ref.apply(v)
    for prec, tol in (("fp64", 1e-12), ("bf16x3", 2e-4)):
        print(f"case pf44 autotune {prec}", flush=True)
        h = Hamiltonian(*args, precision=prec)
        cfg = h.autotune(iters=2)
        print(f"{prec}: {cfg}", flush=True)
        assert cfg["big_kchunk"] in (32, 64, 128, 256)
        assert cfg["big_rows1"] in ((64,) if prec == "fp64" else (64, 128))
        err = np.abs(h.multiply(v) - w_ref).max() / np.abs(w_ref).max()
        assert err < tol, (prec, err)
        h.close()


def test_pf44_lanczos_matches_arpack(pf44):
    from scipy.sparse.linalg import LinearOperator, eigsh

    from nushellx_amd.hv import Hamiltonian
    args = pf44.args(0)
    ref = RefCPU(*args, fp64=True)
    mask = ref.lcoef().astype(float)
    # restrict to the parity-allowed states, as the solver's start vector does
    op = LinearOperator((ref.N, ref.N), matvec=lambda x: mask * ref.apply(mask * np.ravel(x)), dtype=np.float64)
    v0 = mask * np.random.default_rng(1).standard_normal(ref.N)
    ev_ref = np.sort(eigsh(op, k=5, which="SA", tol=1e-11, v0=v0)[0])
    h = Hamiltonian(*args)
    r = h.lanczos(nev=5, tol=1e-7)
    print("device", r["evals"], "arpack", ev_ref, r["matvecs"], flush=True)
    assert np.abs(r["evals"] - ev_ref).max() < 1e-6


@pytest.mark.parametrize("name,J2,big_min,dop", [("sd44", 0, 8, None), ("sd44", 8, 8, None), ("sd23", 3, 4, None),
                                                 ("sd44", 4, 8, "cut"), ("psd33t", 2, 4, None)])
def test_large_block_path_on_small_models(models, dops, name, J2, big_min, dop, monkeypatch):
    """The large-block GEMM kernels (MFMA tiles, K-step masks, split-K atomics,
    several X references per a-term when J > 0) on sd models, a separable
    operator and a truncated (N-hbar-omega window) basis, reached by lowering
    the block side above which they run."""
    from nushellx_amd.hv import Hamiltonian
    monkeypatch.setenv("NSX_BIG_MIN", str(big_min))
    args = models[name].args(J2)
    dopfile = dops[name][dop] if dop else None
    print(f"case large-block {name} 2J={J2} big_min={big_min} dop={dop}", flush=True)
    ref = RefCPU(*args, fp64=True, dopfile=dopfile)
    v = np.random.default_rng(J2 + big_min).standard_normal(ref.N)
    w_ref = ref.apply(v)
    # FP64 tiles, and the FP32 tiles that FP32 and BF16 hi/lo operands run on
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)):
        h = Hamiltonian(*args, precision=prec, dopfile=dopfile)
        err = np.abs(h.multiply(v) - w_ref).max() / np.abs(w_ref).max()
        assert err <= tol, (prec, err)
        h.close()
