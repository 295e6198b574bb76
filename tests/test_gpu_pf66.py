"""pf shell, 6 protons + 6 neutrons, J = 0 (N = 1,777,116; blocks up to 615 x
615): the largest workload, on the large-block matrix-core path.  H*v against
the FP64 oracle on one vector, and size-independent properties of the operator
at full size: symmetry <u, Hv> = <Hu, v> and linearity."""
import tempfile

import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pf66():
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    print("generating pf 6p6n J=0 operator files", flush=True)
    inter = random_interaction(PF, seed=2024, spe_range=(-10.0, -4.0))
    return write_model(PNModel(PNSpace(PF), 6, 6, inter), tempfile.mkdtemp(prefix="nsx_pf66_"), "pf66", [0])


def test_pf66_hv_matches_oracle_and_properties(pf66):
    from nushellx_amd.hv import Hamiltonian
    args = pf66.args(0)
    rng = np.random.default_rng(66)
    print("case pf66 J=0 fp64 setup", flush=True)
    h = Hamiltonian(*args)
    assert h.N == 1777116
    u, v = rng.standard_normal(h.N), rng.standard_normal(h.N)
    hu, hv = h.multiply(u), h.multiply(v)
    print("case pf66 J=0 oracle H*v", flush=True)
    ref = RefCPU(*args, fp64=True)
    w_ref = ref.apply(v)
    scale = np.abs(w_ref).max()
    err = np.abs(hv - w_ref).max() / scale
    print(f"pf66 fp64: relative error {err:.2e}", flush=True)
    assert err <= 1e-12
    # symmetry and linearity at full size
    assert abs(u @ hv - hu @ v) <= 1e-12 * np.linalg.norm(hu) * np.linalg.norm(v)
    hl = h.multiply(2.0 * u - 3.0 * v)
    assert np.abs(hl - (2.0 * hu - 3.0 * hv)).max() <= 1e-12 * np.abs(hl).max()
    h.close()
    print("case pf66 J=0 fp32", flush=True)
    h32 = Hamiltonian(*args, precision="fp32")
    err32 = np.abs(h32.multiply(v) - w_ref).max() / scale
    print(f"pf66 fp32: relative error {err32:.2e}", flush=True)
    assert err32 < 2e-6
    h32.close()
