"""nsx_hv_expect_noalloc (states, images and Gram matrix in the library's static
device scratch, no device allocation) against nsx_hv_expect, which allocates its
buffers, and against the FP64 oracle: two calls in a row, the second on a
smaller input, then the size limit."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative; both paths run the same kernels, the pn part adds with FP64 atomics


@pytest.fixture(scope="module")
def small_models(tmp_path_factory):
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, SideCache, write_model
    out = str(tmp_path_factory.mktemp("noalloc"))
    cache = SideCache()
    sd = PNSpace(SD)
    inter = random_interaction(SD, seed=2024)
    return {
        "sd44": write_model(PNModel(sd, 4, 4, inter), out + "/sd44", "sd44", [2], cache=cache),
        "sd22": write_model(PNModel(sd, 2, 2, inter), out + "/sd22", "sd22", [0], cache=cache),
    }


def _noalloc(h, vecs):
    from nushellx_amd._lib import check, lib
    L = lib()
    vecs = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
    n = vecs.shape[0]
    M = np.empty((n, n), dtype=np.float64)
    check(L.nsx_hv_expect_noalloc(h.handle, n, vecs.ctypes.data, M.ctypes.data), "nsx_hv_expect_noalloc")
    return M.T.copy()


def test_expect_noalloc_matches_expect(small_models):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(11)
    for name, J2, nvec in (("sd44", 2, 12), ("sd22", 0, 3)):  # second call smaller (N and nvec)
        print(f"case noalloc {name} 2J={J2} nvec={nvec}", flush=True)
        args = small_models[name].args(J2)
        h = Hamiltonian(*args)
        V = rng.standard_normal((nvec, h.N))
        M0 = h.expect(V)
        M1 = _noalloc(h, V)
        scale = np.abs(M0).max()
        assert np.abs(M1 - M0).max() <= TOL * scale, (name, np.abs(M1 - M0).max() / scale)
        ref = RefCPU(*args, fp64=True)
        Mo = np.array([[ref.apply(V[i]) @ V[j] for j in range(nvec)] for i in range(nvec)])
        assert np.abs(M1 - Mo).max() <= TOL * np.abs(Mo).max()
        h.close()


def test_expect_noalloc_rejects_oversize(small_models):
    from nushellx_amd._lib import NsxError
    from nushellx_amd.hv import Hamiltonian
    h = Hamiltonian(*small_models["sd44"].args(2))
    nvec = (1 << 20) // (2 * h.N) + 1  # 2 N nvec > 2^20 doubles
    with pytest.raises(NsxError) as e:
        _noalloc(h, np.zeros((nvec, h.N)))
    assert e.value.code == -4 and "static scratch" in str(e.value)
    h.close()
