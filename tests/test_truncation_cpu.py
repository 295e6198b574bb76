"""N-hbar-omega truncated bases (min_MSon/max_MSon of the .nus), on the CPU.

Pins:
* closed form: in 0p1/2-0d5/2-1s1/2 with three protons and three neutrons, the
  window N = 8 keeps only (0p1/2)^2 x one sd nucleon on each side, whose J = 0
  states are (5/2 x 5/2) and (1/2 x 1/2): exactly two, both of positive parity;
* the window selects blocks: on a vector inside the window the truncated
  oracle's H*v equals the full oracle's on every block with a partition pair
  inside the window and is zero on the others (NuTRL's bas_log);
* the window masks elements of the pn input only (single_productm gathers the
  lcoef states and adds the window states, NuCommonX.f90:760-812): for any
  vector v, H_t v = H_t (v in the window) + L_t (v outside it), with L_t the
  like-particle part alone (the same files with every .oph value zeroed);
* the product's plan (CPU emulator) equals the truncated oracle;
* the library's host-only handle reports the same window mask (no GPU).
"""
import ctypes as C

import numpy as np

from oracle.refcpu import RefCPU
from tests.emulator import emu_hv


def test_window_closed_form(models):
    r = RefCPU(*models["psd33w8"].args(0))
    t, m = r.trunc()
    assert t and m.sum() == 2


def _like_only(mf, J2, work):
    """The same model with every pn multipole zeroed: its H*v is the
    like-particle part alone."""
    import os
    from nushellx_amd.synth.diop import read_oph
    from nushellx_amd.synth.fortio import RecordWriter
    src = os.path.dirname(mf.intfile)
    os.makedirs(work, exist_ok=True)
    for f in os.listdir(src):
        if not os.path.lexists(os.path.join(work, f)):
            os.symlink(os.path.join(src, f), os.path.join(work, f))
    oph = os.path.join(work, os.path.basename(mf.intfile) + ".oph")
    items, max_lmo = read_oph(oph)
    os.unlink(oph)
    with RecordWriter(oph) as w:
        w.write(len(items), max_lmo)
        for key, _, pl, pr in items:
            w.write(key, np.float64(0.0), pl, pr)
    return [os.path.join(work, os.path.basename(a)) for a in mf.args(J2)]


def test_truncated_hv_selects_blocks(models):
    rng = np.random.default_rng(13)
    for J2 in models["psd33t"].J2s:
        full = RefCPU(*models["psd33"].args(J2))
        tr = RefCPU(*models["psd33t"].args(J2))
        assert full.N == tr.N
        t, m = tr.trunc()
        assert t and 0 < m.sum() < full.lcoef().sum()
        assert not (m & ~full.lcoef()).any()
        v = rng.standard_normal(full.N) * m
        wf, wt = full.apply(v), tr.apply(v)
        active = inactive = 0
        for (a, b, p, d) in tr.blocks():
            seg = slice(p, p + d)
            if np.any(wt[seg] != 0):
                active += 1
                assert np.abs(wt[seg] - wf[seg]).max() <= 1e-13 * np.abs(wf).max()
            else:
                inactive += 1
        assert active > 0
        # every block holding a window state is computed
        for (a, b, p, d) in tr.blocks():
            if m[p:p + d].any():
                assert np.any(wt[p:p + d] != 0)


def test_truncated_hv_masks_pn_input(models, tmp_path):
    rng = np.random.default_rng(19)
    for J2 in models["psd33t"].J2s:
        tr = RefCPU(*models["psd33t"].args(J2))
        like = RefCPU(*_like_only(models["psd33t"], J2, str(tmp_path / f"like{J2}")))
        _, m = tr.trunc()
        v = rng.standard_normal(tr.N)
        w = tr.apply(v)
        w_in, w_out = tr.apply(v * m), like.apply(v * ~m)
        assert np.abs(w - (w_in + w_out)).max() <= 1e-13 * np.abs(w).max(), J2
        # the pn part does see v outside the window without the mask
        assert np.abs(w - tr.apply(v * m) - tr.apply(v * ~m)).max() <= 1e-13 * np.abs(w).max()
        assert np.abs(tr.apply(v * ~m) - w_out).max() <= 1e-13 * np.abs(w).max()


def test_truncated_plan_matches_oracle(models):
    rng = np.random.default_rng(17)
    for name in ("psd33t", "psd33w8"):
        for J2 in models[name].J2s:
            args = models[name].args(J2)
            r = RefCPU(*args)
            v = rng.standard_normal(r.N)
            w = emu_hv(args, v)
            ref = r.apply(v)
            assert np.abs(w - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0), (name, J2)


def test_library_window_mask_without_gpu(models):
    from nushellx_amd import _lib
    L = _lib.lib()
    args = models["psd33t"].args(2)
    h = C.c_void_p()
    rc = L.nsx_hv_create(*[a.encode() for a in args], _lib.PRECISIONS["fp64"] | _lib.NSX_HOST_ONLY, C.byref(h))
    assert rc == 0, L.nsx_last_error()
    n, basd = C.c_int64(), C.c_int32()
    L.nsx_hv_dims(h, C.byref(n), C.byref(basd))
    mask = np.zeros(n.value, dtype=np.uint8)
    assert L.nsx_hv_trunc(h, mask.ctypes.data) == 1
    t, m = RefCPU(*args).trunc()
    assert np.array_equal(mask.astype(bool), m)
