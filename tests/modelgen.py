"""Seeded synthetic models used by the tests (written in NuShellX file formats)."""
from __future__ import annotations

from nushellx_amd.synth.interaction import random_interaction
from nushellx_amd.synth.mspace import PF, SD, Orbit, PNSpace

PSD = (Orbit(0, 1, 1), Orbit(0, 2, 5), Orbit(1, 0, 1))
from nushellx_amd.synth.writer import PNModel, SideCache, write_model

SEED = 2024


def make_models(outdir: str):
    cache = SideCache()
    sd = PNSpace(SD)
    inter = random_interaction(SD, seed=SEED)
    out = {}
    out["sd22"] = write_model(PNModel(sd, 2, 2, inter), outdir + "/sd22", "sd22", list(range(0, 17, 2)), cache=cache)
    out["sd44"] = write_model(PNModel(sd, 4, 4, inter), outdir + "/sd44", "sd44", list(range(0, 17, 2)), cache=cache)
    out["sd66"] = write_model(PNModel(sd, 6, 6, inter), outdir + "/sd66", "sd66", [0, 4, 8], cache=cache)
    out["sd23"] = write_model(PNModel(sd, 2, 3, inter), outdir + "/sd23", "sd23", [1, 3, 5], cache=cache)
    out["sd32"] = write_model(PNModel(sd, 3, 2, inter), outdir + "/sd32", "sd32", [1, 3], cache=cache)
    # a cross-shell space (0p1/2 with N = 1, 0d5/2 and 1s1/2 with N = 2) for the
    # N-hbar-omega window: positive parity, with and without the window 9..11
    psd = PNSpace(PSD)
    pint = random_interaction(PSD, seed=SEED + 1)
    out["psd33"] = write_model(PNModel(psd, 3, 3, pint), outdir + "/psd33", "psd33", [0, 2, 4], cache=cache,
                               tparity=1)
    out["psd33t"] = write_model(PNModel(psd, 3, 3, pint), outdir + "/psd33t", "pst33", [0, 2, 4], cache=cache,
                                tparity=1, window=(9, 11))
    out["psd33w8"] = write_model(PNModel(psd, 3, 3, pint), outdir + "/psd33w8", "psw33", [0], cache=cache,
                                 tparity=1, window=(8, 8))
    return out


def make_dops(models):
    """Separable operators (.dop) of every model's interaction, as NuDiop writes
    them: "exact" (all terms, 17 digits), "ref" (all terms, NuDiop's f11.7) and
    "cut" (|e|/sqrt(2lam+1) > 0.3 MeV for lam > 0, amplitudes > 0.05)."""
    from nushellx_amd.synth.diop import write_model_dop
    out = {}
    for name, mf in models.items():
        out[name] = {
            "exact": write_model_dop(mf.intfile, mf.intfile + "e", exact=True),
            "ref": write_model_dop(mf.intfile, mf.intfile + "r"),
            "cut": write_model_dop(mf.intfile, mf.intfile + "c", cut=0.3, mcut=0.05),
        }
    return out
