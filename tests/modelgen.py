"""Seeded synthetic models used by the tests (written in NuShellX file formats)."""
from __future__ import annotations

from nushellx_amd.synth.interaction import random_interaction
from nushellx_amd.synth.mspace import PF, SD, PNSpace
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
    return out
