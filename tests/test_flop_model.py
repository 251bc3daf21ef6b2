"""The frozen per-pair flop model (libmobility_amd/flop_model.json) equals the
operation count of the reference-form scalar restatement (oracle/rpy_scalar.py)
for a far-field, non-self pair (+,-,*,/,sqrt = 1; an FMA therefore 2)."""
import json
import os

from oracle import rpy_scalar
from tests.conftest import ROOT

VARIANTS = {
    "open_UF": (0, ("UF",)),
    "open_UF_WF": (0, ("UF", "WF")),
    "open_UT_WT": (0, ("UT", "WT")),
    "open_6x6": (0, ("UF", "UT", "WT", "WF")),
    "wall_UF": (1, ("UF",)),
    "wall_UF_WF": (1, ("UF", "WF")),
    "wall_UT_WT": (1, ("UT", "WT")),
    "wall_6x6": (1, ("UF", "UT", "WT", "WF")),
}


def test_flop_model_frozen():
    model = json.load(open(os.path.join(ROOT, "libmobility_amd", "flop_model.json")))
    for name, (wall, blocks) in VARIANTS.items():
        assert model["flops_per_pair"][name] == rpy_scalar.count_pair_flops(wall, set(blocks)), name
