"""Loaders of the golden fixtures (tests/golden/*.npz)."""
import numpy as np

from conftest import GOLDEN
from rc_amd.tables import GCFTable, VisChunks

RANDOM_FIXTURES = ["random_0", "random_1", "random_2", "random_3"]


def load_random(name):
    z = np.load(GOLDEN / f"{name}.npz")
    vc = VisChunks(uvw=z["uvw"], vis=z["vis_in"], supps=z["supps"],
                   scale=float(z["scale"]), wstep=float(z["wstep"]))
    gcf = GCFTable(data=z["gcf_data"], layer_off=z["gcf_off"], layer_len=z["gcf_len"],
                   center=int(z["gcf_center"]), supps=z["gcf_supps"],
                   max_w_plane=int(z["gcf_M"]), half=bool(z["half"]))
    return vc, gcf, int(z["N"]), z


def load_probe():
    """[(half, scale, wstep, N, pts[n,3], S[n], res[n,5])]"""
    z = np.load(GOLDEN / "pregrid_probe.npz")
    out = []
    for i in range(int(z["n_cases"])):
        g = lambda k: z[f"c{i}_{k}"]
        out.append((bool(g("half")), float(g("scale")), float(g("wstep")), int(g("N")),
                    g("pts"), g("S"), g("res")))
    return out
