"""Copies the reference's tabulated quasicrystal-approximant data files
(QCrystal_Data/*) used as fixtures for the tabulated-graph loader.

These are data (neighbour tables, site coordinates, Ising-like ordered states
and per-site shell counts written by QCrystal_Data/QCrystal_Data_toMC_PBC*.py),
not source.  Run once in the build container, where /root/reference exists:
    python tests/golden/make_golden.py
"""
import os
import shutil

REF = "/root/reference/QCrystal_Data"
HERE = os.path.dirname(os.path.abspath(__file__))

SETS = [("with_PBC", 41), ("no_PBC", 239)]
KINDS = ["nborsList", "sitesList", "stateNeel0", "stateStrp1", "stateStrp2", "stateStrp3",
         "stateStrp4"]


def main():
    for pbc, n in SETS:
        dst = os.path.join(HERE, "qc_%s_%d" % (pbc, n))
        os.makedirs(dst, exist_ok=True)
        src = os.path.join(REF, pbc, str(n))
        for k in KINDS:
            f = "%s_AppxtSz(%d).bin" % (k, n)
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        f = "zvalsList_AppxtSz(%d).dat" % n
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))


if __name__ == "__main__":
    main()
