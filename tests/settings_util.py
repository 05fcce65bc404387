"""Writes the five Settings_*.txt files in the layout of Set_Params.py
(one value per line, then free comment text; CodeExec as "name value")."""
import os

CODE_NAMES = ["ptON", "ptAdapt", "ptTrack", "ANNL_ON", "FFTW_ON", "RKCheck", "PBC_OFF",
              "IsiModel", "qcrystal", "with_recField", "with_DFTcodes", "force_Stripes",
              "refState_Zero", "qct_NeelPhase", "qct_RemovePBC", "rec_orderMaps"]


def write_settings(d, model, mc, smd, disorder, code):
    os.makedirs(os.path.join(d, "outBin"), exist_ok=True)

    def vals(name, values):
        with open(os.path.join(d, name), "w") as f:
            for v in values:
                f.write("%s\n" % v)
            f.write("\n Input data details:\n 0) ...\n")
    vals("Settings_MODEL.txt", model)
    vals("Settings_MC_SIM.txt", mc)
    vals("Settings_SMD_SIM.txt", smd)
    vals("Settings_Disorder.txt", disorder)
    with open(os.path.join(d, "Settings_CodeExec.txt"), "w") as f:
        for k, v in zip(CODE_NAMES, code):
            f.write("%s %d\n" % (k, v))


def default(d, **kw):
    model = [kw.get("L", 6), kw.get("J1", 1.0), kw.get("J2", 0.6), kw.get("JX", 0.0),
             kw.get("J2d", 0.0), kw.get("l1", 1.0), kw.get("l2", 1.0), kw.get("h", 0.0),
             kw.get("T1", 0.5), kw.get("T2", 2.0), kw.get("geom", "square")]
    mc = [kw.get("NTerm", 200), kw.get("NMeas", 2000), kw.get("IState", "high_T"),
          kw.get("TpGrid", "linear"), kw.get("tag", "J2tag"), kw.get("crSkip", 500), 0, 0]
    smd = [kw.get("ntm", 1280), kw.get("tpLstSz", 4), kw.get("tpIndex", 1), kw.get("dtm", 0.05),
           kw.get("win", "nowin"), 0, 0]
    dis = [kw.get("imp", 0.0), 1.0, 1.0, 0.0, 1.0, 0.0, 0.0]
    code = [kw.get("ptON", 1), 0, 0, 0, 0, 0, 0, kw.get("ising", 0), kw.get("qc", 0), kw.get("rec", 1),
            kw.get("dft", 1), 0, kw.get("refzero", 0), kw.get("neelphase", 0), 0, 0]
    write_settings(d, model, mc, smd, dis, code)
