"""The drop-in runs (xMC_run / xSMD_run equivalents): settings guard rails on
the CPU, the full MC -> SMD file pipeline on the GPU."""
import os

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref
from settings_util import default


@pytest.mark.parametrize("kw,msg", [
    (dict(geom="pentagon"), "Unknown geometry"),
    (dict(L=7), "multiple of 2"),
    (dict(geom="triang", L=8), "multiple of 3"),
    (dict(NMeas=2500), "binning size"),
    (dict(NMeas=1000), "greater than 1"),
    (dict(ising=1, h=0.3), "Ising"),
    (dict(crSkip=300), "divisor of BinSz"),
    (dict(dtm=0.08), "too large"),
    (dict(win="boxcar"), "temporal window"),
    (dict(J2d=0.1), "Lieb"),
    (dict(IState="2up1dw"), "not compatible"),
])
def test_settings_guard_rails(tmp_path, kw, msg):
    default(str(tmp_path), **kw)
    with pytest.raises(H.HSMDError) as e:
        H.run_mc(str(tmp_path), 4)
    assert e.value.code == 2 and msg in str(e.value), str(e.value)


def test_smd_frequency_resolution_guard(tmp_path):
    # dwf = 2 pi / (ntm dtm) must not exceed 0.10 (CF/smd_code.cpp:2120-2131)
    default(str(tmp_path), ntm=64, dtm=0.02)
    with pytest.raises(H.HSMDError) as e:
        H.run_smd(str(tmp_path))
    assert e.value.code == 2 and "dwf" in str(e.value)


def test_missing_settings_file(tmp_path):
    with pytest.raises(H.HSMDError) as e:
        H.run_mc(str(tmp_path), 4)
    assert e.value.code == 4


@pytest.mark.gpu
def test_mc_then_smd_pipeline_files_and_dssf(tmp_path):
    d = str(tmp_path)
    # a short evolution (extended mode: dwf above the reference's 0.10) keeps
    # the chaotic trajectories comparable point by point
    default(d, L=6, NTerm=200, NMeas=2000, crSkip=500, ntm=96, dtm=0.02, tpLstSz=4, tpIndex=1)
    H.run_mc(d, 4, precision=H.BATCHED_F64, seed=3)
    label = "_J2(0.600)"
    T = np.fromfile(os.path.join(d, "outBin", "Temp_List%s.bin" % label))
    np.testing.assert_allclose(T, [0.5, 1.0, 1.5, 2.0])
    N = 36
    for t in T:
        tag = "_T(%.5f)%s.bin" % (t, label)
        v1 = np.fromfile(os.path.join(d, "outBin", "MC_Bin_AvgValue1" + tag))
        assert v1.size == 2 * 23
        v1 = v1.reshape(2, 23)
        # energy density below the paramagnet's and bounded by the ground state
        assert np.all(v1[:, 0] < 0.1) and np.all(v1[:, 0] > -1.25)
        v2 = np.fromfile(os.path.join(d, "outBin", "MC_Bin_AvgValue2" + tag))
        assert v2.size == 80
        v2 = v2.reshape(2, 40)
        # S(q) peaks (xy, yz, zz, tt at five points) are positive, lengths >= 0
        assert np.all(v2[:, :20] > 0) and np.all(v2[:, 20:] >= 0)
        # spin stiffness RhoS1 (a bond sum over every sweep) is measured
        assert np.all(v1[:, 20] != 0)
        # MC_Bin_w0SqField: int32 count (every 10th sweep of 2 bins) + Nq x 3 complex
        raw = open(os.path.join(d, "outBin", "MC_Bin_w0SqField" + tag), "rb").read()
        assert len(raw) == 4 + N * 3 * 16
        assert np.frombuffer(raw[:4], np.int32)[0] == 200
        w0 = np.frombuffer(raw[4:], np.complex128).reshape(N, 3)
        assert np.all(np.abs(w0) > 0)
        sf = np.fromfile(os.path.join(d, "outBin", "MC_Bin_SpinField" + tag)).reshape(-1, N, 3)
        assert sf.shape[0] == 2000 // 500
        np.testing.assert_allclose(np.linalg.norm(sf, axis=-1), 1.0, atol=1e-12)
    # MC_Stat_Results: one row per temperature, T and the label value, then
    # (mean, bootstrap deviation) of every Values1 entry over the bins
    st = np.loadtxt(os.path.join(d, "outData", "MC_Stat_Results1%s.dat" % label))
    assert st.shape == (4, 2 + 2 * 23)
    np.testing.assert_allclose(st[:, 0], T, rtol=1e-12)
    assert np.all(st[:, 1] == 0.6)
    for k, t in enumerate(T):
        v1 = np.fromfile(os.path.join(d, "outBin", "MC_Bin_AvgValue1_T(%.5f)%s.bin" % (t, label)))
        m, sd = H.bin_statistics(v1.reshape(-1, 23), bootstrap=True, seed=3 + k)
        np.testing.assert_allclose(st[k, 2::2], m, rtol=1e-11, atol=1e-300)
        np.testing.assert_allclose(st[k, 3::2], sd, rtol=1e-11, atol=1e-300)
    fx = np.loadtxt(os.path.join(d, "outData", "MC_Collect", "MC_Stat_FixTemp1(%.5f).dat" % T[2]),
                    ndmin=2)
    np.testing.assert_allclose(fx[0], st[2], rtol=1e-12)
    nb = np.fromfile(os.path.join(d, "outBin", "nborsList.bin"), dtype=np.int32).reshape(N, 12)
    ref_nb, _, gm, zn = cpu_ref.lattice("square", 6)
    np.testing.assert_array_equal(nb, ref_nb)
    # SMD over the recorded samples at T = TpVec[1]
    H.run_smd(d, precision=H.BATCHED_F64, strict=False)
    tag = "_T(%.5f)%s.bin" % (T[1], label)
    spec = np.fromfile(os.path.join(d, "outBin", "MC_Bin_SpecData" + tag))
    ntm, Nq = 96, 36
    assert spec.size == 3 * Nq * ntm + ntm
    S = spec[:3 * Nq * ntm].reshape(Nq, 3, ntm)
    qss = spec[3 * Nq * ntm:]
    samples = np.fromfile(os.path.join(d, "outBin", "MC_Bin_SpinField" + tag)).reshape(-1, N, 3)
    m = cpu_ref.Model(ref_nb, zn, H.Couplings(J1=1.0, J2=0.6).as_params())
    ref = np.zeros((3, Nq, ntm))
    qref = np.zeros(ntm)
    w = np.ones(ntm)
    for s0 in samples:
        _, tr, q, _ = m.smd_run(s0, 0.02, ntm, gauss_seidel=False)
        ref += cpu_ref.dssf(tr, gm, 6, w, ntm)
        qref += q
    ref /= len(samples)
    qref /= len(samples)
    np.testing.assert_allclose(S.transpose(1, 0, 2), ref, rtol=1e-7, atol=1e-9 * ref.max())
    np.testing.assert_allclose(qss, qref, rtol=1e-9, atol=1e-12)
    # text outputs (CF/smd_code.cpp:3628-4130): quantum-corrected spectrum
    # along the YGMXG path, w-slices, qSS; checked at the Gamma point against
    # the binary file and the closed-form correction factor
    label_dat = "_T(%.5f)%s.dat" % (T[1], label)
    out = os.path.join(d, "outData")
    cq = np.loadtxt(os.path.join(out, "MC_Results", "SpinSpin_Temporal_Corr" + label_dat))
    np.testing.assert_allclose(cq[:, 0], np.arange(ntm) * 0.02, rtol=1e-8)
    np.testing.assert_allclose(cq[:, 1], qss, rtol=1e-11, atol=1e-14)
    zz = np.loadtxt(os.path.join(out, "RK_Results", "YGMXG_wDepn_DFT_zzSpec" + label_dat))
    npw, l2 = ntm // 2, 3
    assert zz.shape == ((4 * l2 + 1) * npw, 3)
    zz = zz.reshape(4 * l2 + 1, npw, 3)
    tmax = ntm * 0.02
    dwf = 2 * np.pi / tmax
    beta = 1.0 / T[1]
    n = np.arange(npw)
    fc = beta * n * dwf
    qt = np.where(n > 0, fc / (1 - np.exp(-np.where(n > 0, fc, 1.0))), 1.0)
    # path point l2 is Gamma (FFT index 0); the spectrum there is S_z(q = 0)
    np.testing.assert_allclose(zz[l2, :, 1], n * dwf, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(zz[l2, :, 2], tmax / (2 * np.pi) * qt * S[0, 2, :npw], rtol=1e-10)
    slices = [f for f in os.listdir(os.path.join(out, "RK_Results")) if f.startswith("DFT_zzSpec_wfreq(")]
    assert len(slices) == len(range(0, npw, int(np.ceil(0.1 * npw))))
    # MC static spectrum: zone map and path at Gamma = |<S_z(0)>|^2 from MC_Bin_w0SqField
    tag0 = "_T(%.5f)%s" % (T[1], label)
    raw = open(os.path.join(d, "outBin", "MC_Bin_w0SqField" + tag0 + ".bin"), "rb").read()
    cnt = np.frombuffer(raw[:4], np.int32)[0]
    w0 = np.frombuffer(raw[4:], np.complex128).reshape(N, 3) / cnt
    zmap = np.loadtxt(os.path.join(out, "MC_Results", "DFT_zzSpec" + tag0 + ".dat"))
    assert zmap.shape == (7 * 7, 3)
    g0 = zmap[(zmap[:, 0] == 0) & (zmap[:, 1] == 0)]
    assert g0.shape[0] == 1
    assert g0[0, 2] == pytest.approx(max(1e-16, abs(w0[0, 2]) ** 2), rel=1e-5)
    zp = np.loadtxt(os.path.join(out, "MC_Results", "YGMXG_wZero_DFT_zzSpec" + tag0 + ".dat"))
    assert zp.shape == (4 * l2 + 1, 2)
    assert zp[l2, 1] == pytest.approx(max(1e-16, abs(w0[0, 2]) ** 2), rel=1e-11)


def test_inputf_state_errors_before_any_gpu_work(tmp_path):
    # IState 'inputf' reads outBin/inputSConf.bin (N x 3 f64); a missing or
    # unnormalised file stops the run (CF/mc_hm_code.cpp:1978-2047)
    d = str(tmp_path)
    default(d, IState="inputf")
    with pytest.raises(H.HSMDError) as e:
        H.run_mc(d, 4)
    assert e.value.code == 4 and "inputSConf" in str(e.value)
    np.full((36, 3), 0.5).tofile(os.path.join(d, "outBin", "inputSConf.bin"))
    with pytest.raises(H.HSMDError) as e:
        H.run_mc(d, 4)
    assert e.value.code == 4 and "normalised" in str(e.value)


@pytest.mark.gpu
def test_inputf_ferromagnet_stays_ordered_at_low_temperature(tmp_path):
    d = str(tmp_path)
    default(d, IState="inputf", J1=-1.0, J2=0.0, T1=0.01, T2=0.02, NTerm=100, NMeas=2000,
            crSkip=1000, rec=0, ptON=0)
    s = np.zeros((36, 3))
    s[:, 0] = 1.0
    s.tofile(os.path.join(d, "outBin", "inputSConf.bin"))
    H.run_mc(d, 2, precision=H.BATCHED_F64, seed=4)
    T = np.fromfile(os.path.join(d, "outBin", "Temp_List_J2(0.000).bin"))
    for t in T:
        v1 = np.fromfile(os.path.join(d, "outBin", "MC_Bin_AvgValue1_T(%.5f)_J2(0.000).bin" % t))
        v1 = v1.reshape(-1, 23)
        assert np.all(v1[:, 15] > 0.98)      # |M|/N
        assert np.all(v1[:, 0] < -1.95)      # E/N close to -2 J1 ... J1 = -1


@pytest.mark.gpu
def test_quasicrystal_run_writes_values3_and_histograms(tmp_path):
    # a tabulated quasicrystal run (qcrystal = 1, the 239-site fixture with its
    # reference states): Values3 holds the StPrj / SqPks / N1fpar blocks with
    # Binder cumulants consistent with their moments, and the three
    # histograms count every measured sweep, accumulating over runs
    import shutil
    d = str(tmp_path)
    default(d, L=239, geom="Aamman", qc=1, J1=1.0, J2=0.4, JX=0.1, T1=0.5, T2=2.0, NTerm=100,
            NMeas=2000, crSkip=1000, rec=0, dft=0, ptON=1)
    src = os.path.join(os.path.dirname(__file__), "golden", "qc_no_PBC_239")
    for f in os.listdir(src):
        if f.endswith(".bin"):
            shutil.copy(os.path.join(src, f), os.path.join(d, "outBin", f))
    os.makedirs(os.path.join(d, "outData"), exist_ok=True)
    H.run_mc(d, 2, precision=H.BATCHED_F64, seed=5)
    ob = os.path.join(d, "outBin")
    v3f = sorted(f for f in os.listdir(ob) if f.startswith("MC_Bin_AvgValue3"))
    assert len(v3f) == 2
    for f in v3f:
        v3 = np.fromfile(os.path.join(ob, f)).reshape(-1, 45)
        assert v3.shape[0] == 2
        for b in range(2):
            for k in range(4):
                p1, p2, p4, bc = v3[b, 5 * k:5 * k + 4]
                assert 0.0 <= p1 <= 1.0 and p2 >= 0.0
                if p2 > 0:
                    assert bc == pytest.approx(1.0 - p4 / (3.0 * p2 * p2), rel=1e-12, abs=1e-12)
                s1, s2, s4, sb = v3[b, 20 + 5 * k:20 + 5 * k + 4]
                if s2 > 0:
                    assert sb == pytest.approx(2.0 - s4 / (s2 * s2), rel=1e-12, abs=1e-12)
            assert 0.0 < v3[b, 40] <= 1.0          # N1fpar in [0, 1]
        tag = f[len("MC_Bin_AvgValue3"):]
        hn = np.fromfile(os.path.join(ob, "MC_Bin_NeelHist" + tag), dtype=np.int32)
        hs = np.fromfile(os.path.join(ob, "MC_Bin_StrpHist" + tag), dtype=np.int32).reshape(41, 4)
        hv = np.fromfile(os.path.join(ob, "MC_Bin_SMagVecHist" + tag), dtype=np.int32)
        assert hn.sum() == 2000 and np.all(hs.sum(0) == 2000) and hv.sum() == 2000
        assert os.path.exists(os.path.join(d, "outData", "MC_Results", "NeelHist" + tag[:-4] + ".dat"))
    # a second run adds to the histogram files
    H.run_mc(d, 2, precision=H.BATCHED_F64, seed=6)
    tag = v3f[0][len("MC_Bin_AvgValue3"):]
    hn = np.fromfile(os.path.join(ob, "MC_Bin_NeelHist" + tag), dtype=np.int32)
    assert hn.sum() == 4000
