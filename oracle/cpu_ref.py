"""ctypes binding of the CPU FP64 oracle (oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpu_ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_I64 = ctypes.POINTER(ctypes.c_longlong)
_P = ctypes.c_void_p
_I = ctypes.c_int

_SIG = {
    "orc_lattice_dims": (_I, [ctypes.c_char_p, _I] + [ctypes.POINTER(_I)] * 5),
    "orc_lattice_generate": (_I, [ctypes.c_char_p, _I, _I, _I32, _D, _I32]),
    "orc_model_create": (_P, [_I, _I, _I, _I, _I32, _I32, _D, _I]),
    "orc_model_destroy": (None, [_P]),
    "orc_bond_count": (_I, [_P]),
    "orc_bonds": (None, [_P, _I32, _I32]),
    "orc_local_fields": (None, [_P, _I, _D, _D]),
    "orc_local_fields_x": (None, [_P, _D, _D]),
    "orc_energies": (None, [_P, _I, _D, _D]),
    "orc_energies_ising": (None, [_P, _I, _D, _D]),
    "orc_reflect": (None, [_D, _D, _D]),
    "orc_heat_bath": (None, [ctypes.c_double, _D, ctypes.c_double, ctypes.c_double, _D]),
    "orc_philox": (None, [ctypes.c_uint64, _U32, _U32]),
    "orc_det_exp": (ctypes.c_double, [ctypes.c_double]),
    "orc_exchange": (None, [_I, _I, _I, ctypes.c_uint64, ctypes.c_uint32, _D, _D, _I32, _I64, _I64,
                            _I32]),
    "orc_mc_replay": (None, [_P, _I, _I, _D, _I32, _I, _I, _I, _I, _I, ctypes.c_uint64,
                             ctypes.c_uint32, _I, _I, _D, _I32, _I64, _I64, _I32, _D]),
    "orc_mc_replay_hb32": (None, [_P, _I, _D, _I32, _I32, _I, _I, ctypes.c_uint64, ctypes.c_uint32,
                                  _D]),
    "orc_mc_reference": (None, [_P, _I, _I, _D, _I, _I, _I, _I, _I, ctypes.c_uint64, _I, _I,
                                _D, _I32, _D, _I, _I64, _I64, _I32, _I]),
    "orc_smd_run": (None, [_P, _D, ctypes.c_double, _I, _I, _D, _D, _D]),
    "orc_dssf": (None, [_I, _I, _I32, _I, _D, _D, _I, _D]),
    "orc_window": (_I, [ctypes.c_char_p, _I, _D]),
    "orc_spin_stiffness": (None, [_P, _D, _I, _D, _D]),
    "orc_sq_grid": (None, [_I, _I, _I32, _I, _D, _D]),
    "orc_sq_direct": (None, [_I, _I, _I, ctypes.c_double, _D, _I, _D, _D]),
    "orc_corr_len": (None, [_I, _I, ctypes.c_double, _D, _D, _D]),
    "orc_dssf_direct": (None, [_I, _D, _I32, _I, _D, _I, _D, _D, _I, _D]),
    "orc_observables": (None, [ctypes.c_char_p, _I, _I, _I32, _I, _I, _I, _I, _D, _D]),
    "orc_qc_values": (None, [_I, _I32, _I, _I, _I, _D, _D, _D, _I, ctypes.c_double, _I32, _D]),
    "orc_qc_peaks": (None, [_I, _D, _D, _I, ctypes.c_double, _I32]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I32)


def lattice(geom, L, pbc=True):
    """Reference lattice by the reference's full scan: (nbors, rvec, gridmap, zn)."""
    N, z1, z2, zx, G = (_I(), _I(), _I(), _I(), _I())
    rc = lib().orc_lattice_dims(geom.encode(), L, ctypes.byref(N), ctypes.byref(z1),
                                ctypes.byref(z2), ctypes.byref(zx), ctypes.byref(G))
    if rc:
        raise ValueError("unknown geometry")
    zn = (z1.value, z2.value, zx.value)
    nb = np.zeros((N.value, sum(zn)), np.int32)
    rv = np.zeros((N.value, 2))
    gm = np.zeros(G.value * G.value, np.int32)
    rc = lib().orc_lattice_generate(geom.encode(), L, int(pbc), _ip(nb), _dp(rv), _ip(gm))
    if rc:
        raise ValueError("oracle lattice generation failed (%d)" % rc)
    return nb, rv, gm, zn


class Model:
    """Couplings + neighbour table (params as Couplings.as_params order)."""

    def __init__(self, nbors, zn, params, imp=None, lieb=False):
        self.nb = np.ascontiguousarray(nbors, dtype=np.int32)
        self.N = self.nb.shape[0]
        self.zn = zn
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.imp = None if imp is None else np.ascontiguousarray(imp, dtype=np.int32)
        self._h = lib().orc_model_create(self.N, zn[0], zn[1], zn[2], _ip(self.nb),
                                         None if self.imp is None else _ip(self.imp),
                                         _dp(self.params), int(lieb))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_model_destroy(self._h)
            self._h = None

    def bonds(self):
        n = lib().orc_bond_count(self._h)
        p = np.zeros((n, 2), np.int32)
        s = np.zeros(n, np.int32)
        lib().orc_bonds(self._h, _ip(p), _ip(s))
        return p, s

    def local_fields(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros_like(s)
        lib().orc_local_fields(self._h, s.shape[0], _dp(s), _dp(out))
        return out

    def local_fields_x(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.N, 3)
        out = np.zeros(self.N)
        lib().orc_local_fields_x(self._h, _dp(s), _dp(out))
        return out

    def energies(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros(s.shape[0])
        lib().orc_energies(self._h, s.shape[0], _dp(s), _dp(out))
        return out

    def energies_ising(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros(s.shape[0])
        lib().orc_energies_ising(self._h, s.shape[0], _dp(s), _dp(out))
        return out

    def mc_replay(self, model, temps, colour, spins, n_sweeps, n_or=5, n_hb=5, n_or_b=2, n_hb_b=8,
                  cap50=True, seed=1, sweep0=0, pt_parity=-1, state=None):
        """Colour-order replay of the GPU sweep; spins (R, N, 3) are updated in
        place (configuration order).  state = (cos, cnt, exc, boost)."""
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        R = len(temps)
        if state is None:
            state = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
                     np.zeros(R, np.int32))
        cos, cnt, exc, boost = state
        tr = np.zeros((n_sweeps, R))
        col = np.ascontiguousarray(colour, dtype=np.int32)
        lib().orc_mc_replay(self._h, model, R, _dp(temps), _ip(col), n_or, n_hb, n_or_b, n_hb_b,
                            int(cap50), seed, sweep0, n_sweeps, pt_parity, _dp(spins), _ip(cos),
                            cnt.ctypes.data_as(_I64), exc.ctypes.data_as(_I64), _ip(boost),
                            _dp(tr))
        return tr, state

    def mc_replay_hb32(self, temps, order, col_off, spins, n_hb=1, seed=1, sweep=0):
        """Heat-bath passes of the batched FP32 sweep replayed in FP64 (the
        FP32 kernel's formulas and uniforms); spins (R, N, 3), configuration
        c at temps[c], updated in place."""
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        od = np.ascontiguousarray(order, dtype=np.int32)
        co = np.ascontiguousarray(col_off, dtype=np.int32)
        lib().orc_mc_replay_hb32(self._h, len(temps), _dp(temps), _ip(od), _ip(co), len(co) - 1,
                                 int(n_hb), int(seed), int(sweep) & 0xffffffff, _dp(spins))

    def mc_reference(self, model, temps, spins, n_sweeps, n_or=5, n_hb=5, n_or_b=2, n_hb_b=8,
                     cap50=True, seed=1, pt_start=100, threads=0, state=None, sweep0=0):
        """Reference-order chain; spins (R, N, 3) in configuration order are
        updated in place.  state = (cos, cnt, exc, boost) continues a chain
        (updated in place); without it every call starts from cos = identity."""
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        R = len(temps)
        if state is None:
            cos = np.arange(R, dtype=np.int32)
            cnt = exc = boost = None
        else:
            cos, cnt, exc, boost = state
        acc = np.zeros((R, 2))
        lib().orc_mc_reference(self._h, model, R, _dp(temps), n_or, n_hb, n_or_b, n_hb_b,
                               int(cap50), seed, n_sweeps, pt_start, _dp(spins), _ip(cos),
                               _dp(acc), threads,
                               None if cnt is None else cnt.ctypes.data_as(_I64),
                               None if exc is None else exc.ctypes.data_as(_I64),
                               None if boost is None else _ip(boost), int(sweep0))
        return cos, acc

    def spin_stiffness(self, rvec, spins):
        """(R, 6): rhoS1vec[0..2] and the unsquared C-sums of rhoS2vec."""
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        rv = np.ascontiguousarray(rvec, dtype=np.float64)
        out = np.zeros((s.shape[0], 6))
        lib().orc_spin_stiffness(self._h, _dp(rv), s.shape[0], _dp(s), _dp(out))
        return out

    def smd_run(self, spins, dt, n_points, gauss_seidel=True, traj=True):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.N, 3).copy()
        tr = np.zeros((n_points, self.N, 3)) if traj else None
        q = np.zeros(n_points)
        e = np.zeros(n_points)
        lib().orc_smd_run(self._h, _dp(s), dt, n_points, int(gauss_seidel),
                          None if tr is None else _dp(tr), _dp(q), _dp(e))
        return s, tr, q, e


def dssf(traj, gridmap, G, window, n_omega):
    traj = np.ascontiguousarray(traj, dtype=np.float64)
    T, N = traj.shape[0], traj.shape[1]
    gm = np.ascontiguousarray(gridmap, dtype=np.int32)
    w = np.ascontiguousarray(window, dtype=np.float64)
    out = np.zeros((3, G * G, n_omega))
    lib().orc_dssf(N, G, _ip(gm), T, _dp(traj), _dp(w), n_omega, _dp(out))
    return out


def sq_grid(gridmap, G, spins):
    """Static S(q) on the DFT grid: complex (R, 3, G*G)."""
    s = np.ascontiguousarray(spins, dtype=np.float64)
    s = s.reshape(-1, s.shape[-2], 3)
    gm = np.ascontiguousarray(gridmap, dtype=np.int32)
    out = np.zeros((s.shape[0], 3, G * G, 2))
    lib().orc_sq_grid(s.shape[1], G, _ip(gm), s.shape[0], _dp(s), _dp(out))
    return out[..., 0] + 1j * out[..., 1]


def sq_direct(rvec, L, Nq, dq, spins):
    """Static S(q) by direct sums: complex (R, 3, Nq)."""
    s = np.ascontiguousarray(spins, dtype=np.float64)
    s = s.reshape(-1, s.shape[-2], 3)
    rv = np.ascontiguousarray(rvec, dtype=np.float64)
    out = np.zeros((s.shape[0], 3, Nq, 2))
    lib().orc_sq_direct(s.shape[1], L, Nq, dq, _dp(rv), s.shape[0], _dp(s), _dp(out))
    return out[..., 0] + 1j * out[..., 1]


def corr_len(sq, G, c3, dq):
    """Key points of one configuration's S(q) (3, G*G complex):
    (peaks[4][5], clen[4][5]) for xy, yz, zz, tt."""
    z = np.ascontiguousarray(np.stack([np.real(sq), np.imag(sq)], -1), dtype=np.float64)
    pk = np.zeros((4, 5))
    cl = np.zeros((4, 5))
    lib().orc_corr_len(G, int(c3), dq, _dp(z), _dp(pk), _dp(cl))
    return pk, cl


def qc_values(nbors, zn1, spins, ref, rvec, Nq, dq, peaks):
    """Quasicrystal order parameters per configuration (R, 11): NeelPrjt,
    StrpPrjt[4] (after the cut), QctMag, signed N1fpar, SqPeaks[4]."""
    nb = np.ascontiguousarray(nbors, dtype=np.int32)
    s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, nb.shape[0], 3)
    rf = np.ascontiguousarray(ref, dtype=np.float64).reshape(5, nb.shape[0], 3)
    rv = np.ascontiguousarray(rvec, dtype=np.float64)
    pk = np.ascontiguousarray(peaks, dtype=np.int32).reshape(8, 2)
    out = np.zeros((s.shape[0], 11))
    lib().orc_qc_values(nb.shape[0], nb.ctypes.data_as(_I32), nb.shape[1], zn1, s.shape[0], _dp(s),
                        _dp(rf), _dp(rv), Nq, dq, pk.ctypes.data_as(_I32), _dp(out))
    return out


def qc_peaks(state, rvec, Lsz, dq):
    """get_qct_SqPkLoc: the eight hotspots (8, 2) of a reference state."""
    st = np.ascontiguousarray(state, dtype=np.float64)
    rv = np.ascontiguousarray(rvec, dtype=np.float64)
    pk = np.zeros((8, 2), np.int32)
    lib().orc_qc_peaks(st.shape[0], _dp(st), _dp(rv), Lsz, dq, pk.ctypes.data_as(_I32))
    return pk


def dssf_direct(traj, rvec, q, window, n_omega, mask=None):
    """DSSF by direct sums at wavevectors q (nq x 2): (3, nq, n_omega)."""
    traj = np.ascontiguousarray(traj, dtype=np.float64)
    T, N = traj.shape[0], traj.shape[1]
    rv = np.ascontiguousarray(rvec, dtype=np.float64)
    qv = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 2)
    w = np.ascontiguousarray(window, dtype=np.float64)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    out = np.zeros((3, qv.shape[0], n_omega))
    lib().orc_dssf_direct(N, _dp(rv), None if mk is None else _ip(mk), qv.shape[0], _dp(qv), T,
                          _dp(traj), _dp(w), n_omega, _dp(out))
    return out


def dssf_direct_fft(traj, rvec, q, window, n_omega, mask=None):
    """orc_dssf_direct with the time transform by FFT and the spatial sums as
    one matrix product (numpy): the same quantity for trajectories and
    q-paths too large for the direct loops.  traj: T x N x 3."""
    traj = np.asarray(traj, dtype=np.float64)
    T, N = traj.shape[0], traj.shape[1]
    rv = np.asarray(rvec, dtype=np.float64)
    qv = np.asarray(q, dtype=np.float64).reshape(-1, 2)
    Y = np.fft.fft(np.asarray(window)[:, None, None] * traj, axis=0)[:n_omega]   # (n, N, 3)
    ph = np.exp(-1j * (qv @ rv.T))                                                # (nq, N)
    if mask is not None:
        ph[:, np.asarray(mask) != 0] = 0.0
    X = np.stack([ph @ Y[:, :, d].T for d in range(3)]) / math.sqrt(N * T)   # (3, nq, n)
    return np.abs(X) ** 2


def window(name, n):
    w = np.zeros(n)
    if lib().orc_window(name.encode(), n, _dp(w)):
        raise ValueError("unknown window " + name)
    return w


def philox(seed, ctr):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox(seed, c.ctypes.data_as(_U32), o.ctypes.data_as(_U32))
    return o


def det_exp(x):
    return lib().orc_det_exp(float(x))


def exchange(E_cfg, beta, parity, cap50=True, seed=1, sweep=0, state=None):
    """One replica-exchange round; state = (cos, cnt, exc, boost), updated in
    place and returned."""
    E = np.ascontiguousarray(E_cfg, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    R = len(b)
    if state is None:
        state = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
                 np.zeros(R, np.int32))
    cos, cnt, exc, boost = state
    lib().orc_exchange(R, int(parity), int(cap50), int(seed), int(sweep) & 0xffffffff, _dp(E), _dp(b),
                       _ip(cos), cnt.ctypes.data_as(_I64), exc.ctypes.data_as(_I64), _ip(boost))
    return state


def dssf_fft(traj, gridmap, G, window, n_omega):
    """orc_dssf by FFTs (numpy): the same transform, O(T log T) per q, for
    trajectories too long for orc_dssf's direct time sums.  traj: T x N x 3."""
    traj = np.asarray(traj, dtype=np.float64)
    T = traj.shape[0]
    gm = np.asarray(gridmap)
    grid = np.zeros((T, 3, G * G))
    ok = gm >= 0
    grid[:, :, ok] = traj[:, gm[ok], :].transpose(0, 2, 1)
    sq = np.fft.fft2(grid.reshape(T, 3, G, G)).reshape(T, 3, G * G) / G
    x = np.fft.fft(np.asarray(window)[:, None, None] * sq, axis=0)[:n_omega] / math.sqrt(T)
    return np.abs(x.transpose(1, 2, 0)) ** 2


def reflect(s, h):
    s = np.ascontiguousarray(s, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    o = np.zeros(3)
    lib().orc_reflect(_dp(s), _dp(h), _dp(o))
    return o


def heat_bath(beta, h, u1, u2):
    h = np.ascontiguousarray(h, dtype=np.float64)
    o = np.zeros(3)
    lib().orc_heat_bath(beta, _dp(h), u1, u2, _dp(o))
    return o


def observables(geom, L, nbors, zn, h_on, spins):
    """[|M|/N, Mz/N, stagg, stripe H, V, D, SMag1, C3-Psi or Z2] of one configuration."""
    nb = np.ascontiguousarray(nbors, dtype=np.int32)
    s = np.ascontiguousarray(spins, dtype=np.float64)
    out = np.zeros(8)
    lib().orc_observables(geom.encode(), L, nb.shape[0], _ip(nb), zn[0], zn[1], zn[2], int(h_on),
                          _dp(s), _dp(out))
    return out
