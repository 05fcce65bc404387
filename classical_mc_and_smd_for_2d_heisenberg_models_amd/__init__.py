"""MI355X-native Monte Carlo sweep engine and spin-dynamics (SMD) engine for
2D classical Heisenberg / Ising models.

Python host mirror of the C ABI in include/hsmd.h (libhsmd.so, built in-tree
by ``build.py``).  The names follow the reference's functions:

  get_localField / get_energyValue   -> Lattice.local_fields / Lattice.energies
  MC sweep + replica exchange        -> MonteCarlo.run
  get_dynSpectrum                    -> SpinDynamics.step / SpinDynamics.dssf

All arrays crossing the boundary are float64 (x, y, z) triples in site order,
the reference's binary sample layout.  There is no CPU fallback: loading fails
loudly when the HIP library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhsmd.so")

PLAIN_F64, BATCHED_F32, BATCHED_F64 = 0, 1, 2
HEISENBERG, ISING = 0, 1
N_MOMENTS = 93
QC_HIST = 1886
MOMENT_NAMES = ["E", "E2", "F1", "F2", "F4", "T1", "T2", "Q1", "Q2", "Q4",
                "S1", "S2", "S4", "SH", "SV", "SD", "X1", "X2", "X4"]


class HSMDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hsmd error %d: %s" % (code, msg))
        self.code = code


class Couplings(ctypes.Structure):
    """Settings_MODEL / Settings_Disorder couplings (Set_Params.py:6-72)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("J1", "J2", "JX", "J2_delta", "lambda1", "lambda2", "h", "imp_ratio",
                 "J1ic", "J2ic", "JXic", "J1ii", "J2ii", "JXii")]

    def __init__(self, J1=1.0, J2=0.0, JX=0.0, J2_delta=0.0, lambda1=1.0, lambda2=1.0, h=0.0,
                 imp_ratio=0.0, J1ic=0.0, J2ic=0.0, JXic=0.0, J1ii=0.0, J2ii=0.0, JXii=0.0):
        super().__init__(J1, J2, JX, J2_delta, lambda1, lambda2, h, imp_ratio,
                         J1ic, J2ic, JXic, J1ii, J2ii, JXii)

    def as_params(self):
        return np.array([getattr(self, f[0]) for f in self._fields_], dtype=np.float64)


class MCSpec(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("precision", ctypes.c_int),
                ("n_or", ctypes.c_int), ("n_hb", ctypes.c_int),
                ("n_or_boost", ctypes.c_int), ("n_hb_boost", ctypes.c_int),
                ("pt_cap50", ctypes.c_int), ("seed", ctypes.c_uint64)]


class SMDSpec(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("dt", ctypes.c_double),
                ("n_steps", ctypes.c_int), ("n_omega", ctypes.c_int),
                ("dssf", ctypes.c_int), ("gauss_seidel", ctypes.c_int)]


class MCTuning(ctypes.Structure):
    _fields_ = [("colour_cap", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("ms_per_sweep", ctypes.c_double), ("default_ms", ctypes.c_double),
                ("n_candidates", ctypes.c_int32), ("cached", ctypes.c_int32)]


class SMDTuning(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int32), ("gemm", ctypes.c_int32),
                ("ms_per_step", ctypes.c_double), ("default_ms_per_step", ctypes.c_double),
                ("gemm_ms", ctypes.c_double), ("default_gemm_ms", ctypes.c_double),
                ("n_candidates", ctypes.c_int32), ("cached", ctypes.c_int32)]


def _as_dict(st):
    return {k: getattr(st, k) for k, _ in st._fields_}


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U32 = ctypes.POINTER(ctypes.c_uint32)

# name -> (restype, argtypes); the full exported surface of include/hsmd.h
SIGNATURES = {
    "hsmd_last_error": (_I, [ctypes.c_char_p, ctypes.c_size_t]),
    "hsmd_device_count": (_I, []),
    "hsmd_lattice_generate": (_I, [ctypes.c_char_p, _I, _I, _I, ctypes.POINTER(Couplings),
                                   ctypes.POINTER(_P)]),
    "hsmd_lattice_from_tables": (_I, [ctypes.c_int32, _I, _I, _I, _I32, _D,
                                      ctypes.POINTER(Couplings), ctypes.POINTER(_P)]),
    "hsmd_lattice_set_impurities": (_I, [_P, _I32]),
    "hsmd_lattice_set_vacancies": (_I, [_P, _I32]),
    "hsmd_lattice_info": (_I, [_P, _I32]),
    "hsmd_lattice_tables": (_I, [_P, _I32, _D]),
    "hsmd_lattice_grid_map": (_I, [_P, _I32]),
    "hsmd_lattice_bonds": (_I, [_P, _I32, _I32]),
    "hsmd_lattice_colours": (_I, [_P, _I32]),
    "hsmd_lattice_set_colour_cap": (_I, [_P, ctypes.c_int32]),
    "hsmd_lattice_set_qc_states": (_I, [_P, _D, _D, _I, _I32]),
    "hsmd_mc_qc_histograms": (_I, [_P, _I32, _I]),
    "hsmd_mc_set_threads": (_I, [_P, _I]),
    "hsmd_mc_sweep_kernel": (_I, [_P, ctypes.c_char_p, _I]),
    "hsmd_mc_autotune": (_I, [_P, _I, _D, ctypes.POINTER(MCSpec), _I, ctypes.POINTER(MCTuning)]),
    "hsmd_smd_set_shape": (_I, [_P, _I, _I]),
    "hsmd_smd_step_kernel": (_I, [_P, ctypes.c_char_p, _I]),
    "hsmd_smd_set_samples_from_mc": (_I, [_P, _P, _I32]),
    "hsmd_smd_autotune": (_I, [ctypes.POINTER(_P), _I, _I32, _I, ctypes.POINTER(SMDSpec), _D, _I,
                               ctypes.POINTER(SMDTuning)]),
    "hsmd_lattice_destroy": (None, [_P]),
    "hsmd_local_fields": (_I, [_P, _I, _I, _D, _D]),
    "hsmd_energies": (_I, [_P, _I, _I, _D, _D]),
    "hsmd_mc_create": (_I, [_P, _I, _D, ctypes.POINTER(MCSpec), ctypes.POINTER(_P)]),
    "hsmd_mc_set_spins": (_I, [_P, _D]),
    "hsmd_mc_get_spins": (_I, [_P, _D]),
    "hsmd_mc_run": (_I, [_P, _I, _I, _I, _D]),
    "hsmd_mc_energies": (_I, [_P, _D]),
    "hsmd_mc_config_energies_noalloc": (_I, [_P, _I, _P, _P, _P]),
    "hsmd_mc_pt_counters": (_I, [_P, _I64, _I64]),
    "hsmd_mc_reset_counters": (_I, [_P]),
    "hsmd_mc_moments": (_I, [_P, _D, _I64, _I]),
    "hsmd_mc_slot_map": (_I, [_P, _I32]),
    "hsmd_mc_set_sq": (_I, [_P, _I]),
    "hsmd_mc_sq_field": (_I, [_P, _D, _I64, _I]),
    "hsmd_static_sq": (_I, [_P, _I, _D, _D]),
    "hsmd_poisson_disk": (_I, [ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _I, _D, _I32]),
    "hsmd_lattice_from_points": (_I, [_I, _D, ctypes.c_double, _D, _P, _P]),
    "hsmd_smd_dssf_q": (_I, [_P, _I, _I, _D, _D]),
    "hsmd_mc_destroy": (None, [_P]),
    "hsmd_smd_create": (_I, [ctypes.POINTER(_P), _I, _I32, _I, ctypes.POINTER(SMDSpec), _D,
                             ctypes.POINTER(_P)]),
    "hsmd_smd_set_samples": (_I, [_P, _D]),
    "hsmd_smd_get_samples": (_I, [_P, _D]),
    "hsmd_smd_step": (_I, [_P, _I]),
    "hsmd_smd_feed": (_I, [_P, _I, _D]),
    "hsmd_smd_dssf": (_I, [_P, _I, _D, _D]),
    "hsmd_smd_energies": (_I, [_P, _D]),
    "hsmd_smd_destroy": (None, [_P]),
    "hsmd_run_mc": (_I, [ctypes.c_char_p, _I, _I, ctypes.c_uint64, _I]),
    "hsmd_run_smd": (_I, [ctypes.c_char_p, _I, _I]),
    "hsmd_smd_window": (_I, [ctypes.c_char_p, _I, _D]),
    "hsmd_debug_philox": (_I, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, _I, _U32]),
    "hsmd_debug_exp": (_I, [_I, _D, _D]),
    "hsmd_debug_mc_sweeps": (_I, [_P, _I]),
    "hsmd_bin_statistics": (_I, [_D, _I, _I, _I, ctypes.c_uint32, _D, _D]),
    "hsmd_debug_mt19937": (_I, [ctypes.c_uint32, _I, _U32]),
}

_lib = None


def lib():
    """Loads libhsmd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libhsmd.so not built: run python -m "
                              "classical_mc_and_smd_for_2d_heisenberg_models_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().hsmd_last_error(buf, 1024)
        raise HSMDError(code, buf.value.decode())


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I32)


def device_count():
    return lib().hsmd_device_count()


class Lattice:
    """Lattice + couplings (make_rvecList / make_nborsTable / make_bondList)."""

    def __init__(self, handle):
        self._h = _P(handle) if not isinstance(handle, _P) else handle
        info = np.zeros(16, np.int32)
        _check(lib().hsmd_lattice_info(self._h, _ip(info)))
        (self.N, self.zn1, self.zn2, self.znx, self.n_bonds, self.n_colours, self.Gsz,
         self.Nq, self.nsubs, self.L) = [int(v) for v in info[:10]]
        self.flags = dict(zip(["j2_on", "jx_on", "xxz_on", "h_on", "disorder", "c3sym"],
                              [bool(v) for v in info[10:16]]))

    @classmethod
    def generate(cls, geometry, L, couplings=None, pbc=True, strict=False):
        c = couplings or Couplings()
        h = _P()
        _check(lib().hsmd_lattice_generate(geometry.encode(), int(L), int(pbc), int(strict),
                                            ctypes.byref(c), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_tables(cls, nbors, zn1, zn2, znx, rvec=None, couplings=None):
        nbors = np.ascontiguousarray(nbors, dtype=np.int32)
        n = nbors.size // (zn1 + zn2 + znx)
        c = couplings or Couplings()
        rv = None if rvec is None else np.ascontiguousarray(rvec, dtype=np.float64)
        h = _P()
        _check(lib().hsmd_lattice_from_tables(n, zn1, zn2, znx, _ip(nbors),
                                               None if rv is None else _dp(rv),
                                               ctypes.byref(c), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_points(cls, points, box, r_shell=(1.15, 1.45, 1.75), couplings=None):
        """Tabulated lattice of points in a periodic box, shells by distance."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
        r = np.ascontiguousarray(r_shell, dtype=np.float64)
        c = couplings or Couplings()
        h = _P()
        _check(lib().hsmd_lattice_from_points(pts.shape[0], _dp(pts), float(box), _dp(r),
                                               ctypes.byref(c), ctypes.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hsmd_lattice_destroy(self._h)
            self._h = None

    def _refresh(self):
        self.__init__(self._h)

    def set_impurities(self, imp):
        imp = np.ascontiguousarray(imp, dtype=np.int32)
        _check(lib().hsmd_lattice_set_impurities(self._h, _ip(imp)))
        self._refresh()

    def set_qc_states(self, neel, stripes, neel_phase=False):
        """Quasicrystal reference states (Neel n x 3, stripes 4 x n x 3) for
        the projection order parameters; returns the eight S(q) hotspots
        (nx, ny) in units of dq."""
        nl = np.ascontiguousarray(neel, dtype=np.float64).reshape(self.N, 3)
        st = np.ascontiguousarray(stripes, dtype=np.float64).reshape(4, self.N, 3)
        pk = np.zeros((8, 2), np.int32)
        _check(lib().hsmd_lattice_set_qc_states(self._h, _dp(nl), _dp(st), int(neel_phase), _ip(pk)))
        return pk

    def set_colour_cap(self, cap):
        """Colour schedule of the GPU sweeps: 0 = default, > 0 = first-fit
        with at most cap sites per colour."""
        _check(lib().hsmd_lattice_set_colour_cap(self._h, int(cap)))
        self._refresh()

    def set_vacancies(self, vac):
        vac = np.ascontiguousarray(vac, dtype=np.int32)
        _check(lib().hsmd_lattice_set_vacancies(self._h, _ip(vac)))
        self._refresh()

    def tables(self):
        nb = np.zeros((self.N, self.zn1 + self.zn2 + self.znx), np.int32)
        _check(lib().hsmd_lattice_tables(self._h, _ip(nb), None))
        return nb

    def sites(self):
        rv = np.zeros((self.N, 2), np.float64)
        _check(lib().hsmd_lattice_tables(self._h, None, _dp(rv)))
        return rv

    def positions(self):
        """Site coordinates rvecList (N x 2)."""
        rv = np.zeros((self.N, 2))
        _check(lib().hsmd_lattice_tables(self._h, None, _dp(rv)))
        return rv

    def grid_map(self):
        g = np.zeros(self.Nq, np.int32)
        _check(lib().hsmd_lattice_grid_map(self._h, _ip(g)))
        return g

    def bonds(self):
        p = np.zeros((self.n_bonds, 2), np.int32)
        s = np.zeros(self.n_bonds, np.int32)
        _check(lib().hsmd_lattice_bonds(self._h, _ip(p), _ip(s)))
        return p, s

    def colours(self):
        c = np.zeros(self.N, np.int32)
        _check(lib().hsmd_lattice_colours(self._h, _ip(c)))
        return c

    def local_fields(self, spins, precision=BATCHED_F64):
        """get_localField at every site of R configurations (R x N x 3)."""
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros_like(s)
        _check(lib().hsmd_local_fields(self._h, precision, s.shape[0], _dp(s), _dp(out)))
        return out

    def energies(self, spins, precision=BATCHED_F64):
        """get_energyValue of R configurations."""
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros(s.shape[0])
        _check(lib().hsmd_energies(self._h, precision, s.shape[0], _dp(s), _dp(out)))
        return out

    def static_sq(self, spins):
        """get_SqData_iMKL / get_qct_SqData of R configurations: complex
        (R, 3, Nq) in the reference's q order."""
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(-1, self.N, 3)
        out = np.zeros((s.shape[0], 3, self.Nq, 2))
        _check(lib().hsmd_static_sq(self._h, s.shape[0], _dp(s), _dp(out)))
        return out[..., 0] + 1j * out[..., 1]

    # reference-named aliases
    get_localField = local_fields
    get_energyValue = energies


class MonteCarlo:
    """All parallel-tempering replicas of one lattice on the GPU."""

    def __init__(self, lattice, temps, model=HEISENBERG, precision=BATCHED_F32, n_or=5, n_hb=5,
                 n_or_boost=2, n_hb_boost=8, pt_cap50=True, seed=1, tune=False, tune_sweeps=10):
        """tune=True runs the sweep autotuner first (colour schedule and
        workgroup shape, timed on the device, cached through HSMD_TUNE_CACHE)
        and applies its choice: the lattice's colour cap changes."""
        self.lattice = lattice
        self.temps = np.ascontiguousarray(temps, dtype=np.float64)
        self.R = len(self.temps)
        self.N = lattice.N
        self.spec = MCSpec(model, precision, n_or, n_hb, n_or_boost, n_hb_boost,
                           int(pt_cap50), seed)
        self.tuning = None
        if tune:
            t = MCTuning()
            _check(lib().hsmd_mc_autotune(lattice._h, self.R, _dp(self.temps),
                                          ctypes.byref(self.spec), int(tune_sweeps), ctypes.byref(t)))
            self.tuning = _as_dict(t)
            lattice.set_colour_cap(t.colour_cap)
        h = _P()
        _check(lib().hsmd_mc_create(lattice._h, self.R, _dp(self.temps), ctypes.byref(self.spec),
                                    ctypes.byref(h)))
        self._h = h
        if self.tuning is not None and precision != PLAIN_F64:
            self.set_threads(self.tuning["threads"])

    def set_threads(self, threads):
        """Threads per replica workgroup of the batched forms."""
        _check(lib().hsmd_mc_set_threads(self._h, int(threads)))

    def sweep_kernel(self):
        """Name of the sweep kernel the engine launches in its current form
        (ends in ' wave-ordered' when the fast kernel orders its colour
        classes per wave instead of by a workgroup barrier)."""
        buf = ctypes.create_string_buffer(256)
        _check(lib().hsmd_mc_sweep_kernel(self._h, buf, 256))
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hsmd_mc_destroy(self._h)
            self._h = None

    def set_spins(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.R, self.N, 3)
        _check(lib().hsmd_mc_set_spins(self._h, _dp(s)))

    def get_spins(self):
        s = np.zeros((self.R, self.N, 3))
        _check(lib().hsmd_mc_get_spins(self._h, _dp(s)))
        return s

    def run(self, n_sweeps, pt_parity=-1, measure=False, trace=False):
        tr = np.zeros((n_sweeps, self.R)) if trace else None
        _check(lib().hsmd_mc_run(self._h, int(n_sweeps), int(pt_parity), int(measure),
                                 None if tr is None else _dp(tr)))
        return tr

    def sweeps_only(self, n_sweeps):
        """n colour sweeps of every replica, no measurement or exchange
        (times the sweep kernel alone)."""
        _check(lib().hsmd_debug_mc_sweeps(self._h, int(n_sweeps)))

    def qc_histograms(self, reset=False):
        """Per-slot quasicrystal histograms: (neel [R, 41], stripes [R, 41, 4],
        stripe vector [R, 41 (imag), 41 (real)])."""
        h = np.zeros((self.R, QC_HIST), np.int32)
        _check(lib().hsmd_mc_qc_histograms(self._h, _ip(h), int(reset)))
        return h[:, :41], h[:, 41:205].reshape(self.R, 41, 4), h[:, 205:].reshape(self.R, 41, 41)

    def energies(self):
        e = np.zeros(self.R)
        _check(lib().hsmd_mc_energies(self._h, _dp(e)))
        return e

    def config_energies_noalloc(self, n_configs, spins_dev, energies_dev, stream=None):
        """get_energyValue of n_configs configurations in device memory
        (device addresses as ints), written to device memory; allocates
        nothing and does not wait for the stream."""
        _check(lib().hsmd_mc_config_energies_noalloc(self._h, int(n_configs), _P(spins_dev),
                                                     _P(energies_dev), _P(stream)))

    def pt_counters(self):
        a = np.zeros(self.R, np.int64)
        b = np.zeros(self.R, np.int64)
        _check(lib().hsmd_mc_pt_counters(self._h, a.ctypes.data_as(_I64), b.ctypes.data_as(_I64)))
        return a, b

    def reset_counters(self):
        _check(lib().hsmd_mc_reset_counters(self._h))

    def moments(self, reset=False):
        out = np.zeros((self.R, N_MOMENTS))
        n = np.zeros(1, np.int64)
        _check(lib().hsmd_mc_moments(self._h, _dp(out), n.ctypes.data_as(_I64), int(reset)))
        return out, int(n[0])

    def slot_map(self):
        m = np.zeros(self.R, np.int32)
        _check(lib().hsmd_mc_slot_map(self._h, _ip(m)))
        return m

    def set_sq(self, every):
        """Static S(q) measurement every `every` measured sweeps (0 = off)."""
        _check(lib().hsmd_mc_set_sq(self._h, int(every)))

    def sq_field(self, reset=False):
        """(R, Nq, 6) sums of |S_x|, |S_y|, |S_z|, arg S_x, arg S_y, arg S_z and
        the number of S(q) measurements."""
        out = np.zeros((self.R, self.lattice.Nq, 6))
        n = np.zeros(1, np.int64)
        _check(lib().hsmd_mc_sq_field(self._h, _dp(out), n.ctypes.data_as(_I64), int(reset)))
        return out, int(n[0])


class SpinDynamics:
    """All SMD samples (any mix of lattices) on the GPU."""

    def __init__(self, lattices, sample_lattice, dt, n_steps, n_omega=None, precision=BATCHED_F32,
                 dssf=True, window=None, tune=False, tune_steps=64):
        """tune=True runs the step autotuner first (workgroup form and DSSF
        contraction tile, timed on the device on a scratch engine, cached
        through HSMD_TUNE_CACHE) and applies its choice."""
        self.lattices = list(lattices)
        self.sample_lattice = np.ascontiguousarray(sample_lattice, dtype=np.int32)
        self.S = len(self.sample_lattice)
        self.n_steps = int(n_steps)
        self.n_omega = int(n_omega if n_omega is not None else n_steps)
        self.spec = SMDSpec(precision, dt, self.n_steps, self.n_omega, int(dssf), 0)
        self.window = None if window is None else np.ascontiguousarray(window, dtype=np.float64)
        arr = (_P * len(self.lattices))(*[l._h for l in self.lattices])
        self.tuning = None
        if tune:
            t = SMDTuning()
            _check(lib().hsmd_smd_autotune(arr, len(self.lattices), _ip(self.sample_lattice), self.S,
                                           ctypes.byref(self.spec),
                                           None if self.window is None else _dp(self.window),
                                           int(tune_steps), ctypes.byref(t)))
            self.tuning = _as_dict(t)
        h = _P()
        _check(lib().hsmd_smd_create(arr, len(self.lattices), _ip(self.sample_lattice), self.S,
                                     ctypes.byref(self.spec),
                                     None if self.window is None else _dp(self.window),
                                     ctypes.byref(h)))
        self._h = h
        if self.tuning is not None:
            self.set_shape(self.tuning["form"], self.tuning["gemm"] if dssf else 1)
        self.sizes = [self.lattices[i].N for i in self.sample_lattice]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hsmd_smd_destroy(self._h)
            self._h = None

    def set_shape(self, form, gemm):
        """Launch shape of the batched forms: form 1 (LDS, one slot class),
        2 (LDS, per-workgroup classes), 3 (global memory); gemm 0 (128 x 128
        contraction tile), 1 (64 x 64), 2 (128 x 256), 3 (128 x 128, two
        LDS stages), 4 (128 x 512, 2 n_omega a multiple of 512), 5 (the
        128 x 512 tile pipelined, same sums as 4) or 6 (the time DFT as a
        pruned FFT: FP32, n_steps a multiple of 512, n_omega <= 512)."""
        _check(lib().hsmd_smd_set_shape(self._h, int(form), int(gemm)))

    def step_kernel(self):
        """Name of the RK4 kernel step() launches in the current form (ends
        in ', wave-ordered on k of n lattices' when the LDS-resident form
        orders its stages per wave instead of by workgroup barriers)."""
        buf = ctypes.create_string_buffer(256)
        _check(lib().hsmd_smd_step_kernel(self._h, buf, 256))
        return buf.value.decode()

    def set_samples_from_mc(self, mc, slots):
        """Sample s starts from the configuration at temperature slot
        slots[s] of the MonteCarlo engine mc, copied on the device."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        if sl.shape != (self.S,):
            raise ValueError("one slot per sample")
        _check(lib().hsmd_smd_set_samples_from_mc(self._h, mc._h, _ip(sl)))

    def set_samples(self, spins_list):
        flat = np.concatenate([np.asarray(s, dtype=np.float64).reshape(-1) for s in spins_list])
        _check(lib().hsmd_smd_set_samples(self._h, _dp(flat)))

    def get_samples(self):
        flat = np.zeros(3 * sum(self.sizes))
        _check(lib().hsmd_smd_get_samples(self._h, _dp(flat)))
        out, o = [], 0
        for n in self.sizes:
            out.append(flat[o:o + 3 * n].reshape(n, 3))
            o += 3 * n
        return out

    def step(self, n):
        _check(lib().hsmd_smd_step(self._h, int(n)))

    def feed(self, points):
        """Appends time points instead of integrating them: points is a
        sequence of per-time-point lists of sample configurations (or an
        array (n, sum N_s, 3)); the DSSF then covers the fed trajectory."""
        if isinstance(points, np.ndarray):
            flat = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
            n = points.shape[0]
        else:
            flat = np.concatenate([np.asarray(s, dtype=np.float64).reshape(-1)
                                   for pt in points for s in pt])
            n = len(points)
        if flat.size != n * 3 * sum(self.sizes):
            raise ValueError("each time point holds every sample's configuration")
        _check(lib().hsmd_smd_feed(self._h, int(n), _dp(flat)))

    def dssf(self, li=0):
        L = self.lattices[li]
        out = np.zeros((3, L.Nq, self.n_omega))
        q = np.zeros(self.n_steps)
        _check(lib().hsmd_smd_dssf(self._h, li, _dp(out), _dp(q)))
        return out, q

    def dssf_q(self, q, li=0):
        """Sample-summed DSSF of lattice li at wavevectors q (nq x 2) by direct
        spatial sums: (3, nq, n_omega)."""
        qv = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((3, qv.shape[0], self.n_omega))
        _check(lib().hsmd_smd_dssf_q(self._h, li, qv.shape[0], _dp(qv), _dp(out)))
        return out

    def energies(self):
        e = np.zeros(self.S)
        _check(lib().hsmd_smd_energies(self._h, _dp(e)))
        return e


def get_dynSpectrum(lattice, samples, dt, n_steps, window=None, n_omega=None,
                    precision=BATCHED_F64):
    """Sample-averaged DSSF and spin autocorrelation of a list of samples on
    one lattice (the reference's get_dynSpectrum summed over its RK loop,
    CF/smd_code.cpp:3079-3396, divided by the sample count)."""
    S = len(samples)
    smd = SpinDynamics([lattice], np.zeros(S, np.int32), dt, n_steps, n_omega, precision,
                       True, window)
    smd.set_samples(samples)
    smd.step(n_steps - 1)
    d, q = smd.dssf(0)
    return d / S, q / S


def poisson_disk(box, r_min=1.0, seed=5, max_points=None):
    """Seeded Poisson-disk points (n x 2) in the periodic box [0, box)^2."""
    cap = int(max_points or (box / r_min) ** 2 * 1.2 + 16)
    out = np.zeros((cap, 2))
    n = np.zeros(1, np.int32)
    _check(lib().hsmd_poisson_disk(float(box), float(r_min), int(seed), cap, _dp(out), _ip(n)))
    return out[:int(n[0])].copy()


def run_mc(workdir, n_replicas, precision=BATCHED_F32, seed=1, strict=True):
    """The reference's MC program (xMC_run) on the Settings_*.txt files of workdir."""
    _check(lib().hsmd_run_mc(os.fsencode(workdir), int(n_replicas), int(precision), int(seed),
                             int(strict)))


def run_smd(workdir, precision=BATCHED_F32, strict=True):
    """The reference's SMD program (xSMD_run) on the Settings_*.txt files of workdir."""
    _check(lib().hsmd_run_smd(os.fsencode(workdir), int(precision), int(strict)))


def temporal_window(name, n):
    """The SMD run's temporal window tWForm on n points (CF/smd_code.cpp:2808-2895)."""
    out = np.zeros(int(n))
    _check(lib().hsmd_smd_window(name.encode(), int(n), _dp(out)))
    return out


def debug_philox(seed, step, inst, stream, n):
    out = np.zeros(4 * n, np.uint32)
    _check(lib().hsmd_debug_philox(seed, step, inst, stream, n, out.ctypes.data_as(_U32)))
    return out.reshape(n, 4)


def bin_statistics(data, bootstrap=True, seed=0):
    """Mean and deviation over bins (data: n_bins x n_values), as the
    reference's MC_Stat_Results (get_statVectors)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    nb, nv = d.shape
    m = np.zeros(nv)
    s = np.zeros(nv)
    _check(lib().hsmd_bin_statistics(_dp(d), nb, nv, int(bootstrap), int(seed) & 0xffffffff, _dp(m), _dp(s)))
    return m, s


def debug_mt19937(seed, n):
    out = np.zeros(n, np.uint32)
    _check(lib().hsmd_debug_mt19937(int(seed) & 0xffffffff, n, out.ctypes.data_as(_U32)))
    return out


def debug_exp(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    _check(lib().hsmd_debug_exp(x.size, _dp(x), _dp(y)))
    return y
