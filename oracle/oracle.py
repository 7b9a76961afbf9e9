"""ctypes bindings to the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement used as the checker by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.  The product
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

F3 = C.c_float * 3


class OrParams(C.Structure):
    _fields_ = [
        ("kspring_cov", C.c_float), ("R_limit", C.c_float),
        ("R_limit_bond", C.c_float),
        ("pairs_cutoff", C.c_float), ("el", C.c_float), ("a", C.c_float),
        ("pairlist_cutoff", C.c_float), ("pairs_threshold", C.c_float),
        ("pairs_freq", C.c_int), ("possiblepairs_freq", C.c_int),
        ("max_covalent", C.c_int), ("max_native", C.c_int), ("max_pairs", C.c_int), ("max_possiblePairs", C.c_int),
        ("timestep", C.c_float), ("zeta", C.c_float), ("temperature", C.c_float),
        ("minimization", C.c_int),
        ("seed", C.c_int),
        ("heating", C.c_int), ("tempFreq", C.c_int),
        ("initialT", C.c_float), ("deltaT", C.c_float),
        ("nav", C.c_int), ("outputtiming", C.c_int), ("dcdfreq", C.c_int),
        ("pulling", C.c_int),
        ("deltax", C.c_float), ("k_trans", C.c_float), ("fconst", C.c_float),
        ("pullFreq", C.c_int),
        ("pullDirectionVector", C.c_int),
        ("pullVector", F3),
        ("fixedEnd", C.c_int), ("pulledEnd", C.c_int),
        ("indentation", C.c_int),
        ("indChip", F3), ("indTip", F3), ("indDirection", F3), ("indSurfR0", F3), ("indSurfN", F3),
        ("indTipR", C.c_float), ("indKs", C.c_float), ("indDeltaX", C.c_float),
        ("indMoveSurface", C.c_int), ("indFixTrans", C.c_int),
        ("indTipSigma", C.c_float), ("indSurfSigma", C.c_float), ("indTipEl", C.c_float), ("indSurfEl", C.c_float),
        ("indTipA", C.c_float), ("indTipB", C.c_float), ("indSurfA", C.c_float), ("indSurfB", C.c_float),
        ("indTipZeta", C.c_float),
        ("indOutputFreq", C.c_int),
        ("indRetractionStep", C.c_longlong),
        ("hi_on", C.c_int), ("hi_exact", C.c_int), ("hi_a", C.c_float),
        ("hi_epsilon_freq", C.c_int), ("hi_capricious", C.c_int), ("hi_unlisted", C.c_int),
        ("hi_epsmax", C.c_float), ("hi_chol_freq", C.c_int), ("hi_faithful_blocks", C.c_int),
        ("updater_freq_min", C.c_int), ("host_forces", C.c_int),
    ]


def available() -> bool:
    return os.path.exists(LIB_PATH)


def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError("oracle/liboracle.so not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        VP = C.c_void_p
        def sig(name, res, *args):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = list(args)
        sig("or_create", VP, C.POINTER(OrParams), C.c_int, C.c_int, VP, C.c_int, VP, VP, VP, C.c_int, VP, VP, VP, VP,
            C.c_int, VP, VP, C.c_int, VP, C.c_int, VP)
        sig("or_destroy", None, VP)
        sig("or_error", C.c_char_p)
        sig("or_n", C.c_int, VP)
        sig("or_get_coords", None, VP, VP)
        sig("or_set_coords", None, VP, VP)
        sig("or_get_T", None, VP, VP)
        sig("or_get_rng", None, VP, VP)
        sig("or_get_ext", None, VP, VP)
        sig("or_step", C.c_longlong, VP)
        sig("or_get_list", C.c_int, VP, C.c_int, VP, VP)
        sig("or_list_cap", C.c_int, VP, C.c_int)
        sig("or_overflow", C.c_int, VP)
        sig("or_forces_f32", None, VP, VP)
        sig("or_forces_f64", None, VP, VP)
        sig("or_energies_f32", None, VP, VP)
        sig("or_energies_f64", None, VP, VP)
        sig("or_sum_energies", None, VP, VP)
        sig("or_build_possible", None, VP)
        sig("or_build_verlet", None, VP)
        sig("or_run", C.c_longlong, VP, C.c_longlong)
        sig("or_run_f64", C.c_longlong, VP, C.c_longlong, VP)
        sig("or_stride", C.c_int, VP)
        sig("or_get_indentation", None, VP, VP)
        sig("or_get_dat", C.c_int, VP, VP, C.c_int)
        sig("or_get_pull_rows", C.c_int, VP, VP, C.c_int)
        sig("or_get_ind_rows", C.c_int, VP, VP, C.c_int)
        sig("or_generate_seeds", None, C.c_int, C.c_int, VP)
        sig("or_hybrid_taus", C.c_uint32, VP)
        sig("or_uint_to_float", C.c_float, C.c_uint32)
        sig("or_logf", C.c_float, C.c_float)
        sig("or_sincosf", None, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float))
        sig("or_rforce", None, VP, VP)
        sig("or_bench_steps", C.c_double, VP, C.c_int, C.c_int)
        sig("or_hi_tea", None, VP, VP, VP)
        sig("or_hi_factor", C.c_int, VP, VP)
        sig("or_rpy_matrix", None, C.c_int, VP, C.c_double, VP)
        _lib = L
    return _lib


def make_params(**kw) -> OrParams:
    """Same defaults as the engine bindings; kw overrides."""
    from_defaults = dict(_DEFAULTS)
    for k in kw:
        if k not in from_defaults:
            raise KeyError(k)
    from_defaults.update(kw)
    p = OrParams()
    for name, _ in OrParams._fields_:
        v = from_defaults[name]
        if isinstance(v, (tuple, list, np.ndarray)):
            getattr(p, name)[:] = [float(x) for x in v]
        else:
            setattr(p, name, v)
    return p


_DEFAULTS = dict(
    kspring_cov=20.0, R_limit=2.0, R_limit_bond=8.0,
    pairs_cutoff=20.0, el=1.0, a=3.8,
    pairlist_cutoff=20.0, pairs_threshold=200.0,
    pairs_freq=1000, possiblepairs_freq=100000,
    max_covalent=8, max_native=128, max_pairs=512, max_possiblePairs=4096,
    timestep=0.16, zeta=50.0, temperature=0.6, minimization=0, seed=1,
    heating=0, tempFreq=1, initialT=0.0, deltaT=0.0,
    nav=1000, outputtiming=10000, dcdfreq=10000,
    pulling=0, deltax=0.0, k_trans=0.05, fconst=0.0, pullFreq=1000, pullDirectionVector=0,
    pullVector=(1.0, 0.0, 0.0), fixedEnd=0, pulledEnd=0,
    indentation=0, indChip=(0, 0, 0), indTip=(0, 0, 0), indDirection=(0, 0, 1), indSurfR0=(0, 0, 0),
    indSurfN=(0, 0, 1), indTipR=50.0, indKs=0.07, indDeltaX=0.0, indMoveSurface=0, indFixTrans=0,
    indTipSigma=1.0, indSurfSigma=1.0, indTipEl=1.0, indSurfEl=1.0,
    indTipA=0.0, indTipB=1.0, indSurfA=0.0, indSurfB=1.0, indTipZeta=5000.0, indOutputFreq=1000,
    indRetractionStep=-1,
    hi_on=0, hi_exact=0, hi_a=1.8, hi_epsilon_freq=10, hi_capricious=1, hi_unlisted=1, hi_epsmax=999.0,
    hi_chol_freq=1, hi_faithful_blocks=1, updater_freq_min=0, host_forces=0,
)


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class Sim:
    """One oracle system (ntr identical trajectories of the topology)."""

    def __init__(self, params: OrParams, system, ntr: int = 1):
        L = lib()
        s = system.normalized() if hasattr(system, "normalized") else system
        self._keep = [s.coords, np.ascontiguousarray(s.bonds[:, 0]), np.ascontiguousarray(s.bonds[:, 1]), s.bond_r0,
                      np.ascontiguousarray(s.natives[:, 0]), np.ascontiguousarray(s.natives[:, 1]), s.nat_r0, s.nat_eh,
                      np.ascontiguousarray(s.pairs[:, 0]), np.ascontiguousarray(s.pairs[:, 1]), s.fixed, s.pulled]
        coords = s.coords
        if ntr > 1:
            coords = np.ascontiguousarray(np.tile(s.coords, (ntr, 1)))
            self._keep.append(coords)
        k = self._keep
        self.params = params
        self.h = L.or_create(C.byref(params), s.n, ntr, _p(coords), len(s.bond_r0), _p(k[1]), _p(k[2]), _p(k[3]),
                             len(s.nat_r0), _p(k[4]), _p(k[5]), _p(k[6]), _p(k[7]), len(s.pairs), _p(k[8]), _p(k[9]),
                             len(s.fixed), _p(k[10]), len(s.pulled), _p(k[11]))
        if not self.h:
            raise RuntimeError(L.or_error().decode())
        self.n = L.or_n(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def coords(self):
        o = np.zeros((self.n, 4), np.float32)
        lib().or_get_coords(self.h, o.ctypes.data)
        return o

    def set_coords(self, c):
        c = np.ascontiguousarray(c, np.float32)
        lib().or_set_coords(self.h, c.ctypes.data)

    def temperature(self):
        o = np.zeros(self.n, np.float32)
        lib().or_get_T(self.h, o.ctypes.data)
        return o

    def rng(self):
        o = np.zeros((self.n, 4), np.uint32)
        lib().or_get_rng(self.h, o.ctypes.data)
        return o

    def ext(self):
        o = np.zeros((self.n, 4), np.float32)
        lib().or_get_ext(self.h, o.ctypes.data)
        return o

    def step(self):
        return lib().or_step(self.h)

    def stride(self):
        return lib().or_stride(self.h)

    def list(self, which):
        cap = lib().or_list_cap(self.h, which)
        cnt = np.zeros(self.n, np.int32)
        ent = np.zeros(cap * self.n, np.int32)
        lib().or_get_list(self.h, which, cnt.ctypes.data, ent.ctypes.data)
        return cnt, ent.reshape(cap, self.n)

    def overflow(self):
        return lib().or_overflow(self.h)

    def forces32(self):
        o = np.zeros((self.n, 4), np.float32)
        lib().or_forces_f32(self.h, o.ctypes.data)
        return o

    def forces64(self):
        o = np.zeros((self.n, 3), np.float64)
        lib().or_forces_f64(self.h, o.ctypes.data)
        return o

    def energies32(self):
        o = np.zeros(3 * self.n, np.float32)
        lib().or_energies_f32(self.h, o.ctypes.data)
        return o.reshape(3, self.n)

    def energies64(self):
        o = np.zeros(3 * self.n, np.float64)
        lib().or_energies_f64(self.h, o.ctypes.data)
        return o.reshape(3, self.n)

    def sum_energies(self, ntr=1):
        o = np.zeros(3 * ntr, np.float32)
        lib().or_sum_energies(self.h, o.ctypes.data)
        return o.reshape(3, ntr)

    def build_possible(self):
        lib().or_build_possible(self.h)

    def build_verlet(self):
        lib().or_build_verlet(self.h)

    def run(self, numsteps):
        return lib().or_run(self.h, numsteps)

    def run64(self, numsteps):
        o = np.zeros((self.n, 3), np.float64)
        done = lib().or_run_f64(self.h, numsteps, o.ctypes.data)
        if done < 0:
            raise RuntimeError(lib().or_error().decode())
        return o

    def indentation(self):
        o = np.zeros(9, np.float32)
        lib().or_get_indentation(self.h, o.ctypes.data)
        return o

    def dat(self, cap=100000):
        o = np.zeros((cap, 6), np.float64)
        m = lib().or_get_dat(self.h, o.ctypes.data, cap)
        return o[:min(m, cap)]

    def pull_rows(self, cap=100000):
        """Rows of the pull .dat: traj, step, |end-to-end|, end-to-end . n,
        F . n, F x, y, z (pulling.cu:166-187)."""
        o = np.zeros((cap, 8), np.float64)
        m = lib().or_get_pull_rows(self.h, o.ctypes.data, cap)
        return o[:min(m, cap)]

    def ind_rows(self, cap=100000):
        """Rows of the indentation .dat (indentation.cu:509-536)."""
        o = np.zeros((cap, 15), np.float64)
        m = lib().or_get_ind_rows(self.h, o.ctypes.data, cap)
        return o[:min(m, cap)]

    def hi_tea(self, ntr=1):
        e = np.zeros(ntr, np.float64)
        b = np.zeros(ntr, np.float32)
        lib().or_hi_tea(self.h, e.ctypes.data, b.ctypes.data)
        return e, b

    def hi_factor(self):
        m = 3 * self.n
        L = np.zeros((m, m), np.float32)
        if lib().or_hi_factor(self.h, L.ctypes.data) != 0:
            raise RuntimeError("no Cholesky factor (hi_exact off)")
        return L

    def bench(self, steps, threads):
        return lib().or_bench_steps(self.h, steps, threads)


def rpy_matrix(coords, a):
    c = np.ascontiguousarray(coords, np.float32)
    n = c.shape[0]
    D = np.zeros((3 * n, 3 * n), np.float64)
    lib().or_rpy_matrix(n, c.ctypes.data, float(a), D.ctypes.data)
    return D


def generate_seeds(seed, n):
    o = np.zeros((n, 4), np.uint32)
    lib().or_generate_seeds(seed, n, o.ctypes.data)
    return o


def ht_stream(seeds, count):
    """uint32 outputs of `count` Hybrid Taus draws for each seed row."""
    s = np.array(seeds, dtype=np.uint32).copy()
    out = np.zeros((s.shape[0], count), np.uint32)
    for i in range(s.shape[0]):
        st = np.ascontiguousarray(s[i])
        for k in range(count):
            out[i, k] = lib().or_hybrid_taus(st.ctypes.data)
        s[i] = st
    return out, s


def gaussians(seeds, count):
    s = np.array(seeds, dtype=np.uint32).copy()
    out = np.zeros((s.shape[0], count, 3), np.float32)
    tmp = np.zeros(4, np.float32)
    for i in range(s.shape[0]):
        st = np.ascontiguousarray(s[i])
        for k in range(count):
            lib().or_rforce(st.ctypes.data, tmp.ctypes.data)
            out[i, k] = tmp[:3]
        s[i] = st
    return out, s


def logf(x):
    return lib().or_logf(float(x))


def sincosf(x):
    s, c = C.c_float(), C.c_float()
    lib().or_sincosf(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def ref_seeds(seed, n):
    """Seed table from the reference's own ran2 (oracle/_ref)."""
    out = subprocess.run([REF_HARNESS, "seeds", str(seed), str(n)], check=True, capture_output=True, text=True).stdout
    return np.array([[int(x) for x in line.split()] for line in out.strip().splitlines()], dtype=np.uint32)


def ref_ran2(idum, count):
    out = subprocess.run([REF_HARNESS, "ran2", str(idum), str(count)], check=True, capture_output=True,
                         text=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    return np.array([float(r[0]) for r in rows]), np.array([int(r[1]) for r in rows])
