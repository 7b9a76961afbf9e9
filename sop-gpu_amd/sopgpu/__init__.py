"""ctypes bindings to libsopgpu.so (include/sopgpu.h).

This module is plumbing for tests, benchmarks and scripting: all compute
happens in the HIP library.  Importing it loads the library and fails
loudly when it is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SOPGPU_LIB names another build of the library (e.g. one compiled with
# different tile parameters); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("SOPGPU_LIB") or os.path.join(os.path.dirname(_HERE), "libsopgpu.so")


class SopError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libsopgpu.so not built at {LIB_PATH}; run __graft_entry__.build()")
    return C.CDLL(LIB_PATH)


lib = _load()

F3 = C.c_float * 3

PARAM_FIELDS = [
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


class Params(C.Structure):
    _fields_ = PARAM_FIELDS


# Defaults of the reference's PARAMETER() declarations (Appendix B of the
# reference's headers); fields the reference marks mandatory get the values
# of its example configs.
DEFAULTS = dict(
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


def make_params(struct_cls=Params, **kw):
    d = dict(DEFAULTS)
    for k in kw:
        if k not in d:
            raise KeyError(f"unknown parameter {k}")
    d.update(kw)
    p = struct_cls()
    for name, ty in struct_cls._fields_:
        v = d[name]
        if isinstance(v, (tuple, list, np.ndarray)):
            getattr(p, name)[:] = [float(x) for x in v]
        else:
            setattr(p, name, v)
    return p


class SystemStruct(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("coords4", C.c_void_p),
        ("nbonds", C.c_int), ("bond_i", C.c_void_p), ("bond_j", C.c_void_p), ("bond_r0", C.c_void_p),
        ("nnative", C.c_int), ("nat_i", C.c_void_p), ("nat_j", C.c_void_p), ("nat_r0", C.c_void_p),
        ("nat_eh", C.c_void_p),
        ("npairs", C.c_int), ("pair_i", C.c_void_p), ("pair_j", C.c_void_p),
        ("nfixed", C.c_int), ("fixed", C.c_void_p), ("npulled", C.c_int), ("pulled", C.c_void_p),
        ("rng_state4", C.c_void_p),
    ]


class SynthStruct(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("coords4", C.POINTER(C.c_float)),
        ("nbonds", C.c_int), ("bond_i", C.POINTER(C.c_int)), ("bond_j", C.POINTER(C.c_int)),
        ("bond_r0", C.POINTER(C.c_float)),
        ("nnative", C.c_int), ("nat_i", C.POINTER(C.c_int)), ("nat_j", C.POINTER(C.c_int)),
        ("nat_r0", C.POINTER(C.c_float)), ("nat_eh", C.POINTER(C.c_float)),
        ("npairs", C.c_int), ("pair_i", C.POINTER(C.c_int)), ("pair_j", C.POINTER(C.c_int)),
    ]


@dataclass
class System:
    """One SOP system in .top order (numpy arrays)."""

    coords: np.ndarray                      # (n,4) float32
    bonds: np.ndarray                       # (nb,2) int32
    bond_r0: np.ndarray                     # (nb,) float32
    natives: np.ndarray                     # (nn,2) int32
    nat_r0: np.ndarray
    nat_eh: np.ndarray
    pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    fixed: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pulled: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    @property
    def n(self) -> int:
        return int(self.coords.shape[0])

    def normalized(self) -> "System":
        c = np.ascontiguousarray(self.coords, dtype=np.float32)
        if c.shape[1] == 3:
            c = np.ascontiguousarray(np.concatenate([c, np.zeros((c.shape[0], 1), np.float32)], axis=1))
        return System(
            c,
            np.ascontiguousarray(self.bonds, dtype=np.int32).reshape(-1, 2),
            np.ascontiguousarray(self.bond_r0, dtype=np.float32),
            np.ascontiguousarray(self.natives, dtype=np.int32).reshape(-1, 2),
            np.ascontiguousarray(self.nat_r0, dtype=np.float32),
            np.ascontiguousarray(self.nat_eh, dtype=np.float32),
            np.ascontiguousarray(self.pairs, dtype=np.int32).reshape(-1, 2),
            np.ascontiguousarray(self.fixed, dtype=np.int32),
            np.ascontiguousarray(self.pulled, dtype=np.int32),
        )

    def struct(self):
        s = self.normalized()
        keep = [s.coords, np.ascontiguousarray(s.bonds[:, 0]), np.ascontiguousarray(s.bonds[:, 1]), s.bond_r0,
                np.ascontiguousarray(s.natives[:, 0]), np.ascontiguousarray(s.natives[:, 1]), s.nat_r0, s.nat_eh,
                np.ascontiguousarray(s.pairs[:, 0]), np.ascontiguousarray(s.pairs[:, 1]), s.fixed, s.pulled]
        p = lambda a: a.ctypes.data if a.size else None  # noqa: E731
        st = SystemStruct(s.n, p(keep[0]), len(s.bond_r0), p(keep[1]), p(keep[2]), p(keep[3]),
                          len(s.nat_r0), p(keep[4]), p(keep[5]), p(keep[6]), p(keep[7]),
                          len(s.pairs), p(keep[8]), p(keep[9]), len(s.fixed), p(keep[10]), len(s.pulled), p(keep[11]),
                          None)
        return st, keep


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


VP = C.c_void_p
_sig("sop_last_error", C.c_char_p)
_sig("sop_version", C.c_char_p)
_sig("sop_engine_create", C.c_int, C.c_int, C.c_ulonglong, C.POINTER(VP))
_sig("sop_engine_destroy", None, VP)
_sig("sop_engine_add_job", C.c_int, VP, C.POINTER(Params), C.POINTER(SystemStruct), C.c_longlong, C.POINTER(C.c_int))
_sig("sop_engine_remove_job", C.c_int, VP, C.c_int)
_sig("sop_engine_job_bytes", C.c_int, C.POINTER(Params), C.POINTER(SystemStruct), C.POINTER(C.c_ulonglong))
_sig("sop_engine_set_tuning", C.c_int, VP, C.c_int, C.c_int)
_sig("sop_engine_set_cells", C.c_int, VP, C.c_int, C.c_int)
_sig("sop_engine_set_lds_step", C.c_int, VP, C.c_int)
class Tuning(C.Structure):
    _fields_ = [("max_blocks", C.c_int), ("unit_rows", C.c_int), ("shared_units", C.c_int),
                ("pair_batch", C.c_int), ("verlet_chunk", C.c_int), ("sort_every", C.c_int),
                ("split_rows", C.c_int), ("split_global", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_sig("sop_engine_autotune", C.c_int, VP, C.POINTER(Tuning), C.POINTER(C.c_double), C.POINTER(C.c_double),
     C.POINTER(C.c_int))
_sig("sop_engine_get_shape", C.c_int, VP, C.POINTER(Tuning))
_sig("sop_engine_set_shape", C.c_int, VP, C.POINTER(Tuning))
_sig("sop_engine_run_blocks", C.c_int, VP, C.c_int)
_sig("sop_engine_run_job", C.c_int, VP, C.c_int)
_sig("sop_engine_block_steps", C.c_int, VP, C.POINTER(C.c_int))
_sig("sop_engine_synchronize", C.c_int, VP)
_sig("sop_engine_job_status", C.c_int, VP, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
     C.POINTER(C.c_int))
_sig("sop_engine_active_jobs", C.c_int, VP, C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("sop_engine_get_coords", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_set_coords", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_get_temperature", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_get_rng", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_get_ext", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_list_cap", C.c_int, VP, C.c_int, C.c_int, C.POINTER(C.c_int))
_sig("sop_engine_get_list", C.c_int, VP, C.c_int, C.c_int, VP, VP)
_sig("sop_engine_get_indentation", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_compute_forces", C.c_int, VP, C.c_int, VP)
_sig("sop_engine_compute_energies", C.c_int, VP, C.c_int, VP, VP)
_sig("sop_engine_rebuild_lists", C.c_int, VP, C.c_int, C.c_int)
_sig("sop_engine_profile", C.c_int, VP, C.c_int)
_sig("sop_engine_profile_read", C.c_int, VP, C.POINTER(C.c_longlong), C.POINTER(C.c_double))
_sig("sop_engine_list_stats", C.c_int, VP, VP, C.POINTER(C.c_longlong))
_sig("sop_engine_set_chol_mode", C.c_int, VP, C.c_int)
_sig("sop_engine_get_hi", C.c_int, VP, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float))
_sig("sop_engine_get_hi_factor", C.c_int, VP, C.c_int, VP)
_sig("sop_chol_factor", C.c_int, C.c_int, C.c_int, VP, C.c_int, VP, C.c_int, C.POINTER(C.c_double))
_sig("sop_chol_factor_noalloc", C.c_int, C.c_int, C.c_int, VP, C.c_ulonglong, C.c_int, VP)
_sig("sop_server_create", C.c_int, C.c_int, C.c_ulonglong, C.POINTER(VP))
_sig("sop_server_destroy", None, VP)
_sig("sop_server_set_tuning", C.c_int, VP, C.c_int, C.c_int)
_sig("sop_server_set_shape", C.c_int, VP, C.POINTER(Tuning))
_sig("sop_server_autotune", C.c_int, VP, C.POINTER(Tuning), C.POINTER(C.c_double), C.POINTER(C.c_double),
     C.POINTER(C.c_int))
_sig("sop_server_submit", C.c_int, VP, C.POINTER(Params), C.POINTER(SystemStruct), C.c_longlong, C.POINTER(C.c_int))
_sig("sop_server_pump", C.c_int, VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("sop_server_poll", C.c_int, VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
     C.POINTER(C.c_double))
_sig("sop_server_result", C.c_int, VP, C.c_int, VP)
_sig("sop_server_stats", C.c_int, VP, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
     C.POINTER(C.c_int), C.POINTER(C.c_ulonglong))
_sig("sop_copy_bandwidth", C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.POINTER(C.c_double))
_sig("sop_generate_seeds", C.c_int, C.c_int, C.c_int, VP)
_sig("sop_rng_probe", C.c_int, C.c_int, VP, C.c_int, C.c_int, C.c_int, VP, VP)
_sig("sop_check_division", C.c_int, C.c_int, C.c_uint, C.c_uint, VP, C.c_int, C.POINTER(C.c_ulonglong))
_sig("sop_synth_chain", C.c_int, C.c_int, C.c_ulonglong, C.c_float, C.c_float, C.c_float, C.POINTER(SynthStruct))
_sig("sop_synth_shell", C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_float, C.c_float, C.c_float,
     C.POINTER(SynthStruct))
_sig("sop_free_system", None, C.POINTER(SynthStruct))
_sig("sop_read_topology", C.c_int, C.c_char_p, C.c_char_p, C.POINTER(SynthStruct))
_sig("sop_write_pdb", C.c_int, C.c_char_p, VP, C.c_int, C.c_char_p)
_sig("sop_write_dcd", C.c_int, C.c_char_p, C.c_int, C.c_int, VP, C.c_int, C.c_float)
_sig("sop_run_config", C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p))

# host plugins (sop_plugin.h)
PLUGIN_FN = C.CFUNCTYPE(C.c_int, VP, VP)
ENERGY_FN = C.CFUNCTYPE(C.c_float, VP, C.c_int, C.c_int, VP)
RELEASE_FN = C.CFUNCTYPE(None, VP)
_sig("sop_sim_create", C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.POINTER(VP))
_sig("sop_sim_run", C.c_int, VP)
_sig("sop_sim_destroy", None, VP)
_sig("sop_sim_fail", None, C.c_char_p)
_sig("sop_sim_add_updater", C.c_int, VP, C.c_char_p, C.c_int, PLUGIN_FN, VP, RELEASE_FN)
_sig("sop_sim_add_potential", C.c_int, VP, C.c_char_p, C.c_int, PLUGIN_FN, ENERGY_FN, VP, RELEASE_FN)
_sig("sop_sim_potential_count", C.c_int, VP, C.POINTER(C.c_int))
_sig("sop_sim_updater_count", C.c_int, VP, C.POINTER(C.c_int))
_sig("sop_sim_potential_name", C.c_char_p, VP, C.c_int)
_sig("sop_sim_updater_name", C.c_char_p, VP, C.c_int)
_sig("sop_sim_updater_frequency", C.c_int, VP, C.c_int, C.POINTER(C.c_int))
_sig("sop_sim_potential_by_name", C.c_int, VP, C.c_char_p, C.POINTER(C.c_int))
_sig("sop_sim_updater_by_name", C.c_int, VP, C.c_char_p, C.POINTER(C.c_int))
_sig("sop_sim_integrator_name", C.c_char_p, VP)
_sig("sop_sim_get_energy", C.c_int, VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float))
_sig("sop_sim_state", C.c_int, VP, C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("sop_sim_get_coords", C.c_int, VP, C.c_int, VP)
_sig("sop_sim_add_forces", C.c_int, VP, C.c_int, VP)

EXPORTED = [
    "sop_last_error", "sop_version", "sop_engine_create", "sop_engine_destroy", "sop_engine_add_job",
    "sop_engine_remove_job", "sop_engine_job_bytes", "sop_engine_set_tuning", "sop_engine_set_cells", "sop_engine_set_lds_step", "sop_engine_autotune",
    "sop_engine_get_shape", "sop_engine_set_shape",
    "sop_engine_run_blocks", "sop_engine_run_job", "sop_engine_block_steps", "sop_engine_synchronize",
    "sop_engine_job_status", "sop_engine_active_jobs", "sop_engine_get_coords", "sop_engine_set_coords",
    "sop_engine_get_temperature", "sop_engine_get_rng", "sop_engine_get_ext", "sop_engine_list_cap",
    "sop_engine_get_list", "sop_engine_get_indentation", "sop_engine_compute_forces",
    "sop_engine_compute_energies", "sop_engine_rebuild_lists", "sop_engine_profile", "sop_engine_profile_read",
    "sop_engine_list_stats", "sop_engine_set_chol_mode", "sop_engine_get_hi", "sop_engine_get_hi_factor",
    "sop_chol_factor", "sop_chol_factor_noalloc", "sop_server_create", "sop_server_destroy", "sop_server_set_tuning", "sop_server_set_shape",
    "sop_server_autotune", "sop_server_submit",
    "sop_server_pump", "sop_server_poll", "sop_server_result", "sop_server_stats", "sop_copy_bandwidth",
    "sop_generate_seeds", "sop_rng_probe", "sop_check_division", "sop_synth_chain", "sop_synth_shell",
    "sop_free_system", "sop_run_config", "sop_read_topology", "sop_write_pdb", "sop_write_dcd",
    "sop_sim_create", "sop_sim_run", "sop_sim_destroy", "sop_sim_fail", "sop_sim_add_updater",
    "sop_sim_add_potential", "sop_sim_potential_count", "sop_sim_updater_count", "sop_sim_potential_name",
    "sop_sim_updater_name", "sop_sim_updater_frequency", "sop_sim_potential_by_name", "sop_sim_updater_by_name",
    "sop_sim_integrator_name", "sop_sim_get_energy", "sop_sim_state", "sop_sim_get_coords", "sop_sim_add_forces",
]


def check(rc: int):
    if rc != 0:
        raise SopError(rc, lib.sop_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data


def _from_synth(s: SynthStruct) -> System:
    n = s.n
    coords = np.ctypeslib.as_array(s.coords4, shape=(n * 4,)).reshape(n, 4).copy()
    nb, nn = s.nbonds, s.nnative
    bonds = np.stack([np.ctypeslib.as_array(s.bond_i, shape=(max(nb, 1),))[:nb],
                      np.ctypeslib.as_array(s.bond_j, shape=(max(nb, 1),))[:nb]], axis=1).astype(np.int32)
    br0 = np.ctypeslib.as_array(s.bond_r0, shape=(max(nb, 1),))[:nb].copy()
    nat = np.stack([np.ctypeslib.as_array(s.nat_i, shape=(max(nn, 1),))[:nn],
                    np.ctypeslib.as_array(s.nat_j, shape=(max(nn, 1),))[:nn]], axis=1).astype(np.int32)
    nr0 = np.ctypeslib.as_array(s.nat_r0, shape=(max(nn, 1),))[:nn].copy()
    neh = np.ctypeslib.as_array(s.nat_eh, shape=(max(nn, 1),))[:nn].copy()
    npr = s.npairs
    if npr:
        pairs = np.stack([np.ctypeslib.as_array(s.pair_i, shape=(npr,)),
                          np.ctypeslib.as_array(s.pair_j, shape=(npr,))], axis=1).astype(np.int32)
    else:
        pairs = np.zeros((0, 2), np.int32)
    return System(coords, bonds, br0, nat, nr0, neh, pairs)


def synth_chain(n: int, seed: int, density: float = 175.0, eh=(0.9, 1.5)) -> System:
    s = SynthStruct()
    check(lib.sop_synth_chain(n, seed, density, eh[0], eh[1], C.byref(s)))
    try:
        return _from_synth(s)
    finally:
        lib.sop_free_system(C.byref(s))


def synth_shell(nchains: int, n: int, seed: int, radius: float = 140.0, eh=(0.9, 1.5)) -> System:
    s = SynthStruct()
    check(lib.sop_synth_shell(nchains, n, seed, radius, eh[0], eh[1], C.byref(s)))
    try:
        return _from_synth(s)
    finally:
        lib.sop_free_system(C.byref(s))


def read_topology(top_path: str, coords_path: Optional[str] = None) -> System:
    """SOP::load of a .top file (src/IO/topio.cpp:36-80), plus readCoord of a
    PDB/XYZ file (src/IO/pdbio.cpp:292-337) when coords_path is given."""
    s = SynthStruct()
    check(lib.sop_read_topology(top_path.encode(), coords_path.encode() if coords_path else None, C.byref(s)))
    try:
        return _from_synth(s)
    finally:
        lib.sop_free_system(C.byref(s))


def write_pdb(top_path: str, coords: np.ndarray, out_path: str) -> None:
    """savePDB (src/IO/pdbio.cpp:194-219) of the .top's atoms at coords."""
    c = np.ascontiguousarray(np.asarray(coords, np.float32)[:, :3])
    c4 = np.ascontiguousarray(np.concatenate([c, np.zeros((c.shape[0], 1), np.float32)], axis=1))
    check(lib.sop_write_pdb(top_path.encode(), _ptr(c4), c4.shape[0], out_path.encode()))


def write_dcd(path: str, frames: np.ndarray, freq: int = 1000, delta: float = 0.0) -> None:
    """DCD file (src/IO/dcdio.cpp:100-194); frames is [nframes, 3, natoms]."""
    f = np.ascontiguousarray(frames, dtype=np.float32)
    check(lib.sop_write_dcd(path.encode(), f.shape[2], f.shape[0], _ptr(f), freq, delta))


def generate_seeds(seed: int, n: int) -> np.ndarray:
    out = np.zeros((n, 4), np.uint32)
    check(lib.sop_generate_seeds(seed, n, _ptr(out)))
    return out


def rng_probe(seeds: np.ndarray, count: int, gaussian: bool, device: int = 0):
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    n = seeds.shape[0]
    out = np.zeros((n, count, 3) if gaussian else (n, count), np.float32 if gaussian else np.uint32)
    st = np.zeros((n, 4), np.uint32)
    check(lib.sop_rng_probe(device, _ptr(seeds), n, count, int(gaussian), _ptr(out), _ptr(st)))
    return out, st


def check_division(lo: int, hi: int, a, device: int = 0) -> int:
    """Mismatches of the step kernel's short division against IEEE a/b over
    the float bit patterns b in [lo, hi) and the numerators a."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    bad = C.c_ulonglong(0)
    check(lib.sop_check_division(device, lo, hi, _ptr(a), len(a), C.byref(bad)))
    return int(bad.value)


class Engine:
    """An engine on one GPU holding any number of jobs."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = VP()
        check(lib.sop_engine_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self._n = {}

    def close(self):
        if self.h:
            lib.sop_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_job(self, params: Params, system: System, numsteps: int) -> int:
        st, keep = system.struct()
        jid = C.c_int()
        check(lib.sop_engine_add_job(self.h, C.byref(params), C.byref(st), numsteps, C.byref(jid)))
        self._n[jid.value] = system.n
        del keep
        return jid.value

    def remove_job(self, jid: int):
        check(lib.sop_engine_remove_job(self.h, jid))
        self._n.pop(jid, None)

    def set_tuning(self, max_blocks: int = 0, use_graph: bool = True):
        check(lib.sop_engine_set_tuning(self.h, max_blocks, int(use_graph)))

    def set_cells(self, min_beads: int = 0, enabled: bool = True):
        check(lib.sop_engine_set_cells(self.h, min_beads, int(enabled)))

    def set_lds_step(self, enabled: bool = True):
        check(lib.sop_engine_set_lds_step(self.h, int(enabled)))

    def autotune(self) -> dict:
        """Searches the launch shapes on the admitted jobs (they do not
        advance) and keeps the fastest; returns the shape and the ms per step
        before and after."""
        t, a, b, n = Tuning(), C.c_double(), C.c_double(), C.c_int()
        check(lib.sop_engine_autotune(self.h, C.byref(t), C.byref(a), C.byref(b), C.byref(n)))
        return dict(shape=t.as_dict(), ms_per_step_start=a.value, ms_per_step_best=b.value, trials=n.value)

    def shape(self) -> dict:
        t = Tuning()
        check(lib.sop_engine_get_shape(self.h, C.byref(t)))
        return t.as_dict()

    def set_shape(self, **kw):
        t = Tuning()
        check(lib.sop_engine_get_shape(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib.sop_engine_set_shape(self.h, C.byref(t)))

    def run_blocks(self, nblocks: int):
        check(lib.sop_engine_run_blocks(self.h, nblocks))

    def run_job(self, jid: int):
        check(lib.sop_engine_run_job(self.h, jid))

    def block_steps(self) -> int:
        L = C.c_int()
        check(lib.sop_engine_block_steps(self.h, C.byref(L)))
        return L.value

    def synchronize(self):
        check(lib.sop_engine_synchronize(self.h))

    def status(self, jid: int):
        s, e, o = C.c_longlong(), C.c_longlong(), C.c_int()
        check(lib.sop_engine_job_status(self.h, jid, C.byref(s), C.byref(e), C.byref(o)))
        return s.value, e.value, o.value

    def active(self):
        c, m = C.c_int(), C.c_int()
        check(lib.sop_engine_active_jobs(self.h, C.byref(c), C.byref(m)))
        return c.value, m.value

    def coords(self, jid: int) -> np.ndarray:
        out = np.zeros((self._n[jid], 4), np.float32)
        check(lib.sop_engine_get_coords(self.h, jid, _ptr(out)))
        return out

    def set_coords(self, jid: int, c: np.ndarray):
        c = np.ascontiguousarray(c, dtype=np.float32)
        check(lib.sop_engine_set_coords(self.h, jid, _ptr(c)))

    def temperature(self, jid: int) -> np.ndarray:
        out = np.zeros(self._n[jid], np.float32)
        check(lib.sop_engine_get_temperature(self.h, jid, _ptr(out)))
        return out

    def rng(self, jid: int) -> np.ndarray:
        out = np.zeros((self._n[jid], 4), np.uint32)
        check(lib.sop_engine_get_rng(self.h, jid, _ptr(out)))
        return out

    def ext(self, jid: int) -> np.ndarray:
        out = np.zeros((self._n[jid], 4), np.float32)
        check(lib.sop_engine_get_ext(self.h, jid, _ptr(out)))
        return out

    def list(self, jid: int, which: int):
        cap = C.c_int()
        check(lib.sop_engine_list_cap(self.h, jid, which, C.byref(cap)))
        n = self._n[jid]
        cnt = np.zeros(n, np.int32)
        ent = np.zeros(cap.value * n, np.int32)
        check(lib.sop_engine_get_list(self.h, jid, which, _ptr(cnt), _ptr(ent)))
        return cnt, ent.reshape(cap.value, n)

    def indentation(self, jid: int) -> np.ndarray:
        out = np.zeros(13, np.float32)
        check(lib.sop_engine_get_indentation(self.h, jid, _ptr(out)))
        return out

    def forces(self, jid: int) -> np.ndarray:
        out = np.zeros((self._n[jid], 4), np.float32)
        check(lib.sop_engine_compute_forces(self.h, jid, _ptr(out)))
        return out

    def energies(self, jid: int):
        n = self._n[jid]
        per = np.zeros(3 * n, np.float32)
        sums = np.zeros(3, np.float32)
        check(lib.sop_engine_compute_energies(self.h, jid, _ptr(per), _ptr(sums)))
        return per.reshape(3, n), sums

    def rebuild(self, jid: int, mask: int = 3):
        check(lib.sop_engine_rebuild_lists(self.h, jid, mask))

    def profile(self, on: bool):
        check(lib.sop_engine_profile(self.h, int(on)))

    def profile_read(self):
        n, ms = C.c_longlong(), C.c_double()
        check(lib.sop_engine_profile_read(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_chol_mode(self, mode: int):
        check(lib.sop_engine_set_chol_mode(self.h, mode))

    def hi(self, jid: int):
        e, b = C.c_double(), C.c_float()
        check(lib.sop_engine_get_hi(self.h, jid, C.byref(e), C.byref(b)))
        return e.value, b.value

    def hi_factor(self, jid: int) -> np.ndarray:
        m = 3 * self._n[jid]
        L = np.zeros((m, m), np.float32)
        check(lib.sop_engine_get_hi_factor(self.h, jid, _ptr(L)))
        return L

    def list_stats(self):
        avg = np.zeros(4, np.float64)
        nb = C.c_longlong()
        check(lib.sop_engine_list_stats(self.h, _ptr(avg), C.byref(nb)))
        return avg, nb.value


CHOL_BF16X3, CHOL_F32 = 0, 1


def chol_factor(A: np.ndarray, mode: int = CHOL_BF16X3, repeats: int = 1, device: int = 0):
    """Blocked Cholesky of a symmetric positive definite matrix on the GPU.
    Returns (L, ms_per_factor)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    n = A.shape[0]
    L = np.zeros_like(A)
    ms = C.c_double()
    check(lib.sop_chol_factor(device, n, _ptr(A), mode, _ptr(L), repeats, C.byref(ms)))
    return L, ms.value


def copy_bandwidth(nbytes: int = 1 << 30, repeats: int = 20, device: int = 0) -> float:
    """GB/s (read + write) of a plain streaming copy kernel."""
    g = C.c_double()
    check(lib.sop_copy_bandwidth(device, nbytes, repeats, C.byref(g)))
    return g.value


def run_config(path: str, **overrides) -> None:
    """Run the sop-gpu program on a parameter file (src/sop-gpu.cpp:18-88);
    keyword arguments override parameters of the file."""
    names = [k.encode() for k in overrides]
    values = [str(v).encode() for v in overrides.values()]
    n = len(names)
    check(lib.sop_run_config(path.encode(), n, (C.c_char_p * max(n, 1))(*names), (C.c_char_p * max(n, 1))(*values)))


def job_bytes(params: Params, system: System) -> int:
    st, keep = system.struct()
    b = C.c_ulonglong()
    check(lib.sop_engine_job_bytes(C.byref(params), C.byref(st), C.byref(b)))
    return b.value


class Server:
    """Continuous-batching request loop on one GPU (sop_server_*)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = VP()
        check(lib.sop_server_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self._n = {}

    def close(self):
        if self.h:
            lib.sop_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tuning(self, max_blocks: int = 0, use_graph: bool = True):
        check(lib.sop_server_set_tuning(self.h, max_blocks, int(use_graph)))

    def set_shape(self, shape: dict):
        t = Tuning(**{k: int(v) for k, v in shape.items()})
        check(lib.sop_server_set_shape(self.h, C.byref(t)))

    def autotune(self) -> dict:
        t, a, b, n = Tuning(), C.c_double(), C.c_double(), C.c_int()
        check(lib.sop_server_autotune(self.h, C.byref(t), C.byref(a), C.byref(b), C.byref(n)))
        return dict(shape=t.as_dict(), ms_per_step_start=a.value, ms_per_step_best=b.value, trials=n.value)

    def submit(self, params: Params, system: System, numsteps: int) -> int:
        st, keep = system.struct()
        t = C.c_int()
        check(lib.sop_server_submit(self.h, C.byref(params), C.byref(st), numsteps, C.byref(t)))
        self._n[t.value] = system.n
        return t.value

    def pump(self, nblocks: int = 1):
        r, q = C.c_int(), C.c_int()
        check(lib.sop_server_pump(self.h, nblocks, C.byref(r), C.byref(q)))
        return r.value, q.value

    def poll(self, ticket: int):
        st, a, b, c = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        check(lib.sop_server_poll(self.h, ticket, C.byref(st), C.byref(a), C.byref(b), C.byref(c)))
        return st.value, a.value, b.value, c.value

    def result(self, ticket: int) -> np.ndarray:
        out = np.zeros((self._n[ticket], 4), np.float32)
        check(lib.sop_server_result(self.h, ticket, _ptr(out)))
        return out

    def stats(self):
        bs, busy, adm, fin, peak = C.c_double(), C.c_double(), C.c_int(), C.c_int(), C.c_ulonglong()
        check(lib.sop_server_stats(self.h, C.byref(bs), C.byref(busy), C.byref(adm), C.byref(fin), C.byref(peak)))
        return dict(bead_steps=bs.value, busy_s=busy.value, admitted=adm.value, finished=fin.value,
                    peak_bytes=peak.value)


class Simulation:
    """The sop-gpu program with its plugin registries (sop_sim_*): create
    from a parameter file, add host updaters / potentials (Python callables
    wrapped as C callbacks), run once."""

    def __init__(self, path: str, **overrides):
        names = [k.encode() for k in overrides]
        values = [str(v).encode() for v in overrides.values()]
        n = len(names)
        h = VP()
        check(lib.sop_sim_create(path.encode(), n, (C.c_char_p * max(n, 1))(*names),
                                 (C.c_char_p * max(n, 1))(*values), C.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib.sop_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _wrap(fn):
        def tramp(sim, user):
            try:
                r = fn()
                return 0 if r is None else int(r)
            except Exception as e:  # the plugin's failure ends the run with its text
                lib.sop_sim_fail(str(e).encode())
                return -1
        return PLUGIN_FN(tramp)

    def add_updater(self, name: str, frequency: int, update):
        cb = self._wrap(update)
        self._keep.append(cb)
        check(lib.sop_sim_add_updater(self.h, name.encode(), frequency, cb, None, RELEASE_FN()))

    def add_potential(self, name: str, compute, energies_count: int = 0, energy=None):
        cb = self._wrap(compute)
        ecb = ENERGY_FN(lambda sim, traj, i, user: float(energy(traj, i))) if energy else ENERGY_FN()
        self._keep += [cb, ecb]
        check(lib.sop_sim_add_potential(self.h, name.encode(), energies_count, cb, ecb, None, RELEASE_FN()))

    def run(self):
        check(lib.sop_sim_run(self.h))

    def potentials(self):
        c = C.c_int()
        check(lib.sop_sim_potential_count(self.h, C.byref(c)))
        return [lib.sop_sim_potential_name(self.h, k).decode() for k in range(c.value)]

    def updaters(self):
        c = C.c_int()
        check(lib.sop_sim_updater_count(self.h, C.byref(c)))
        out = []
        for k in range(c.value):
            f = C.c_int()
            check(lib.sop_sim_updater_frequency(self.h, k, C.byref(f)))
            out.append((lib.sop_sim_updater_name(self.h, k).decode(), f.value))
        return out

    def potential_by_name(self, name: str) -> int:
        i = C.c_int()
        check(lib.sop_sim_potential_by_name(self.h, name.encode(), C.byref(i)))
        return i.value

    def updater_by_name(self, name: str) -> int:
        i = C.c_int()
        check(lib.sop_sim_updater_by_name(self.h, name.encode(), C.byref(i)))
        return i.value

    def integrator(self) -> str:
        return lib.sop_sim_integrator_name(self.h).decode()

    def state(self):
        st, n, ntr = C.c_longlong(), C.c_int(), C.c_int()
        check(lib.sop_sim_state(self.h, C.byref(st), C.byref(n), C.byref(ntr)))
        return st.value, n.value, ntr.value

    def coords(self, traj: int = 0) -> np.ndarray:
        n = self.state()[1]
        out = np.zeros((n, 4), np.float32)
        check(lib.sop_sim_get_coords(self.h, traj, _ptr(out)))
        return out

    def add_forces(self, traj: int, f4: np.ndarray):
        f = np.ascontiguousarray(f4, dtype=np.float32)
        check(lib.sop_sim_add_forces(self.h, traj, _ptr(f)))

    def energy(self, index: int, traj: int = 0, term: int = 0) -> float:
        e = C.c_float()
        check(lib.sop_sim_get_energy(self.h, index, traj, term, C.byref(e)))
        return e.value
