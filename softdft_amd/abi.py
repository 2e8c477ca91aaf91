"""ctypes binding of libsoftdft_hip (include/softdft_hip.h).

The library is the product path: there is no CPU fallback.  Loading fails
loudly when the shared object is missing; creating a context fails loudly
when no HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import List, Optional, Sequence

import numpy as np

from .problem import StatePoint

LIB_NAME = "libsoftdft_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

SDFT_ABI_VERSION = 5
SDFT_OK = 0
ERRORS = {-1: "SDFT_E_ARG", -2: "SDFT_E_HIP", -3: "SDFT_E_NORMALIZATION", -4: "SDFT_E_NOCONV",
          -5: "SDFT_E_UNSUPPORTED", -6: "SDFT_E_STATE"}
F_MFMT, F_TPT1MFMT, F_MSA, F_TPT1MSA, F_SW = 1, 2, 4, 8, 16
PREC_FP64, PREC_FP32, PREC_MIXED = 64, 32, 3264
FIELD_RHO_BEADS, FIELD_RHO_BONDS, FIELD_X, FIELD_FX = 1, 2, 3, 4
FIELD_RHO_SEGMENTS = 16  # + species index: rho_K.segments[u], (NP, S_u, N_u)

EXPORTED = ("sdft_abi_version", "sdft_last_error", "sdft_create", "sdft_destroy", "sdft_solution_length",
            "sdft_solution_offsets", "sdft_objective", "sdft_solve", "sdft_observables", "sdft_get_field",
            "sdft_last_timing", "sdft_last_pipeline_iters", "sdft_set_operator_band", "sdft_kernel_profile",
            "sdft_set_launch_mode", "sdft_autotune", "sdft_set_tuning", "sdft_observables_noalloc",
            "sdft_last_switch_iters", "sdft_bulk_solve", "sdft_set_chemistry", "sdft_set_graft_density")
TUNE_PROPAGATOR, TUNE_CONV_FWD, TUNE_CONV_PRE, TUNE_CONV_BACK = 1, 2, 3, 4
CONV_TALL, CONV_WIDE64, CONV_WIDE128, CONV_MID, CONV_NARROW, CONV_FFT = 0, 1, 2, 3, 4, 5
KP_PROP_FP32, KP_PROP_FP64, KP_CONV_FWD_FP32, KP_CONV_FWD_FP64, KP_CONV_BACK_FP32, KP_CONV_BACK_FP64, \
    KP_POINTWISE_FP64, KP_CHAIN_FP32, KP_CHAIN_FP64, KP_ACCUM_FP32, KP_ACCUM_FP64 = range(11)

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class SdftDesc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("n_points", C.c_int32), ("NB", C.c_int32), ("MB", C.c_int32),
                ("n_species", C.c_int32), ("functionals", C.c_int32), ("bjerrum", C.c_double),
                ("diameters", _f64p), ("valences", _f64p), ("bond_types", _i32p),
                ("species_nseg", _i32p), ("species_model", _i32p), ("species_nstates", _i32p),
                ("seg_states", _i32p), ("species_fixed_seg", _i32p), ("fixed_density", _f64p),
                ("max_distance", _f64p),
                ("NP", _i32p), ("dz", _f64p), ("delta_muH", _f64p), ("mu_species", _f64p),
                ("total_charge", _f64p), ("fixed_coord", _i32p),
                ("Ext", _f64p), ("trapez", _f64p), ("surf", _f64p),
                ("hard_walls", C.c_int32), ("wall_lo", _i32p), ("wall_hi", _i32p), ("wall_charge", _f64p),
                ("n_sw_pairs", C.c_int32), ("sw_pairs", _i32p), ("sw_energy", _f64p), ("sw_lambda", _f64p),
                ("sw_wall_idx", _i32p), ("sw_wall_fac", _f64p),
                ("ndim", C.c_int32), ("grid", _i32p), ("bins", _f64p)]


class SdftSolverOpts(C.Structure):
    _fields_ = [("tole", C.c_double), ("damping", C.c_double), ("mixing_max", C.c_double),
                ("switch_tole", C.c_double), ("anderm", C.c_int32), ("max_iters", C.c_int32),
                ("precision", C.c_int32), ("fp32_switch", C.c_double), ("band_fp32", C.c_int32),
                ("check_every", C.c_int32)]


class SdftTuneEntry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("key", C.c_int32), ("precision", C.c_int32), ("choice", C.c_int32),
                ("us", C.c_double * 6)]


class SdftError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    lib.sdft_abi_version.restype = C.c_int32
    lib.sdft_last_error.restype = C.c_char_p
    lib.sdft_last_error.argtypes = [C.c_void_p]
    lib.sdft_create.restype = C.c_int32
    lib.sdft_create.argtypes = [C.POINTER(SdftDesc), C.POINTER(C.c_void_p)]
    lib.sdft_destroy.restype = None
    lib.sdft_destroy.argtypes = [C.c_void_p]
    lib.sdft_solution_length.restype = C.c_int64
    lib.sdft_solution_length.argtypes = [C.c_void_p]
    lib.sdft_solution_offsets.restype = C.c_int32
    lib.sdft_solution_offsets.argtypes = [C.c_void_p, _i64p]
    lib.sdft_objective.restype = C.c_int32
    lib.sdft_objective.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int32]
    lib.sdft_solve.restype = C.c_int32
    lib.sdft_solve.argtypes = [C.c_void_p, C.POINTER(SdftSolverOpts), _f64p, _i32p, _f64p]
    lib.sdft_observables.restype = C.c_int32
    lib.sdft_observables.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    lib.sdft_last_switch_iters.restype = C.c_int32
    lib.sdft_last_switch_iters.argtypes = [C.c_void_p, _i32p]
    lib.sdft_observables_noalloc.restype = C.c_int32
    lib.sdft_observables_noalloc.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    lib.sdft_get_field.restype = C.c_int32
    lib.sdft_get_field.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, C.c_int64]
    lib.sdft_last_timing.restype = C.c_int32
    lib.sdft_last_timing.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.sdft_last_pipeline_iters.restype = C.c_int32
    lib.sdft_last_pipeline_iters.argtypes = [C.c_void_p, _i32p, _i32p]
    lib.sdft_set_operator_band.restype = C.c_int32
    lib.sdft_set_operator_band.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.sdft_kernel_profile.restype = C.c_int32
    lib.sdft_kernel_profile.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, _f64p]
    lib.sdft_autotune.restype = C.c_int32
    lib.sdft_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(SdftTuneEntry), C.c_int32, _i32p]
    lib.sdft_set_tuning.restype = C.c_int32
    lib.sdft_set_tuning.argtypes = [C.c_void_p, C.POINTER(SdftTuneEntry), C.c_int32]
    lib.sdft_set_chemistry.restype = C.c_int32
    lib.sdft_set_chemistry.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.sdft_set_graft_density.restype = C.c_int32
    lib.sdft_set_graft_density.argtypes = [C.c_void_p, _f64p]
    lib.sdft_set_launch_mode.restype = C.c_int32
    lib.sdft_set_launch_mode.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i64p]
    if lib.sdft_abi_version() != SDFT_ABI_VERSION:
        raise RuntimeError("libsoftdft_hip ABI version mismatch")
    _lib = lib
    return lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def functional_flags(names: Sequence[str]) -> int:
    f = 0
    for n in names:
        f |= {"mfmt": F_MFMT, "TPT1mfmt": F_TPT1MFMT, "msa": F_MSA, "TPT1msa": F_TPT1MSA, "sw": F_SW}[n]
    return f


def build_descriptor(points: List[StatePoint]):
    """The sdft_desc_t of a batch of state points that share one molecular-system layout, with the
    numpy arrays it points into (keep them alive while the descriptor is used)."""
    if not points:
        raise ValueError("empty batch")
    p0 = points[0]
    ms = p0.ms
    for sp in points[1:]:
        m2 = sp.ms
        if (m2.monomers != ms.monomers or m2.bond_types != ms.bond_types or len(m2.species) != len(ms.species)
                or not np.array_equal(m2.diameters, ms.diameters) or not np.array_equal(m2.valences, ms.valences)
                or sp.functionals != p0.functionals or abs(m2.bjerrum - ms.bjerrum) > 0
                or m2.sw != ms.sw):
            raise ValueError("state points in one batch must share the molecular-system layout")
    NB, MB, nsp = ms.NB, ms.MB, len(ms.species)
    P = len(points)
    keep = {}
    keep["diam"] = np.ascontiguousarray(ms.diameters, dtype=np.float64)
    keep["val"] = np.ascontiguousarray(ms.valences, dtype=np.float64)
    keep["bonds"] = np.ascontiguousarray(np.array(ms.bond_types, dtype=np.int32).reshape(-1))
    nseg, model, nst, states, fseg, fden, maxd = [], [], [], [], [], [], []
    for u, s in enumerate(ms.species):
        for q in range(s.n_seg):
            if len(s.topology.children[q]) > 1:
                raise SdftError(-5, f"species {u}: branched chains are outside this build")
        S = max(len(f) for f in s.state_family)
        nseg.append(s.n_seg)
        model.append(0 if s.model == "freely_jointed" else 1)
        nst.append(S)
        for f in s.state_family:
            states.extend(list(f) + [-1] * (S - len(f)))
        fx = p0.fixed[u]
        if fx.any:
            segs = np.nonzero(fx.segments)[0]
            if len(segs) != 1 or segs[0] != 0:
                raise SdftError(-5, f"species {u}: only a grafted first segment is supported")
            fseg.append(0)
            fden.append(fx.density)
            maxd.extend(list(s.max_distance[:, 0]) if s.max_distance is not None else [0.0] * s.n_seg)
        else:
            fseg.append(-1)
            fden.append(0.0)
            maxd.extend([0.0] * s.n_seg)
    keep["nseg"] = np.array(nseg, dtype=np.int32)
    keep["model"] = np.array(model, dtype=np.int32)
    keep["nst"] = np.array(nst, dtype=np.int32)
    keep["states"] = np.array(states, dtype=np.int32)
    keep["fseg"] = np.array(fseg, dtype=np.int32)
    keep["fden"] = np.array(fden, dtype=np.float64)
    keep["maxd"] = np.array(maxd, dtype=np.float64)
    keep["NP"] = np.array([sp.NP for sp in points], dtype=np.int32)
    keep["dz"] = np.array([sp.dz for sp in points], dtype=np.float64)
    keep["dmu"] = np.ascontiguousarray(np.stack([sp.ms.delta_muH for sp in points]).astype(np.float64))
    keep["mu"] = np.ascontiguousarray(np.stack([sp.mu_species for sp in points]).astype(np.float64))
    keep["qtot"] = np.array([sp.total_charge for sp in points], dtype=np.float64)
    keep["fcoord"] = np.ascontiguousarray(np.array([[int(sp.fixed[u].coords[0]) if sp.fixed[u].any else -1
                                                     for u in range(nsp)] for sp in points], dtype=np.int32))
    keep["Ext"] = np.ascontiguousarray(np.concatenate([sp.Ext.ravel(order="F") for sp in points]))
    keep["trapez"] = np.ascontiguousarray(np.concatenate([sp.trapez.ravel(order="F") for sp in points]))
    keep["surf"] = np.ascontiguousarray(np.concatenate([sp.surf[:sp.NP] for sp in points]))
    keep["lo"] = np.ascontiguousarray(np.stack([sp.wall_lo for sp in points]).astype(np.int32))
    keep["hi"] = np.ascontiguousarray(np.stack([sp.wall_hi for sp in points]).astype(np.int32))
    has_q = any(sp.wall_charges is not None for sp in points)
    keep["wq"] = np.ascontiguousarray(np.array([sp.wall_charges or (0.0, 0.0) for sp in points], dtype=np.float64))
    d = SdftDesc()
    d.abi_version = SDFT_ABI_VERSION
    d.n_points = P
    d.NB, d.MB, d.n_species = NB, MB, nsp
    d.functionals = functional_flags(p0.functionals)
    d.bjerrum = ms.bjerrum
    d.diameters = _ptr(keep["diam"], C.c_double)
    d.valences = _ptr(keep["val"], C.c_double)
    d.bond_types = _ptr(keep["bonds"], C.c_int32)
    d.species_nseg = _ptr(keep["nseg"], C.c_int32)
    d.species_model = _ptr(keep["model"], C.c_int32)
    d.species_nstates = _ptr(keep["nst"], C.c_int32)
    d.seg_states = _ptr(keep["states"], C.c_int32)
    d.species_fixed_seg = _ptr(keep["fseg"], C.c_int32)
    d.fixed_density = _ptr(keep["fden"], C.c_double)
    d.max_distance = _ptr(keep["maxd"], C.c_double)
    d.NP = _ptr(keep["NP"], C.c_int32)
    d.dz = _ptr(keep["dz"], C.c_double)
    d.delta_muH = _ptr(keep["dmu"], C.c_double)
    d.mu_species = _ptr(keep["mu"], C.c_double)
    d.total_charge = _ptr(keep["qtot"], C.c_double)
    d.fixed_coord = _ptr(keep["fcoord"], C.c_int32)
    d.Ext = _ptr(keep["Ext"], C.c_double)
    d.trapez = _ptr(keep["trapez"], C.c_double)
    d.surf = _ptr(keep["surf"], C.c_double)
    d.hard_walls = int(any(sp.hard_walls for sp in points))
    d.wall_lo = _ptr(keep["lo"], C.c_int32)
    d.wall_hi = _ptr(keep["hi"], C.c_int32)
    d.wall_charge = _ptr(keep["wq"], C.c_double) if has_q else None
    if any(sp.sw_wall_idx is not None for sp in points):
        idx = np.full((P, NB, 2), -1, dtype=np.int32)
        fac = np.zeros((P, NB, 2))
        for k, sp in enumerate(points):
            if sp.sw_wall_idx is not None:
                idx[k] = sp.sw_wall_idx
                fac[k] = sp.sw_wall_fac
        keep["swwi"] = np.ascontiguousarray(idx)
        keep["swwf"] = np.ascontiguousarray(fac)
        d.sw_wall_idx = _ptr(keep["swwi"], C.c_int32)
        d.sw_wall_fac = _ptr(keep["swwf"], C.c_double)
    if ms.sw is not None and "sw" in p0.functionals:
        keep["sw_pairs"] = np.ascontiguousarray(np.array(ms.sw.pairs, dtype=np.int32).reshape(-1))
        keep["sw_eps"] = np.ascontiguousarray(np.array(ms.sw.energys[:len(ms.sw.pairs)], dtype=np.float64))
        keep["sw_lam"] = np.ascontiguousarray(np.array(ms.sw.lambdas[:len(ms.sw.pairs)], dtype=np.float64))
        d.n_sw_pairs = len(ms.sw.pairs)
        d.sw_pairs = _ptr(keep["sw_pairs"], C.c_int32)
        d.sw_energy = _ptr(keep["sw_eps"], C.c_double)
        d.sw_lambda = _ptr(keep["sw_lam"], C.c_double)
    nds = {sp.ndim for sp in points}
    if len(nds) != 1:
        raise ValueError("a batch mixes one- and multi-dimensional grids")
    D = nds.pop()
    if D > 1:
        # multi-dimensional grids: per-point grid and bins; the contact values come from the host
        keep["grid"] = np.ascontiguousarray(np.array([list(sp.grid) + [1] * (3 - D) for sp in points], dtype=np.int32))
        keep["bins"] = np.ascontiguousarray(np.array([list(sp.bins) + [1.0] * (3 - D) for sp in points], dtype=np.float64))
        d.ndim = D
        d.grid = _ptr(keep["grid"], C.c_int32)
        d.bins = _ptr(keep["bins"], C.c_double)
        d.hard_walls = 0
        d.wall_charge = None
    return d, keep, nst, nseg


class Batch:
    """A library context over state points that share one molecular-system layout."""

    def __init__(self, points: List[StatePoint]):
        if not points:
            raise ValueError("empty batch")
        lib = load_library()
        self.lib = lib
        self.points = points
        d, keep, nst, nseg = build_descriptor(points)
        ms = points[0].ms
        self.NB, self.MB, self.nsp = ms.NB, ms.MB, len(ms.species)
        P = len(points)
        self._keep = keep
        self._desc = d
        h = C.c_void_p()
        rc = lib.sdft_create(C.byref(d), C.byref(h))
        if rc != SDFT_OK:
            raise SdftError(rc, lib.sdft_last_error(None).decode())
        self.h = h
        self.n_total = lib.sdft_solution_length(h)
        offs = np.zeros(P + 1, dtype=np.int64)
        lib.sdft_solution_offsets(h, _ptr(offs, C.c_int64))
        self.offsets = offs
        self.S_max = int(max(nst))
        self.species_states = [int(v) for v in nst]
        self.species_segments = [int(v) for v in nseg]

    # ------------------------------------------------------------------
    def _check(self, rc: int):
        if rc != SDFT_OK:
            raise SdftError(rc, self.lib.sdft_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def x0(self) -> np.ndarray:
        return np.concatenate([sp.x0() for sp in self.points])

    def split(self, x: np.ndarray) -> List[np.ndarray]:
        return [x[self.offsets[p]:self.offsets[p + 1]] for p in range(len(self.points))]

    def objective(self, x: np.ndarray, precision: int = PREC_FP64) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.n_total,):
            raise ValueError(f"x must have length {self.n_total}")
        fx = np.zeros_like(x)
        self._check(self.lib.sdft_objective(self.h, _ptr(x, C.c_double), _ptr(fx, C.c_double), precision))
        return fx

    def solve(self, x0: Optional[np.ndarray] = None, *, tole=1e-8, damping=0.1, mixing_max=0.02, switch_tole=100.0,
              anderm=5, max_iters=3000, precision=PREC_FP64, fp32_switch=1e-5, band_fp32=True, check_every=10,
              raise_on_noconv=True):
        x = np.ascontiguousarray(self.x0() if x0 is None else x0, dtype=np.float64).copy()
        P = len(self.points)
        iters = np.zeros(P, dtype=np.int32)
        err = np.zeros(P, dtype=np.float64)
        o = SdftSolverOpts(tole=tole, damping=damping, mixing_max=mixing_max, switch_tole=switch_tole,
                           anderm=anderm, max_iters=max_iters, precision=precision, fp32_switch=fp32_switch,
                           band_fp32=int(band_fp32), check_every=check_every)
        rc = self.lib.sdft_solve(self.h, C.byref(o), _ptr(x, C.c_double), _ptr(iters, C.c_int32), _ptr(err, C.c_double))
        if rc != SDFT_OK and (raise_on_noconv or rc != -4):
            self._check(rc)
        return x, iters, err

    def set_chemistry(self, delta_muH: np.ndarray, mu_species: np.ndarray):
        """Replace every point's delta_muH [P][NB] and mu_species [P][n_species] (and the chain
        normalisations that follow from them), e.g. after a pH step: the state of a continuation."""
        P = len(self.points)
        dmu = np.ascontiguousarray(delta_muH, dtype=np.float64).reshape(P, self.NB)
        mus = np.ascontiguousarray(mu_species, dtype=np.float64).reshape(P, self.nsp)
        self._check(self.lib.sdft_set_chemistry(self.h, _ptr(dmu, C.c_double), _ptr(mus, C.c_double)))

    def set_graft_density(self, density: np.ndarray):
        """Replace every point's grafting density per species [P][n_species]."""
        P = len(self.points)
        d = np.ascontiguousarray(density, dtype=np.float64).reshape(P, self.nsp)
        self._check(self.lib.sdft_set_graft_density(self.h, _ptr(d, C.c_double)))

    def last_timing(self):
        ms = C.c_double()
        it = C.c_double()
        self._check(self.lib.sdft_last_timing(self.h, C.byref(ms), C.byref(it)))
        return ms.value, it.value

    def set_operator_band(self, fp32: bool = True, fp64: bool = False):
        """Bond operator of objective() per precision: banded (True) or dense (False)."""
        self._check(self.lib.sdft_set_operator_band(self.h, int(fp32), int(fp64)))

    def autotune(self, reps: int = 3) -> List[dict]:
        """Measure the propagator mapping per chain species and precision and the convolution tile
        shape per grid-length bin and stage, keep the fastest; returns the decisions."""
        cap = 4 * 64 + 3 * 4096
        buf = (SdftTuneEntry * cap)()
        n = C.c_int32()
        self._check(self.lib.sdft_autotune(self.h, reps, buf, cap, C.byref(n)))
        return [dict(kind=e.kind, key=e.key, precision=e.precision, choice=e.choice, us=list(e.us))
                for e in buf[:min(n.value, cap)]]

    def set_tuning(self, entries: Sequence[dict]):
        """Apply decisions (from autotune() or forced): dicts with kind, key, precision, choice."""
        arr = (SdftTuneEntry * max(1, len(entries)))()
        for i, e in enumerate(entries):
            arr[i].kind = e["kind"]
            arr[i].key = e.get("key", 0)
            arr[i].precision = e.get("precision", 0)
            arr[i].choice = e["choice"]
        self._check(self.lib.sdft_set_tuning(self.h, arr, len(entries)))

    def set_conv_shape(self, choice: int):
        """Force one convolution path for every grid length and stage: an operator-GEMM tile shape
        (CONV_TALL ... CONV_NARROW) or the batched FFT (CONV_FFT, taken where the mirror length 2 NP
        is a power of two from 4096 to 65536; elsewhere the fixed rule applies); -1 restores the
        fixed rule."""
        self.set_tuning([dict(kind=k, key=0, choice=choice) for k in (TUNE_CONV_FWD, TUNE_CONV_PRE, TUNE_CONV_BACK)])

    def set_launch_mode(self, graph: Optional[bool] = None, concurrent: Optional[bool] = None) -> int:
        """Launch mode of solve(): graph windows (True, default) or kernel-by-kernel launches; FP32 and
        FP64 pipelines side by side (True, default) or in turn; None keeps a setting. Returns the
        graphs built so far."""
        n = C.c_int64()
        g = -1 if graph is None else int(bool(graph))
        cc = -1 if concurrent is None else int(bool(concurrent))
        self._check(self.lib.sdft_set_launch_mode(self.h, g, cc, C.byref(n)))
        return n.value

    def set_graph_capture(self, enable: Optional[bool] = None) -> int:
        """Graph windows on / off (None: unchanged); returns the graphs built so far."""
        return self.set_launch_mode(graph=enable)

    def switch_iters(self) -> np.ndarray:
        """Per point, the iterations done at the FP32 -> FP64 switch of the last mixed solve (0: none)."""
        out = np.zeros(len(self.points), dtype=np.int32)
        self._check(self.lib.sdft_last_switch_iters(self.h, _ptr(out, C.c_int32)))
        return out

    def pipeline_iters(self):
        """(FP32, FP64) batch iterations of the last solve during which each pipeline ran."""
        a = C.c_int32()
        b = C.c_int32()
        self._check(self.lib.sdft_last_pipeline_iters(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_profile(self, kernel_class: int, reps: int = 20):
        """(average ms per launch, algorithmic flops per launch) measured with HIP events on the
        context's stream."""
        ms = C.c_double()
        fl = C.c_double()
        self._check(self.lib.sdft_kernel_profile(self.h, kernel_class, reps, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def observables(self):
        P = len(self.points)
        om = np.zeros(P)
        al = np.zeros(P * self.nsp * self.S_max)
        pc = np.zeros(P * 2)
        self._check(self.lib.sdft_observables(self.h, _ptr(om, C.c_double), _ptr(al, C.c_double), _ptr(pc, C.c_double)))
        return om, al.reshape(P, self.nsp, self.S_max), pc.reshape(P, 2)

    def observables_noalloc(self):
        """observables() through sdft_observables_noalloc: no device allocation, partial sums in the
        library's static per-GPU scratch array."""
        P = len(self.points)
        om = np.zeros(P)
        al = np.zeros(P * self.nsp * self.S_max)
        pc = np.zeros(P * 2)
        self._check(self.lib.sdft_observables_noalloc(self.h, _ptr(om, C.c_double), _ptr(al, C.c_double),
                                                      _ptr(pc, C.c_double)))
        return om, al.reshape(P, self.nsp, self.S_max), pc.reshape(P, 2)

    def get_field(self, field_id: int, point: int) -> np.ndarray:
        sp = self.points[point]
        n = {FIELD_RHO_BEADS: sp.NP * self.NB, FIELD_RHO_BONDS: sp.NP * 2 * self.MB,
             FIELD_X: sp.n, FIELD_FX: sp.n}[field_id]
        out = np.zeros(n)
        self._check(self.lib.sdft_get_field(self.h, field_id, point, _ptr(out, C.c_double), n))
        if field_id == FIELD_RHO_BEADS:
            return out.reshape((sp.NP, self.NB), order="F")
        if field_id == FIELD_RHO_BONDS:
            return out.reshape((sp.NP, 2, self.MB), order="F")
        return out

    def segment_profile(self, species: int, point: int) -> np.ndarray:
        """Segment densities of one species at one point, (NP, S_u, N_u) with S_u the species'
        states per segment (rho_K.segments, accumulate_densities.jl:95-184)."""
        if not 0 <= species < self.nsp:
            raise ValueError(f"species index {species} out of range")
        sp = self.points[point]
        S, N = self.species_states[species], self.species_segments[species]
        out = np.zeros(sp.NP * S * N)
        self._check(self.lib.sdft_get_field(self.h, FIELD_RHO_SEGMENTS + species, point,
                                            _ptr(out, C.c_double), out.size))
        return out.reshape((sp.NP, S, N), order="F")
