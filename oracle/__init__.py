"""FP64 CPU oracle for the batched gas-phase kinetics path.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  The product library never
imports, links or runs anything under ``oracle/``.

``kmd.load`` restates the reference's YAML loader in Python; ``oracle.c``
restates its numerics in C (see the citations there).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import amd as amd_mod
from . import kmd as kmd_mod

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class _KMD(ctypes.Structure):
    _fields_ = [
        ("nSpec", ctypes.c_int), ("nConstSpec", ctypes.c_int), ("nReac", ctypes.c_int),
        ("maxSpecInReac", ctypes.c_int),
        ("reacNreac", ctypes.c_void_p), ("reacNprod", ctypes.c_void_p), ("reacSidx", ctypes.c_void_p),
        ("reacNuki", ctypes.c_void_p), ("convExponent", ctypes.c_void_p),
        ("nArr", ctypes.c_int), ("arrIdx", ctypes.c_void_p), ("arrPar", ctypes.c_void_p),
        ("nTroe", ctypes.c_int), ("troeIdx", ctypes.c_void_p), ("troePar", ctypes.c_void_p),
        ("nCmaq", ctypes.c_int), ("cmaqIdx", ctypes.c_void_p), ("cmaqPar", ctypes.c_void_p),
        ("nEmis", ctypes.c_int), ("emisSpecies", ctypes.c_void_p), ("emisRate", ctypes.c_void_p),
        ("nPhoto", ctypes.c_int), ("photoReac", ctypes.c_void_p),
        ("CONV_PPM", ctypes.c_double), ("lu_order", ctypes.c_void_p),
    ]


class _AMD(ctypes.Structure):
    _fields_ = [
        ("nSpec", ctypes.c_int), ("nParticles", ctypes.c_int), ("nSimpol", ctypes.c_int),
        ("mw", ctypes.c_void_p), ("density", ctypes.c_void_p), ("gas", ctypes.c_void_p),
        ("aero", ctypes.c_void_p), ("par", ctypes.c_void_p),
    ]


class TimeAdvance(ctypes.Structure):
    """Byte layout of TChem::TimeAdvance (src/core/TChem_Util.hpp:469-476)."""
    _fields_ = [
        ("tbeg", ctypes.c_double), ("tend", ctypes.c_double), ("dt", ctypes.c_double),
        ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double),
        ("max_num_newton_iterations", ctypes.c_int), ("num_time_iterations_per_interval", ctypes.c_int),
        ("num_outer_time_iterations_per_interval", ctypes.c_int), ("jacobian_interval", ctypes.c_int),
    ]


TADV_DTYPE = np.dtype([
    ("tbeg", "f8"), ("tend", "f8"), ("dt", "f8"), ("dtmin", "f8"), ("dtmax", "f8"),
    ("max_num_newton_iterations", "i4"), ("num_time_iterations_per_interval", "i4"),
    ("num_outer_time_iterations_per_interval", "i4"), ("jacobian_interval", "i4"),
])
STATS_DTYPE = np.dtype([("accepted", "i4"), ("rejected", "i4"), ("newton_failures", "i4"),
                        ("newton_iterations", "i4"), ("factorizations", "i4"), ("failed", "i4")])


def build(force: bool = False) -> str:
    """Compile oracle.c with the recipe in oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        _lib.oracle_kforward.argtypes = [vp, d, d, vp, vp]
        _lib.oracle_omega.argtypes = [vp, vp, vp, vp]
        _lib.oracle_jacobian.argtypes = [vp, vp, vp, vp]
        _lib.oracle_lu.argtypes = [i, vp, i, vp]
        _lib.oracle_lu.restype = i
        _lib.oracle_lu_solve.argtypes = [i, vp, i, vp, vp]
        _lib.oracle_sourceterm_batch.argtypes = [vp, i, vp, vp, vp]
        _lib.oracle_jacobian_batch.argtypes = [vp, i, vp, vp, vp]
        _lib.oracle_run_batch.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, vp, i, i]
        _lib.oracle_trbdf2_constants.argtypes = [vp]
        _lib.oracle_record_steps.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i]
        _lib.oracle_record_steps.restype = i
        _lib.oracle_tlm_fixed_steps.argtypes = [vp, vp, i, i, vp, vp, vp, vp, i, i, vp, vp, vp]
        _lib.oracle_tlm_fixed_steps.restype = i
        _lib.oracle_simpol_constants.argtypes = [vp, d, d, vp]
        _lib.oracle_aerosol_omega.argtypes = [vp, vp, d, d, vp, vp, vp, vp, vp]
        _lib.oracle_aerosol_jacobian.argtypes = [vp, vp, d, d, vp, vp, vp, vp, vp]
        _lib.oracle_aerosol_rhs_batch.argtypes = [vp, vp, i, vp, vp, vp]
        _lib.oracle_aerosol_run_batch.argtypes = [vp, vp, i, vp, vp, vp, vp, vp, vp, vp, vp, i]
        _lib.oracle_detmath_batch.argtypes = [i, i, vp, vp, vp]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """Binds one mechanism's tables to the C oracle."""

    def __init__(self, mech, lu_order=True):
        """lu_order: True = factor the Newton systems in the fill-reducing
        elimination order (kmd.elimination_order), False = species order, or
        an explicit order [m]."""
        self.k = mech if isinstance(mech, kmd_mod.OracleKMD) else kmd_mod.load(mech)
        k = self.k
        if lu_order is True:
            self.lu_order = kmd_mod.elimination_order(k)
        elif lu_order is False or lu_order is None:
            self.lu_order = None
        else:
            self.lu_order = np.ascontiguousarray(lu_order, np.int32)
        self._keep = [k.reacNreac, k.reacNprod, k.reacSidx, k.reacNuki, k.convExponent, k.arrIdx, k.arrPar,
                      k.troeIdx, k.troePar, k.cmaqIdx, k.cmaqPar, k.emisSpecies, k.emisRate, k.photoReac,
                      self.lu_order]
        self.s = _KMD(k.nSpec, k.nConstSpec, k.nReac, k.maxSpecInReac,
                      _p(k.reacNreac), _p(k.reacNprod), _p(k.reacSidx), _p(k.reacNuki), _p(k.convExponent),
                      len(k.arrIdx), _p(k.arrIdx), _p(k.arrPar),
                      len(k.troeIdx), _p(k.troeIdx), _p(k.troePar),
                      len(k.cmaqIdx), _p(k.cmaqIdx), _p(k.cmaqPar),
                      len(k.emisSpecies), _p(k.emisSpecies), _p(k.emisRate),
                      len(k.photoReac), _p(k.photoReac), kmd_mod.CONV_PPM, _p(self.lu_order))
        self.ptr = ctypes.addressof(self.s)

    @property
    def m(self):
        return self.k.m

    def kforward(self, T, p, photo=None):
        kfor = np.zeros(self.k.nReac)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        lib().oracle_kforward(self.ptr, float(T), float(p), _p(ph), _p(kfor))
        return kfor

    def omega(self, kfor, c):
        c = np.ascontiguousarray(c, np.float64)
        out = np.zeros(self.m)
        lib().oracle_omega(self.ptr, _p(np.ascontiguousarray(kfor)), _p(c), _p(out))
        return out

    def jacobian(self, kfor, c):
        c = np.ascontiguousarray(c, np.float64)
        out = np.zeros((self.m, self.m))
        lib().oracle_jacobian(self.ptr, _p(np.ascontiguousarray(kfor)), _p(c), _p(out))
        return out

    def sourceterm_batch(self, state, photo=None):
        state = np.ascontiguousarray(state, np.float64)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        out = np.zeros((state.shape[0], self.m))
        lib().oracle_sourceterm_batch(self.ptr, state.shape[0], _p(state), _p(ph), _p(out))
        return out

    def jacobian_batch(self, state, photo=None):
        state = np.ascontiguousarray(state, np.float64)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        out = np.zeros((state.shape[0], self.m, self.m))
        lib().oracle_jacobian_batch(self.ptr, state.shape[0], _p(state), _p(ph), _p(out))
        return out

    def run_batch(self, tol_newton, tol_time, tadv, state, t_out, dt_out, photo=None, stats=None,
                  jacobian_mode=0, nthreads=0):
        """One runDeviceBatch interval; state/t_out/dt_out updated in place."""
        assert state.dtype == np.float64 and state.flags.c_contiguous
        assert tadv.dtype == TADV_DTYPE
        tn = np.ascontiguousarray(tol_newton, np.float64)
        tt = np.ascontiguousarray(tol_time, np.float64)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        lib().oracle_run_batch(self.ptr, state.shape[0], _p(tn), _p(tt), _p(tadv), _p(state), _p(ph),
                               _p(t_out), _p(dt_out), _p(stats), int(jacobian_mode), int(nthreads))

    def record_steps(self, tol_newton, tol_time, ta, row, photo=None, lscale=None, max_steps=4096):
        """One cell (analytic Jacobian) with rate constants scaled by exp(lscale);
        returns (final row, accepted step sizes) or (row, None) on failure."""
        assert ta.dtype == TADV_DTYPE and ta.shape == (1,)
        row = np.array(row, np.float64)
        h = np.zeros(max_steps)
        tn = np.ascontiguousarray(tol_newton, np.float64)
        tt = np.ascontiguousarray(tol_time, np.float64)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        ls = None if lscale is None else np.ascontiguousarray(lscale, np.float64)
        n = lib().oracle_record_steps(self.ptr, _p(tn), _p(tt), _p(ta), _p(row), _p(ph), _p(ls), _p(h), max_steps)
        return row, (h[:n].copy() if n >= 0 else None)


    def tlm_fixed_steps(self, tol_newton, max_newton, jint, row, h, sens=(), photo=None, lscale=None):
        """Fixed step sequence h with the discrete tangent linear of the
        parameters in sens (>= 0: reaction log-scale, < 0: initial species -s-1).
        Returns (u_final [m], S [m][len(sens)]); raises on a Newton failure."""
        row = np.ascontiguousarray(row, np.float64)
        h = np.ascontiguousarray(h, np.float64)
        sens = np.ascontiguousarray(sens, np.int32)
        u = np.zeros(self.m)
        S = np.zeros((self.m, max(len(sens), 1)))
        tn = np.ascontiguousarray(tol_newton, np.float64)
        ph = None if photo is None else np.ascontiguousarray(photo, np.float64)
        ls = None if lscale is None else np.ascontiguousarray(lscale, np.float64)
        rc = lib().oracle_tlm_fixed_steps(self.ptr, _p(tn), int(max_newton), int(jint), _p(row), _p(ph), _p(ls), _p(h),
                                          len(h), len(sens), _p(sens), _p(u), _p(S))
        if rc != 0:
            raise RuntimeError("oracle_tlm_fixed_steps: Newton failure")
        return u, S[:, :len(sens)]


class AerosolOracle:
    """Binds a gas mechanism and an aerosol file to the C oracle's
    single-particle SIMPOL.1 functions."""

    def __init__(self, chem, aero):
        self.g = chem if isinstance(chem, Oracle) else Oracle(chem)
        self.a = aero if isinstance(aero, amd_mod.OracleAMD) else amd_mod.load(aero, list(self.g.k.names))
        a = self.a
        self._keep = [a.mw, a.density, a.gas, a.aero, a.par]
        self.s = _AMD(a.nSpec, a.nParticles, a.nSimpol, _p(a.mw), _p(a.density), _p(a.gas), _p(a.aero),
                      _p(a.par))
        self.ptr = ctypes.addressof(self.s)

    @property
    def mg(self):
        return self.g.m

    @property
    def mt(self):
        return self.g.m + self.a.nSpec * self.a.nParticles

    @property
    def dim(self):
        return 3 + self.g.k.nSpec + self.a.nSpec * self.a.nParticles

    def simpol_constants(self, s, T, p):
        out = np.zeros(4)
        lib().oracle_simpol_constants(_p(np.ascontiguousarray(self.a.par[s])), float(T), float(p), _p(out))
        return out

    def omega(self, T, p, x, cconst, nc):
        kfor = self.g.kforward(T, p)
        out = np.zeros(self.mt)
        lib().oracle_aerosol_omega(self.g.ptr, self.ptr, float(T), float(p), _p(kfor),
                                   _p(np.ascontiguousarray(cconst, np.float64)),
                                   _p(np.ascontiguousarray(nc, np.float64)),
                                   _p(np.ascontiguousarray(x, np.float64)), _p(out))
        return out

    def jacobian(self, T, p, x, cconst, nc):
        kfor = self.g.kforward(T, p)
        out = np.zeros((self.mt, self.mt))
        lib().oracle_aerosol_jacobian(self.g.ptr, self.ptr, float(T), float(p), _p(kfor),
                                      _p(np.ascontiguousarray(cconst, np.float64)),
                                      _p(np.ascontiguousarray(nc, np.float64)),
                                      _p(np.ascontiguousarray(x, np.float64)), _p(out))
        return out

    def rhs_batch(self, state, nc):
        state = np.ascontiguousarray(state, np.float64)
        nc = np.ascontiguousarray(nc, np.float64)
        out = np.zeros((state.shape[0], self.mt))
        lib().oracle_aerosol_rhs_batch(self.g.ptr, self.ptr, state.shape[0], _p(state), _p(nc), _p(out))
        return out

    def run_batch(self, tol_newton, tol_time, tadv, state, nc, t_out, dt_out, stats=None, nthreads=0):
        """One AerosolChemistry::runDeviceBatch interval, in place."""
        assert state.dtype == np.float64 and state.flags.c_contiguous and state.shape[1] == self.dim
        assert tadv.dtype == TADV_DTYPE
        nc = np.ascontiguousarray(nc, np.float64)
        lib().oracle_aerosol_run_batch(self.g.ptr, self.ptr, state.shape[0],
                                       _p(np.ascontiguousarray(tol_newton, np.float64)),
                                       _p(np.ascontiguousarray(tol_time, np.float64)), _p(tadv), _p(state),
                                       _p(nc), _p(t_out), _p(dt_out), _p(stats), int(nthreads))


def detmath(fn: str, x, y=None):
    """The portable exp / log / log10 / pow the rate constants use, on arrays."""
    code = {"exp": 0, "log": 1, "log10": 2, "pow": 3}[fn]
    x = np.ascontiguousarray(x, np.float64).ravel()
    yy = np.ascontiguousarray(np.broadcast_to(0.0 if y is None else y, x.shape), np.float64).ravel()
    out = np.zeros_like(x)
    lib().oracle_detmath_batch(code, x.size, _p(x), _p(yy), _p(out))
    return out


def lu(A):
    A = np.array(A, np.float64, order="C")
    n = A.shape[0]
    piv = np.zeros(n, np.int32)
    info = lib().oracle_lu(n, _p(A), n, _p(piv))
    return A, piv, info


def lu_solve(LU, piv, b):
    b = np.array(b, np.float64)
    lib().oracle_lu_solve(LU.shape[0], _p(np.ascontiguousarray(LU)), LU.shape[0], _p(piv), _p(b))
    return b


def trbdf2_constants():
    out = np.zeros(6)
    lib().oracle_trbdf2_constants(_p(out))
    return out


def run_sliced(o: Oracle, workload, state, photo, n_slices=1, jacobian_mode=0, nthreads=0, ends=None,
               checkpoints=None):
    """The workload's hour [tbeg, tend] on the given rows, cut into n_slices
    equal runDeviceBatch intervals, or intervals ending at `ends` (each run
    to its end; t_out / dt_out carried into the next interval's
    TimeAdvance).  Returns (state, t_out, dt_out, stats summed over the
    slices; failed = failed in any); with a list as `checkpoints`, the state
    after every slice is appended to it."""
    w = workload
    st = np.ascontiguousarray(state, np.float64).copy()
    ph = None if photo is None or photo.shape[1] == 0 else np.ascontiguousarray(photo, np.float64)
    n = st.shape[0]
    t = np.full(n, w.tbeg)
    dt = np.full(n, w.dt_in)
    tn = np.array([w.atol_newton, w.rtol_newton])
    tt = np.tile([w.atol_time, w.rtol_time], (o.m, 1))
    total = np.zeros(n, STATS_DTYPE)
    if ends is None:
        ends = [w.tend if k == n_slices - 1 else w.tbeg + (w.tend - w.tbeg) * (k + 1) / n_slices
                for k in range(n_slices)]
    for tend in ends:
        ta = np.zeros(n, TADV_DTYPE)
        ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = t, tend, dt, w.dtmin, w.dtmax
        ta["max_num_newton_iterations"] = w.max_newton
        ta["num_time_iterations_per_interval"] = 10_000_000
        ta["jacobian_interval"] = w.jacobian_interval
        s = np.zeros(n, STATS_DTYPE)
        o.run_batch(tn, tt, ta, st, t, dt, photo=ph, stats=s, jacobian_mode=jacobian_mode, nthreads=nthreads)
        for f in ("accepted", "rejected", "newton_failures", "newton_iterations", "factorizations"):
            total[f] += s[f]
        total["failed"] = np.maximum(total["failed"], s["failed"])
        if checkpoints is not None:
            checkpoints.append(st.copy())
    return st, t, dt, total
