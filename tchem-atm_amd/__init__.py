"""Host-side mirror of TChem-atm's batched gas-phase kinetics interface.

Python classes named after the reference's C++ types, calling the C ABI of
``libtchem_atm.so`` (declared in include/tchem_atm.h) through ctypes:

  KineticModelData(mechfile)                   src/core/TChem_KineticModelData.cpp:53-62
  KineticModelData.createConstData()           createNCAR_KineticModelConstData, ...ConstData.hpp:115-148
  setScenarioConditions(file, kmd)             src/core/TChem_AtmosphericChemistry.cpp:429-504
  AtmosphericChemistry.runDeviceBatch(...)     src/core/TChem_AtmosphericChemistry.hpp:83-97
  AtmosphericChemistry.getWorkSpaceSize(kmcd)  src/core/TChem_AtmosphericChemistry.hpp:47-81
  SourceTerm.runDeviceBatch(...)               src/examples/TChem_GasChemistry_RHSs.cpp:184-236
  Jacobian.runDeviceBatch(...)                 src/examples/TChem_GasChemistry_RHSs.cpp:254-309
  ReactionRates / RateofProgress / NetProductionRates.runDeviceBatch
                                               src/core/TChem_{ReactionRates,RateofProgress,NetProductionRates}.hpp
  setScenarioConditionsPhotolysisReactions / ExternalForcing
                                               src/core/TChem_AtmosphericChemistry.cpp:511-631

Device arrays are torch tensors on a HIP device (float64, contiguous); torch
provides device memory and streams only.  Every compute call goes to the HIP
kernels: there is no CPU fallback, and a missing library raises at import.
Errors raise TChemError with the library's message (the reference throws
std::logic_error from TCHEM_CHECK_ERROR, src/core/TChem_Util.hpp:108-114).

The package directory name contains a dash; import it with
``importlib.import_module("tchem-atm_amd")``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtchem_atm.so")

PRECISION_FP64 = 0
PRECISION_MIXED = 1
SCHEDULE_GRID = 0
SCHEDULE_PERSISTENT = 1
LU_AUTO = 0     # sparse when the mechanism has a sparse schedule, dense otherwise
LU_DENSE = 1    # dense LU with partial pivoting (the oracle's arithmetic)
LU_SPARSE = 2   # level-scheduled sparse LU, pivots fixed by the fill-reducing order
LU_MFMA = 3     # dense FP32 LU on the matrix cores (mixed precision, m <= 255), one cell per 4 wavefronts
PROGRAMS_AUTO = 0    # sparse form: programs in LDS when they leave room for 4 cells
PROGRAMS_LDS = 1     # staged into LDS once per workgroup
PROGRAMS_MEMORY = 2  # read from device memory (L2-resident), all LDS for cells

TIME_ADVANCE_DTYPE = np.dtype([
    ("_tbeg", "f8"), ("_tend", "f8"), ("_dt", "f8"), ("_dtmin", "f8"), ("_dtmax", "f8"),
    ("_max_num_newton_iterations", "i4"), ("_num_time_iterations_per_interval", "i4"),
    ("_num_outer_time_iterations_per_interval", "i4"), ("_jacobian_interval", "i4"),
])
assert TIME_ADVANCE_DTYPE.itemsize == 56
CELL_STATS_DTYPE = np.dtype([("accepted", "i4"), ("rejected", "i4"), ("newton_failures", "i4"),
                             ("newton_iterations", "i4"), ("factorizations", "i4"), ("failed", "i4"),
                             ("t_start", "u8"), ("t_end", "u8")])
assert CELL_STATS_DTYPE.itemsize == 40
COUNTERS = ("accepted", "rejected", "newton_failures", "newton_iterations", "factorizations", "failed")
REALTIME_HZ = 100e6  # device real-time counter of t_start / t_end

# every symbol include/tchem_atm.h declares
EXPORTED_SYMBOLS = (
    "tchem_last_error", "tchem_kmd_load_yaml", "tchem_kmd_destroy", "tchem_kmd_info", "tchem_kmd_species_name",
    "tchem_kmd_photolysis_reactions", "tchem_kmd_elimination_order", "tchem_scenario_load", "tchem_kmcd_create", "tchem_kmcd_destroy",
    "tchem_atmchem_workspace_bytes", "tchem_atmchem_run_device_batch", "tchem_atmchem_run_device_batch_noalloc", "tchem_dense_lu_solve_mfma",
    "tchem_atmchem_update_time_advance",
    "tchem_sourceterm_run_device_batch", "tchem_jacobian_run_device_batch", "tchem_adjoint_workspace_bytes",
    "tchem_adjoint_forward", "tchem_adjoint_backward", "tchem_adjoint_steps", "tchem_atmchem_graph_create",
    "tchem_graph_launch", "tchem_graph_status", "tchem_graph_destroy", "tchem_atmchem_autotune",
    "tchem_reactionrates_run_device_batch", "tchem_rateofprogress_run_device_batch",
    "tchem_netproductionrates_run_device_batch", "tchem_scenario_photolysis_rates", "tchem_scenario_external_forcing",
    "tchem_atmchem_e3sm_run_device_batch", "tchem_amd_load_yaml", "tchem_amd_destroy", "tchem_amd_info",
    "tchem_amd_set_number_of_particles", "tchem_amd_species_name", "tchem_amd_species_props",
    "tchem_scenario_particles", "tchem_amcd_create", "tchem_amcd_destroy", "tchem_aerosol_rhs_run_device_batch",
    "tchem_aerosolchem_run_device_batch",
)


class TChemError(RuntimeError):
    pass


class _Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("schedule", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("lu", ctypes.c_int), ("cells_per_block", ctypes.c_int), ("programs", ctypes.c_int),
                ("reserved", ctypes.c_int * 2)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C tchem-atm_amd/csrc)")
    # torch brings its own HIP runtime under another file name; loaded first,
    # it is the one the library binds to.  Loaded after the library, a second
    # runtime would share the process with the library's, and the one that
    # initialises second finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
    lib.tchem_last_error.restype = cp
    sig = {
        "tchem_kmd_load_yaml": [cp, cp, ctypes.POINTER(vp)],
        "tchem_kmd_destroy": [vp],
        "tchem_kmd_info": [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)],
        "tchem_kmd_species_name": [vp, i, cp, ctypes.c_size_t],
        "tchem_kmd_photolysis_reactions": [vp, vp, i],
        "tchem_kmd_elimination_order": [vp, vp, i],
        "tchem_scenario_load": [vp, cp, vp, i, ctypes.POINTER(i)],
        "tchem_kmcd_create": [vp, ctypes.POINTER(vp)],
        "tchem_kmcd_destroy": [vp],
        "tchem_atmchem_workspace_bytes": [vp, ctypes.POINTER(_Options), ctypes.POINTER(ctypes.c_size_t)],
        "tchem_atmchem_run_device_batch": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                           vp],
        "tchem_atmchem_run_device_batch_noalloc": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                                   vp, vp, vp],
        "tchem_atmchem_update_time_advance": [i, vp, vp, vp, vp],
        "tchem_dense_lu_solve_mfma": [i, i, vp, vp, vp, vp],
        "tchem_sourceterm_run_device_batch": [i, vp, vp, vp, vp, vp, vp],
        "tchem_jacobian_run_device_batch": [i, vp, vp, vp, vp, vp],
        "tchem_adjoint_workspace_bytes": [vp, i, i, ctypes.POINTER(ctypes.c_size_t)],
        "tchem_adjoint_forward": [i, vp, vp, vp, vp, vp, vp, i, vp, vp, vp, vp, i, vp, vp],
        "tchem_adjoint_backward": [i, vp, vp, vp, vp, i, vp, vp],
        "tchem_adjoint_steps": [i, vp, i, vp, vp, vp, vp],
        "tchem_atmchem_graph_create": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, vp, vp, vp, vp, i,
                                       ctypes.c_double, vp, ctypes.POINTER(vp)],
        "tchem_graph_launch": [vp, vp],
        "tchem_graph_status": [vp, ctypes.POINTER(i), ctypes.POINTER(i)],
        "tchem_graph_destroy": [vp],
        "tchem_atmchem_autotune": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, i, i,
                                   ctypes.POINTER(ctypes.c_double), vp, vp],
        "tchem_reactionrates_run_device_batch": [i, vp, vp, vp, vp],
        "tchem_rateofprogress_run_device_batch": [i, vp, vp, i, vp, vp, vp],
        "tchem_netproductionrates_run_device_batch": [i, vp, vp, i, vp, vp, i, vp, vp],
        "tchem_scenario_photolysis_rates": [vp, cp, i, vp, i, ctypes.POINTER(i)],
        "tchem_scenario_external_forcing": [vp, cp, i, vp, ctypes.POINTER(i)],
        "tchem_atmchem_e3sm_run_device_batch": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, vp, i, vp, vp, vp,
                                                vp, vp, vp, vp],
        "tchem_amd_load_yaml": [cp, vp, ctypes.POINTER(vp)],
        "tchem_amd_destroy": [vp],
        "tchem_amd_info": [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)],
        "tchem_amd_set_number_of_particles": [vp, i],
        "tchem_amd_species_name": [vp, i, cp, ctypes.c_size_t],
        "tchem_amd_species_props": [vp, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)],
        "tchem_scenario_particles": [vp, cp, i, i, vp, vp, ctypes.POINTER(i)],
        "tchem_amcd_create": [vp, vp, ctypes.POINTER(vp)],
        "tchem_amcd_destroy": [vp],
        "tchem_aerosol_rhs_run_device_batch": [i, vp, vp, vp, vp, vp],
        "tchem_aerosolchem_run_device_batch": [ctypes.POINTER(_Options), i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                               vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = i
    return lib


_lib = _load()


def lib():
    return _lib


def _check(rc):
    if rc != 0:
        raise TChemError(f"tchem error {rc}: {_lib.tchem_last_error().decode()}")


def _ptr(t, dtype_name="float64", what="array"):
    """Device pointer of a contiguous torch tensor (None passes NULL)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not t.is_cuda:
        raise TChemError(f"{what} must be a device tensor")
    if not t.is_contiguous():
        raise TChemError(f"{what} must be contiguous")
    if dtype_name and str(t.dtype) != f"torch.{dtype_name}":
        raise TChemError(f"{what} must be {dtype_name}, got {t.dtype}")
    return t.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream if isinstance(stream, int) else stream.cuda_stream


class KineticModelData:
    """NCAR-format mechanism loaded on the host (KineticModelData)."""

    def __init__(self, mechfile: str, echofile: str | None = None):
        h = ctypes.c_void_p()
        _check(_lib.tchem_kmd_load_yaml(mechfile.encode(), echofile.encode() if echofile else None,
                                        ctypes.byref(h)))
        self._h = h
        n = [ctypes.c_int() for _ in range(4)]
        _check(_lib.tchem_kmd_info(h, *[ctypes.byref(x) for x in n]))
        self.nSpec, self.nConstSpec, self.nReac, self.nPhoto = (x.value for x in n)
        buf = ctypes.create_string_buffer(64)
        self.speciesNames = []
        for k in range(self.nSpec):
            _check(_lib.tchem_kmd_species_name(h, k, buf, 64))
            self.speciesNames.append(buf.value.decode())
        pr = np.zeros(max(self.nPhoto, 1), np.int32)
        _check(_lib.tchem_kmd_photolysis_reactions(h, pr.ctypes.data, self.nPhoto))
        self.photolysisReactions = pr[: self.nPhoto].copy()

    @property
    def nActive(self) -> int:
        return self.nSpec - self.nConstSpec

    def eliminationOrder(self) -> np.ndarray:
        """Order in which the device LU forms eliminate the active species
        (position -> species index); computed on the host."""
        out = np.zeros(max(self.nActive, 1), np.int32)
        _check(_lib.tchem_kmd_elimination_order(self._h, out.ctypes.data, out.size))
        return out[: self.nActive].copy()

    def createConstData(self) -> "KineticModelConstData":
        return KineticModelConstData(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tchem_kmd_destroy(h)
            self._h = None


class KineticModelConstData:
    """Device-resident compiled mechanism (KineticModelNCAR_ConstData)."""

    def __init__(self, kmd: KineticModelData):
        h = ctypes.c_void_p()
        _check(_lib.tchem_kmcd_create(kmd._h, ctypes.byref(h)))
        self._h = h
        self.nSpec, self.nConstSpec, self.nReac, self.nPhoto = kmd.nSpec, kmd.nConstSpec, kmd.nReac, kmd.nPhoto

    @property
    def nActive(self) -> int:
        return self.nSpec - self.nConstSpec

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tchem_kmcd_destroy(h)
            self._h = None


def createNCAR_KineticModelConstData(kmd: KineticModelData) -> KineticModelConstData:
    return KineticModelConstData(kmd)


def setScenarioConditions(filename: str, kmd: KineticModelData) -> np.ndarray:
    """Host state rows [nBatch][nSpec+3] from a scenario YAML."""
    n = ctypes.c_int()
    _check(_lib.tchem_scenario_load(kmd._h, filename.encode(), None, 0, ctypes.byref(n)))
    state = np.zeros((n.value, kmd.nSpec + 3))
    _check(_lib.tchem_scenario_load(kmd._h, filename.encode(), state.ctypes.data, n.value, ctypes.byref(n)))
    return state


class AtmosphericChemistry:
    @staticmethod
    def getWorkSpaceSize(kmcd: KineticModelConstData, precision: int = PRECISION_FP64, lu: int = LU_AUTO,
                         cells_per_block: int = 0, programs: int = PROGRAMS_AUTO) -> int:
        """Bytes of LDS one workgroup uses in the given form."""
        b = ctypes.c_size_t()
        opts = _Options(precision, SCHEDULE_GRID, 0, lu, cells_per_block, programs)
        _check(_lib.tchem_atmchem_workspace_bytes(kmcd._h, ctypes.byref(opts), ctypes.byref(b)))
        return b.value

    @staticmethod
    def runDeviceBatch(tol_newton, tol_time, fac, tadv, state, t_out, dt_out, state_out, kmcd, photo_rates=None,
                       stats=None, precision=PRECISION_FP64, schedule=SCHEDULE_GRID, blocks_per_cu=0, lu=LU_AUTO,
                       cells_per_block=0, programs=PROGRAMS_AUTO, stream=None):
        """One integration interval over the batch (asynchronous on `stream`).

        tadv is a uint8 device tensor holding TIME_ADVANCE_DTYPE records (or any
        tensor with 56 bytes per cell); stats likewise CELL_STATS_DTYPE."""
        nBatch = state.shape[0]
        opts = _Options(precision, schedule, blocks_per_cu, lu, cells_per_block, programs)
        _check(_lib.tchem_atmchem_run_device_batch(
            ctypes.byref(opts), nBatch, _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(fac, None, "fac"), _ptr(tadv, None, "tadv"), _ptr(state, what="state"),
            _ptr(photo_rates, what="photo_rates"), _ptr(t_out, what="t_out"), _ptr(dt_out, what="dt_out"),
            _ptr(state_out, what="state_out"), _ptr(stats, None, "stats"), kmcd._h, _stream_ptr(stream)))

    @staticmethod
    def captureGraph(tol_newton, tol_time, tadv, state, t_out, dt_out, state_out, kmcd, n_intervals,
                     stop_fraction=0.9999, photo_rates=None, stats=None, precision=PRECISION_FP64,
                     schedule=SCHEDULE_GRID, blocks_per_cu=0, lu=LU_AUTO, cells_per_block=0,
                     programs=PROGRAMS_AUTO) -> "Graph":
        """The interval loop (runDeviceBatch + updateTimeAdvance) x n_intervals
        captured into a HIP graph bound to these arrays; with stop_fraction > 0
        the driver's stop test (mean t > stop_fraction * tend) runs on the
        device after each interval."""
        opts = _Options(precision, schedule, blocks_per_cu, lu, cells_per_block, programs)
        h = ctypes.c_void_p()
        _check(_lib.tchem_atmchem_graph_create(
            ctypes.byref(opts), state.shape[0], _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(tadv, None, "tadv"), _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"),
            _ptr(t_out, what="t_out"), _ptr(dt_out, what="dt_out"), _ptr(state_out, what="state_out"),
            _ptr(stats, None, "stats"), int(n_intervals), float(stop_fraction), kmcd._h, ctypes.byref(h)))
        return Graph(h, (tol_newton, tol_time, tadv, state, photo_rates, t_out, dt_out, state_out, stats, kmcd))

    @staticmethod
    def autotune(tol_newton, tol_time, tadv, state, kmcd, photo_rates=None, precision=PRECISION_MIXED,
                 schedule=SCHEDULE_PERSISTENT, lu=LU_AUTO, sample_cells=0, sample_iterations=0, stream=None):
        """Fastest (lu, cells_per_block, blocks_per_cu) on a sample of this
        batch; returns a dict of runDeviceBatch keyword arguments plus
        'seconds' (the sample's best time)."""
        opts = _Options(precision, schedule, 0, lu, 0)
        best = ctypes.c_double()
        _check(_lib.tchem_atmchem_autotune(
            ctypes.byref(opts), state.shape[0], _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(tadv, None, "tadv"), _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"),
            int(sample_cells), int(sample_iterations), ctypes.byref(best), kmcd._h, _stream_ptr(stream)))
        return dict(precision=opts.precision, schedule=opts.schedule, lu=opts.lu,
                    cells_per_block=opts.cells_per_block, blocks_per_cu=opts.blocks_per_cu, programs=opts.programs,
                    seconds=best.value)

    @staticmethod
    def updateTimeAdvance(tadv, t_out, dt_out, stream=None):
        """tadv._tbeg = t_out, tadv._dt = dt_out (the driver's between-interval step)."""
        n = t_out.shape[0]
        _check(_lib.tchem_atmchem_update_time_advance(n, _ptr(tadv, None, "tadv"), _ptr(t_out), _ptr(dt_out),
                                                      _stream_ptr(stream)))


class Graph:
    """A captured interval loop (tchem_graph_t); keeps its arrays and the
    kmcd whose tables it reads alive."""

    def __init__(self, h, keep):
        self._h = h
        self._keep = keep

    def launch(self, stream=None):
        _check(_lib.tchem_graph_launch(self._h, _stream_ptr(stream)))

    def status(self):
        """(intervals run, stopped) of the last launch (synchronous)."""
        n, st = ctypes.c_int(), ctypes.c_int()
        _check(_lib.tchem_graph_status(self._h, ctypes.byref(n), ctypes.byref(st)))
        return n.value, bool(st.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tchem_graph_destroy(h)
            self._h = None


class AtmosphericChemistryE3SM:
    """The box model with the E3SM kinetics (AtmosphericChemistryE3SM::runDeviceBatch):
    per-cell photolysis rates [nBatch][nPhoto] and external sources [nBatch][m]."""

    @staticmethod
    def runDeviceBatch(tol_newton, tol_time, fac, tadv, state, photo_rates, external_sources, t_out, dt_out,
                       state_out, kmcd, stats=None, precision=PRECISION_FP64, schedule=SCHEDULE_GRID,
                       blocks_per_cu=0, lu=LU_AUTO, cells_per_block=0, programs=PROGRAMS_AUTO, stream=None):
        opts = _Options(precision, schedule, blocks_per_cu, lu, cells_per_block, programs)
        n_photo = 0 if photo_rates is None else photo_rates.shape[1]
        _check(_lib.tchem_atmchem_e3sm_run_device_batch(
            ctypes.byref(opts), state.shape[0], _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(fac, None, "fac"), _ptr(tadv, None, "tadv"), _ptr(state, what="state"),
            _ptr(photo_rates, what="photo_rates"), n_photo, _ptr(external_sources, what="external_sources"),
            _ptr(t_out, what="t_out"), _ptr(dt_out, what="dt_out"), _ptr(state_out, what="state_out"),
            _ptr(stats, None, "stats"), kmcd._h, _stream_ptr(stream)))


class SourceTerm:
    @staticmethod
    def runDeviceBatch(state, photo_rates, external_sources, omega, kmcd, stream=None):
        _check(_lib.tchem_sourceterm_run_device_batch(
            state.shape[0], _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"),
            _ptr(external_sources, what="external_sources"), _ptr(omega, what="omega"), kmcd._h,
            _stream_ptr(stream)))


class ReactionRates:
    """E3SM kinetics rate constants (ReactionRates::runDeviceBatch)."""

    @staticmethod
    def runDeviceBatch(state, kfwd, kmcd, stream=None):
        _check(_lib.tchem_reactionrates_run_device_batch(state.shape[0], _ptr(state, what="state"),
                                                         _ptr(kfwd, what="kfwd"), kmcd._h, _stream_ptr(stream)))


class RateofProgress:
    """E3SM kinetics rates of progress (RateofProgress::runDeviceBatch)."""

    @staticmethod
    def runDeviceBatch(state, photo_rates, rate_of_progress, kmcd, stream=None):
        n_photo = 0 if photo_rates is None else photo_rates.shape[1]
        _check(_lib.tchem_rateofprogress_run_device_batch(
            state.shape[0], _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"), n_photo,
            _ptr(rate_of_progress, what="rate_of_progress"), kmcd._h, _stream_ptr(stream)))


class NetProductionRates:
    """E3SM kinetics net production rates (NetProductionRates::runDeviceBatch);
    net_production_rates is [nBatch][ld] with ld >= active species (the
    reference's callers use nSpec; entries past the active species are 0)."""

    @staticmethod
    def runDeviceBatch(state, photo_rates, external_sources, net_production_rates, kmcd, stream=None):
        n_photo = 0 if photo_rates is None else photo_rates.shape[1]
        _check(_lib.tchem_netproductionrates_run_device_batch(
            state.shape[0], _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"), n_photo,
            _ptr(external_sources, what="external_sources"),
            _ptr(net_production_rates, what="net_production_rates"), net_production_rates.shape[1], kmcd._h,
            _stream_ptr(stream)))


def setScenarioConditionsPhotolysisReactions(filename: str, kmd: KineticModelData, nBatch: int) -> np.ndarray:
    """Host [nBatch][count] photolysis rates of a scenario file (count 0 without the block)."""
    n = ctypes.c_int()
    _check(_lib.tchem_scenario_photolysis_rates(kmd._h, filename.encode(), nBatch, None, 0, ctypes.byref(n)))
    out = np.zeros((nBatch, n.value))
    if n.value:
        _check(_lib.tchem_scenario_photolysis_rates(kmd._h, filename.encode(), nBatch, out.ctypes.data, n.value,
                                                    ctypes.byref(n)))
    return out


def setScenarioConditionsExternalForcing(filename: str, kmd: KineticModelData, nBatch: int) -> np.ndarray:
    """Host [nBatch][active species] external forcing of a scenario file."""
    out = np.zeros((nBatch, kmd.nActive))
    n = ctypes.c_int()
    _check(_lib.tchem_scenario_external_forcing(kmd._h, filename.encode(), nBatch, out.ctypes.data,
                                                ctypes.byref(n)))
    return out


class Jacobian:
    @staticmethod
    def runDeviceBatch(state, photo_rates, jacobian, kmcd, stream=None):
        _check(_lib.tchem_jacobian_run_device_batch(
            state.shape[0], _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"),
            _ptr(jacobian, what="jacobian"), kmcd._h, _stream_ptr(stream)))


class Adjoint:
    """Discrete adjoint of the recorded TR-BDF2 trajectory (no counterpart in
    the reference): gradients of a loss on the final concentrations with
    respect to the initial active species and per-reaction log-scale
    parameters theta (kf_r = k_r exp(theta_r)).  The autograd wrapper is
    adjoint.KineticsAdjoint."""

    @staticmethod
    def workspaceBytes(kmcd, nBatch, maxSteps):
        b = ctypes.c_size_t()
        _check(_lib.tchem_adjoint_workspace_bytes(kmcd._h, int(nBatch), int(maxSteps), ctypes.byref(b)))
        return b.value

    @staticmethod
    def workspace(kmcd, nBatch, maxSteps, device="cuda"):
        """Device workspace tensor (uint8) for nBatch cells and maxSteps accepted steps."""
        import torch
        return torch.empty(max(Adjoint.workspaceBytes(kmcd, nBatch, maxSteps), 1), dtype=torch.uint8,
                           device=device)

    @staticmethod
    def forward(tol_newton, tol_time, tadv, state, t_out, dt_out, workspace, maxSteps, kmcd, photo_rates=None,
                log_rate_scale=None, stats=None, stream=None):
        """Recording forward run from tadv._tbeg to ._tend (state integrated in
        place).  log_rate_scale: [nReac] shared, or [nBatch][nReac] per cell."""
        stride = 0 if log_rate_scale is None or log_rate_scale.dim() == 1 else log_rate_scale.shape[-1]
        _check(_lib.tchem_adjoint_forward(
            state.shape[0], _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(tadv, None, "tadv"), _ptr(state, what="state"), _ptr(photo_rates, what="photo_rates"),
            _ptr(log_rate_scale, what="log_rate_scale"), stride, _ptr(t_out, what="t_out"),
            _ptr(dt_out, what="dt_out"),
            _ptr(stats, None, "stats"), _ptr(workspace, None, "workspace"), int(maxSteps), kmcd._h,
            _stream_ptr(stream)))

    @staticmethod
    def backward(lam, lam_out, grad, workspace, maxSteps, kmcd, stream=None):
        """lam [nBatch][nSpec] = dL/d(final species) -> lam_out = dL/d(initial
        species), grad [nBatch][nReac] = dL/dtheta per cell."""
        _check(_lib.tchem_adjoint_backward(
            lam.shape[0], _ptr(lam, what="lambda"), _ptr(lam_out, what="lambda_out"), _ptr(grad, what="grad"),
            _ptr(workspace, None, "workspace"), int(maxSteps), kmcd._h, _stream_ptr(stream)))

    @staticmethod
    def steps(workspace, nBatch, maxSteps, kmcd, with_h=True, stream=None):
        """(nsteps [nBatch] int32, h [nBatch][maxSteps] or None) of the record."""
        import torch
        ns = torch.empty(max(nBatch, 1), dtype=torch.int32, device=workspace.device)
        h = torch.empty((max(nBatch, 1), maxSteps), dtype=torch.float64, device=workspace.device) if with_h else None
        _check(_lib.tchem_adjoint_steps(nBatch, _ptr(workspace, None, "workspace"), int(maxSteps),
                                        _ptr(ns, None, "nsteps"), _ptr(h, "float64", "h") if with_h else None,
                                        kmcd._h, _stream_ptr(stream)))
        return ns[:nBatch], (h[:nBatch] if with_h else None)


class AerosolModelData:
    """Single-particle SIMPOL.1 aerosol model read against a gas model
    (AerosolModelData, src/core/TChem_AerosolModelData.cpp:30-249)."""

    def __init__(self, mechfile: str, kmd: KineticModelData):
        h = ctypes.c_void_p()
        _check(_lib.tchem_amd_load_yaml(mechfile.encode(), kmd._h, ctypes.byref(h)))
        self._h = h
        self.kmd = kmd
        self._refresh()

    def _refresh(self):
        n = [ctypes.c_int() for _ in range(3)]
        _check(_lib.tchem_amd_info(self._h, *[ctypes.byref(x) for x in n]))
        self.nSpec, self.nParticles, self.nSimpol = (x.value for x in n)
        buf = ctypes.create_string_buffer(128)
        self.speciesNames, self.density, self.molecularWeights, self.kappa = [], [], [], []
        d, w, k = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        for i in range(self.nSpec):
            _check(_lib.tchem_amd_species_name(self._h, i, buf, 128))
            self.speciesNames.append(buf.value.decode())
            _check(_lib.tchem_amd_species_props(self._h, i, ctypes.byref(d), ctypes.byref(w), ctypes.byref(k)))
            self.density.append(d.value)
            self.molecularWeights.append(w.value)
            self.kappa.append(k.value)

    def setNumberofParticles(self, n: int):
        _check(_lib.tchem_amd_set_number_of_particles(self._h, int(n)))
        self._refresh()

    @property
    def nUnknowns(self) -> int:
        """Active gas species + particle species (getNumberOfTimeODEs)."""
        return self.kmd.nActive + self.nSpec * self.nParticles

    @property
    def stateDim(self) -> int:
        return 3 + self.kmd.nSpec + self.nSpec * self.nParticles

    def scenarioConditionParticles(self, filename: str, state: np.ndarray):
        """Particle block of a scenario file: writes particle species into the
        host state rows (in place) and returns number concentrations
        [nBatch][nParticles], or None when the file has no particle block."""
        assert state.dtype == np.float64 and state.flags.c_contiguous and state.shape[1] == self.stateDim
        nc = np.zeros((state.shape[0], max(self.nParticles, 1)))
        found = ctypes.c_int()
        _check(_lib.tchem_scenario_particles(self._h, filename.encode(), state.shape[0], self.stateDim,
                                             state.ctypes.data, nc.ctypes.data, ctypes.byref(found)))
        return nc[:, : self.nParticles].copy() if found.value else None

    def createConstData(self, kmcd: KineticModelConstData) -> "AerosolModelConstData":
        return AerosolModelConstData(self, kmcd)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tchem_amd_destroy(h)
            self._h = None


class AerosolModelConstData:
    """Device tables of an aerosol model, bound to the gas model's kmcd
    (create_AerosolModelConstData)."""

    def __init__(self, amd: AerosolModelData, kmcd: KineticModelConstData):
        h = ctypes.c_void_p()
        _check(_lib.tchem_amcd_create(amd._h, kmcd._h, ctypes.byref(h)))
        self._h = h
        self.kmcd = kmcd  # must outlive this object
        self.nUnknowns, self.stateDim, self.nParticles = amd.nUnknowns, amd.stateDim, amd.nParticles

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tchem_amcd_destroy(h)
            self._h = None


def create_AerosolModelConstData(amd: AerosolModelData, kmcd: KineticModelConstData) -> AerosolModelConstData:
    return AerosolModelConstData(amd, kmcd)


class AerosolChemistry:
    """Box model with single-particle SIMPOL.1 partitioning
    (AerosolChemistry::runDeviceBatch, src/core/TChem_AerosolChemistry.cpp:31-258)."""

    @staticmethod
    def runDeviceBatch(tol_newton, tol_time, fac, tadv, state, number_conc, t_out, dt_out, state_out, amcd,
                       stats=None, lu=LU_AUTO, stream=None):
        """lu: AUTO / SPARSE (compiled programs over gas + particle unknowns)
        or DENSE (dense LU with partial pivoting); FP64 either way."""
        nBatch = state.shape[0]
        opts = _Options(PRECISION_FP64, SCHEDULE_GRID, 0, lu, 0, PROGRAMS_AUTO)
        _check(_lib.tchem_aerosolchem_run_device_batch(
            ctypes.byref(opts), nBatch, _ptr(tol_newton, what="tol_newton"), _ptr(tol_time, what="tol_time"),
            _ptr(fac, None, "fac"), _ptr(tadv, None, "tadv"), _ptr(state, what="state"),
            _ptr(number_conc, what="number_conc"), _ptr(t_out, what="t_out"), _ptr(dt_out, what="dt_out"),
            _ptr(state_out, what="state_out"), _ptr(stats, None, "stats"), amcd._h, _stream_ptr(stream)))


class AerosolRHS:
    """Batched gas + SIMPOL source term (Aerosol_RHS; the reference's
    AerosolChemistry_RHSs example)."""

    @staticmethod
    def runDeviceBatch(state, number_conc, omega, amcd, stream=None):
        _check(_lib.tchem_aerosol_rhs_run_device_batch(
            state.shape[0], _ptr(state, what="state"), _ptr(number_conc, what="number_conc"),
            _ptr(omega, what="omega"), amcd._h, _stream_ptr(stream)))


def time_advance_tensor(n, tbeg, tend, dt, dtmin, dtmax, max_newton, iters_per_interval, jacobian_interval=1,
                        device="cuda"):
    """Device tensor (uint8, 56 bytes per cell) of TimeAdvance records."""
    import torch
    rec = np.zeros(n, TIME_ADVANCE_DTYPE)
    rec["_tbeg"], rec["_tend"], rec["_dt"], rec["_dtmin"], rec["_dtmax"] = tbeg, tend, dt, dtmin, dtmax
    rec["_max_num_newton_iterations"] = max_newton
    rec["_num_time_iterations_per_interval"] = iters_per_interval
    rec["_jacobian_interval"] = jacobian_interval
    return torch.from_numpy(rec.view(np.uint8).copy()).to(device)


def time_advance_from_tensor(t) -> np.ndarray:
    return t.cpu().numpy().view(TIME_ADVANCE_DTYPE)


def stats_tensor(n, device="cuda"):
    import torch
    return torch.zeros(n * CELL_STATS_DTYPE.itemsize, dtype=torch.uint8, device=device)


def stats_from_tensor(t) -> np.ndarray:
    return t.cpu().numpy().view(CELL_STATS_DTYPE)
