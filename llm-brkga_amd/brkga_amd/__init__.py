"""Python (ctypes) binding of the MI355X BRKGA engine C ABI (include/brkga_hip.h).

The engine itself is the HIP library ``libbrkga_hip.so`` next to this file;
this module only marshals numpy arrays and handles.  There is no CPU
fallback: if the library is missing or cannot be loaded, every call raises.

Object model (mirrors the reference's ``Solver<T>`` / ``Population<T>`` /
``BRKGAConfig<T>``, core/solver.hpp, core/population.hpp, core/config.hpp):

    eng = Engine(0)
    prob = Problem.tsp_coords(eng, xy)          # configs/tsp_config.hpp
    pop = Population(prob, Params(population_size=1024, elite_size=152, ...))
    pop.initialize(); pop.evolve(100)
    fitness, keys, tour = pop.best()
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbrkga_hip.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_OOM = 2
ERR_DEVICE = 3
ERR_BUSY = 4
ERR_DEADLINE = 5
ERR_NOT_FOUND = 6
ERR_NO_DEVICE = 7
ERR_CANCELLED = 8

TSP_COORDS = 0
TSP_MATRIX = 1
KNAPSACK_THRESHOLD = 2
KNAPSACK_GREEDY = 3
TSPJ = 4
BITSP_COORDS = 5
BITSP_MATRIX = 6
CONTINUOUS = 7
JOB_SCHEDULING = 8
ZDT1 = 9
VRP_DAYS = 10
MOVRP = 11
HOST = 12

SPHERE = 0
RASTRIGIN = 1
ROSENBROCK = 2


class BrkgaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"brkga status {status}: {message}")
        self.status = status


class CParams(ctypes.Structure):
    _fields_ = [
        ("population_size", ctypes.c_int32),
        ("elite_size", ctypes.c_int32),
        ("mutant_size", ctypes.c_int32),
        ("elite_prob", ctypes.c_double),
        ("num_islands", ctypes.c_int32),
        ("migration_interval", ctypes.c_int32),
        ("num_migrants", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("ls_interval", ctypes.c_int32),
        ("ls_individuals", ctypes.c_int32),
        ("ls_moves", ctypes.c_int32),
        ("diversity_interval", ctypes.c_int32),
    ]


@dataclass
class Params:
    """GA parameters (BRKGAConfig fields, core/config.hpp:22-26, plus the
    island settings core/solver.hpp:164-167 hard-codes)."""

    population_size: int = 100
    elite_size: int = 20
    mutant_size: int = 10
    elite_prob: float = 0.7
    num_islands: int = 1
    migration_interval: int = 0
    num_migrants: int = 0
    seed: int = 0
    ls_interval: int = 0      # TSPJ local search (tspj_config.hpp:77-135); 0 = off
    ls_individuals: int = 0
    ls_moves: int = 0
    diversity_interval: int = 0  # NSGA-II ensure_minimum_diversity (population.hpp:130-182); 0 = off, 25 = reference

    def to_c(self) -> CParams:
        return CParams(self.population_size, self.elite_size, self.mutant_size, float(self.elite_prob),
                       self.num_islands, self.migration_interval, self.num_migrants,
                       self.seed & 0xFFFFFFFFFFFFFFFF, self.ls_interval, self.ls_individuals, self.ls_moves,
                       self.diversity_interval)


class CServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_resident", ctypes.c_int32),
                ("ticks_per_round", ctypes.c_int32), ("autotune_reps", ctypes.c_int32)]


class CRequestInfo(ctypes.Structure):
    _fields_ = [("submit_us", ctypes.c_double), ("admit_us", ctypes.c_double), ("done_us", ctypes.c_double),
                ("generations", ctypes.c_int32), ("status", ctypes.c_int32), ("device_bytes", ctypes.c_uint64)]


class CServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("admitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("resident", ctypes.c_int32), ("pending", ctypes.c_int32),
                ("bytes_in_use", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64), ("ticks", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
                ("wait_us", ctypes.c_double), ("host_us", ctypes.c_double), ("lanes", ctypes.c_int32)]


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_PF = ctypes.POINTER(ctypes.c_float)
_PI = ctypes.POINTER(ctypes.c_int32)

_SIGNATURES = {
    "brkga_version": ([], ctypes.c_char_p),
    "brkga_last_error": ([], ctypes.c_char_p),
    "brkga_device_count": ([_PI], ctypes.c_int),
    "brkga_engine_create": ([_I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_engine_destroy": ([_P], ctypes.c_int),
    "brkga_problem_create_tsp_coords": ([_P, _PF, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_tsp_matrix": ([_P, _PF, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_knapsack": ([_P, _PI, _PI, _I32, ctypes.c_int64, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_tspj": ([_P, _PF, _PF, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_bitsp_coords": ([_P, _PF, _PF, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_bitsp_matrix": ([_P, _PF, _PF, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_continuous": ([_P, _I32, _I32, ctypes.c_float, ctypes.c_float, ctypes.POINTER(_P)],
                                        ctypes.c_int),
    "brkga_problem_create_job_scheduling": ([_P, _PF, _I32, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_zdt1": ([_P, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_vrp_days": ([_P, _PF, _I32, _I32, ctypes.c_double, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_movrp": ([_P, _PF, _PI, _I32, ctypes.c_int64, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_create_host": ([_P, _I32, _I32, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_problem_info": ([_P, _PI, _PI, _PI, _PI], ctypes.c_int),
    "brkga_problem_destroy": ([_P], ctypes.c_int),
    "brkga_evaluate": ([_P, _P, _P, _I32, _I32, _P], ctypes.c_int),
    "brkga_decode": ([_P, _P, _P, _I32, _I32, _P], ctypes.c_int),
    "brkga_evaluate_host": ([_P, _PF, _I32, _PF, _PI], ctypes.c_int),
    "brkga_evaluate_host_noalloc": ([_P, _PF, _I32, _PF, _PI], ctypes.c_int),
    "brkga_population_create": ([_P, ctypes.POINTER(CParams), ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_population_destroy": ([_P], ctypes.c_int),
    "brkga_population_initialize": ([_P, _P], ctypes.c_int),
    "brkga_evolve": ([_P, _I32, _P], ctypes.c_int),
    "brkga_evolve_graph": ([_P, _I32, _P], ctypes.c_int),
    "brkga_population_host_begin": ([_P, _I32, _PF, _PF, _PI, _P], ctypes.c_int),
    "brkga_population_host_end": ([_P, _PF, _P], ctypes.c_int),
    "brkga_population_generation": ([_P, _PI], ctypes.c_int),
    "brkga_population_read": ([_P, _PF, _PF, _PI, _PI, _PF, _P], ctypes.c_int),
    "brkga_population_write": ([_P, _PF, _P], ctypes.c_int),
    "brkga_get_best": ([_P, _PF, _PF, _PI, _P], ctypes.c_int),
    "brkga_population_restore": ([_P, _PF, _PF, _PI, _PF, _I32, _P], ctypes.c_int),
    "brkga_population_device_buffers": ([_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P), _PI],
                                        ctypes.c_int),
    "brkga_population_profile": ([_P, _I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _P],
                                 ctypes.c_int),
    "brkga_population_autotune": ([_P, _I32, _PI, _P], ctypes.c_int),
    "brkga_population_set_block": ([_P, _I32], ctypes.c_int),
    "brkga_population_set_children": ([_P, _I32], ctypes.c_int),
    "brkga_population_step_shape": ([_P, _PI, _PI], ctypes.c_int),
    "brkga_batch_create": ([_P, _I32, ctypes.POINTER(_P), ctypes.POINTER(CParams), ctypes.POINTER(_P)],
                           ctypes.c_int),
    "brkga_batch_destroy": ([_P], ctypes.c_int),
    "brkga_batch_initialize": ([_P, _I32, _P], ctypes.c_int),
    "brkga_batch_evolve": ([_P, _I32, _I32, _P], ctypes.c_int),
    "brkga_batch_autotune": ([_P, _I32, _P], ctypes.c_int),
    "brkga_batch_population": ([_P, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_batch_num_buckets": ([_P, _PI], ctypes.c_int),
    "brkga_batch_bucket_info": ([_P, _I32, _PI, _PI, _PI, _PI, _PI, _PI], ctypes.c_int),
    "brkga_batch_lanes": ([_P, _PI], ctypes.c_int),
    "brkga_server_start": ([_P, ctypes.POINTER(CServerConfig), ctypes.POINTER(_P)], ctypes.c_int),
    "brkga_server_submit": ([_P, _P, ctypes.POINTER(CParams), _I32, _I32, ctypes.POINTER(ctypes.c_uint64)],
                            ctypes.c_int),
    "brkga_server_wait": ([_P, ctypes.c_uint64, _I32, _PF, _PI, ctypes.POINTER(CRequestInfo)], ctypes.c_int),
    "brkga_server_stats_get": ([_P, ctypes.POINTER(CServerStats)], ctypes.c_int),
    "brkga_server_cancel": ([_P, ctypes.c_uint64], ctypes.c_int),
    "brkga_server_stop": ([_P], ctypes.c_int),
}


def exported_symbols():
    """Names the C ABI exports (every function declared in include/brkga_hip.h)."""
    return list(_SIGNATURES)


def lib():
    """Loads libbrkga_hip.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(status: int):
    if status != OK:
        msg = lib().brkga_last_error().decode(errors="replace")
        raise BrkgaError(status, msg)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_PF)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_PI)


def device_count() -> int:
    c = ctypes.c_int32(0)
    _check(lib().brkga_device_count(ctypes.byref(c)))
    return c.value


class Engine:
    """One engine per device (Solver::initialize_gpu_system, core/solver.hpp:1307-1394)."""

    def __init__(self, device: int = 0):
        h = _P()
        _check(lib().brkga_engine_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().brkga_engine_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """A problem plug-in with engine-owned device data (BRKGAConfig subclasses)."""

    def __init__(self, engine: Engine, handle):
        self.engine = engine
        self.handle = handle
        kind, L, nobj, mx = (ctypes.c_int32() for _ in range(4))
        _check(lib().brkga_problem_info(handle, ctypes.byref(kind), ctypes.byref(L), ctypes.byref(nobj),
                                        ctypes.byref(mx)))
        self.kind, self.chrom_len, self.num_obj, self.maximize = kind.value, L.value, nobj.value, mx.value

    @staticmethod
    def _make(engine, fn, *args):
        h = _P()
        _check(fn(engine.handle, *args, ctypes.byref(h)))
        return Problem(engine, h)

    @classmethod
    def tsp_coords(cls, engine, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
        p = cls._make(engine, lib().brkga_problem_create_tsp_coords, _fptr(xy), xy.shape[0])
        p._data = (xy,)
        return p

    @classmethod
    def tsp_matrix(cls, engine, dist):
        dist = np.ascontiguousarray(dist, dtype=np.float32)
        n = dist.shape[0]
        return cls._make(engine, lib().brkga_problem_create_tsp_matrix, _fptr(dist), n)

    @classmethod
    def knapsack(cls, engine, weights, values, capacity, greedy=True):
        w = np.ascontiguousarray(weights, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.int32)
        return cls._make(engine, lib().brkga_problem_create_knapsack, _iptr(w), _iptr(v), w.shape[0],
                         int(capacity), 1 if greedy else 0)

    @classmethod
    def tspj(cls, engine, travel, jobs):
        t = np.ascontiguousarray(travel, dtype=np.float32)
        j = np.ascontiguousarray(jobs, dtype=np.float32)
        return cls._make(engine, lib().brkga_problem_create_tspj, _fptr(t), _fptr(j), t.shape[0])

    @classmethod
    def bitsp_coords(cls, engine, xy0, xy1):
        a = np.ascontiguousarray(xy0, dtype=np.float32).reshape(-1, 2)
        b = np.ascontiguousarray(xy1, dtype=np.float32).reshape(-1, 2)
        return cls._make(engine, lib().brkga_problem_create_bitsp_coords, _fptr(a), _fptr(b), a.shape[0] - 1)

    @classmethod
    def bitsp_matrix(cls, engine, m0, m1):
        a = np.ascontiguousarray(m0, dtype=np.float32)
        b = np.ascontiguousarray(m1, dtype=np.float32)
        return cls._make(engine, lib().brkga_problem_create_bitsp_matrix, _fptr(a), _fptr(b), a.shape[0])

    @classmethod
    def continuous(cls, engine, function, n, lower, upper):
        """ContinuousOptGPUConfig: function SPHERE / RASTRIGIN / ROSENBROCK over n variables."""
        return cls._make(engine, lib().brkga_problem_create_continuous, int(function), int(n), float(lower),
                         float(upper))

    @classmethod
    def job_scheduling(cls, engine, job_times, machines):
        t = np.ascontiguousarray(job_times, dtype=np.float32)
        return cls._make(engine, lib().brkga_problem_create_job_scheduling, _fptr(t), t.shape[0], int(machines))

    @classmethod
    def zdt1(cls, engine, n):
        return cls._make(engine, lib().brkga_problem_create_zdt1, int(n))

    @classmethod
    def vrp_days(cls, engine, time_matrix, max_stops_per_day, max_drive_minutes):
        """VrpBlazerConfig: time_matrix[dim][dim], location 0 the depot."""
        t = np.ascontiguousarray(time_matrix, dtype=np.float32)
        return cls._make(engine, lib().brkga_problem_create_vrp_days, _fptr(t), t.shape[0], int(max_stops_per_day),
                         float(max_drive_minutes))

    @classmethod
    def movrp(cls, engine, xy, demands, capacity, vehicles):
        """MultiObjectiveVRPConfig: xy[dim][2], demands[dim], location 0 the depot."""
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
        d = np.ascontiguousarray(demands, dtype=np.int32)
        return cls._make(engine, lib().brkga_problem_create_movrp, _fptr(xy), _iptr(d), xy.shape[0], int(capacity),
                         int(vehicles))

    @classmethod
    def host(cls, engine, chrom_len, num_obj=1, maximize=False):
        """A problem the caller evaluates on the host (a config with a host
        fitness_function); its populations step with host_generation()."""
        return cls._make(engine, lib().brkga_problem_create_host, int(chrom_len), int(num_obj),
                         1 if maximize else 0)

    def evaluate_device(self, d_keys: int, d_fitness: int, pop_size: int, stream: int = 0):
        """evaluate_population_gpu on caller-owned device pointers (ints)."""
        _check(lib().brkga_evaluate(self.handle, d_keys, d_fitness, pop_size, self.chrom_len, stream or None))

    def evaluate(self, keys):
        """Fitness [rows(, num_obj)] and decoded solutions [rows][chrom_len] of host keys."""
        keys = np.ascontiguousarray(keys, dtype=np.float32).reshape(-1, self.chrom_len)
        fit = np.empty((keys.shape[0], self.num_obj), np.float32)
        dec = np.empty(keys.shape, np.int32)
        _check(lib().brkga_evaluate_host(self.handle, _fptr(keys), keys.shape[0], _fptr(fit), _iptr(dec)))
        return (fit[:, 0] if self.num_obj == 1 else fit), dec

    def decode_device(self, d_keys: int, d_decoded: int, pop_size: int, stream: int = 0):
        _check(lib().brkga_decode(self.handle, d_keys, d_decoded, pop_size, self.chrom_len, stream or None))

    def close(self):
        if self.handle:
            lib().brkga_problem_destroy(self.handle)
            self.handle = _P()


class Population:
    """Device-resident population (Population<T> / GPUPopulation<T>)."""

    def __init__(self, problem: Problem, params: Params):
        self.problem = problem
        self.params = params
        self._cp = params.to_c()
        h = _P()
        _check(lib().brkga_population_create(problem.handle, ctypes.byref(self._cp), ctypes.byref(h)))
        self.handle = h
        self._owned = True

    @classmethod
    def _borrowed(cls, problem: Problem, params: Params, handle, owner):
        """View of a population owned by a batch (not destroyed by this object)."""
        self = cls.__new__(cls)
        self.problem = problem
        self.params = params
        self._cp = params.to_c()
        self.handle = handle
        self._owned = False
        self._owner = owner
        return self

    @property
    def size(self):
        return self.params.population_size

    def initialize(self, stream: int = 0):
        _check(lib().brkga_population_initialize(self.handle, stream or None))

    def evolve(self, generations: int, graph: bool = False, stream: int = 0):
        fn = lib().brkga_evolve_graph if graph else lib().brkga_evolve
        _check(fn(self.handle, generations, stream or None))

    @property
    def generation(self) -> int:
        g = ctypes.c_int32()
        _check(lib().brkga_population_generation(self.handle, ctypes.byref(g)))
        return g.value

    def read(self, stream: int = 0):
        P, L, M = self.size, self.problem.chrom_len, self.problem.num_obj
        keys = np.empty((P, L), np.float32)
        fit = np.empty((P, M), np.float32)
        order = np.empty(P, np.int32)
        rank = np.full(P, -1, np.int32)
        crowd = np.zeros(P, np.float32)
        _check(lib().brkga_population_read(self.handle, _fptr(keys), _fptr(fit), _iptr(order), _iptr(rank),
                                           _fptr(crowd), stream or None))
        out = {"keys": keys, "fitness": fit if M > 1 else fit[:, 0], "order": order}
        if M > 1:
            out["rank"] = rank
            out["crowding"] = crowd
        return out

    def write(self, keys, stream: int = 0):
        keys = np.ascontiguousarray(keys, dtype=np.float32)
        assert keys.shape == (self.size, self.problem.chrom_len)
        _check(lib().brkga_population_write(self.handle, _fptr(keys), stream or None))

    def checkpoint(self):
        """State to resume from: read() plus the generation counter."""
        state = self.read()
        state["generation"] = self.generation
        return state

    def restore(self, state, stream: int = 0):
        """Resume from checkpoint() of a population with the same problem and
        parameters; later steps are bit-identical to the original run's."""
        P, L, M = self.size, self.problem.chrom_len, self.problem.num_obj
        keys = np.ascontiguousarray(state["keys"], dtype=np.float32).reshape(P, L)
        fit = np.ascontiguousarray(state["fitness"], dtype=np.float32).reshape(P, M)
        rank = crowd = None
        if M > 1:
            rank = np.ascontiguousarray(state["rank"], dtype=np.int32)
            crowd = np.ascontiguousarray(state["crowding"], dtype=np.float32)
        _check(lib().brkga_population_restore(self.handle, _fptr(keys), _fptr(fit),
                                              None if rank is None else _iptr(rank),
                                              None if crowd is None else _fptr(crowd),
                                              int(state["generation"]), stream or None))

    def best(self, stream: int = 0):
        M, L = self.problem.num_obj, self.problem.chrom_len
        fit = np.empty(M, np.float32)
        keys = np.empty(L, np.float32)
        host = self.problem.kind == HOST  # no device decoder: no decoded solution
        dec = None if host else np.empty(L, np.int32)
        _check(lib().brkga_get_best(self.handle, _fptr(fit), _fptr(keys), None if host else _iptr(dec),
                                    stream or None))
        return (fit if M > 1 else float(fit[0])), keys, dec

    def host_begin(self, initialize: bool = False, stream: int = 0):
        """Keys [rows][L] of the next generation (or generation 0), the fitness
        of copied rows and the mask of rows the caller must evaluate."""
        L, M = self.problem.chrom_len, self.problem.num_obj
        rows = self.size
        keys = np.empty((rows, L), np.float32)
        fit = np.empty((rows, M), np.float32)
        needs = np.empty(rows, np.int32)
        _check(lib().brkga_population_host_begin(self.handle, 1 if initialize else 0, _fptr(keys), _fptr(fit),
                                                 _iptr(needs), stream or None))
        return keys, fit, needs.astype(bool)

    def host_end(self, fitness, stream: int = 0):
        M = self.problem.num_obj
        f = np.ascontiguousarray(fitness, dtype=np.float32).reshape(self.size, M)
        _check(lib().brkga_population_host_end(self.handle, _fptr(f), stream or None))

    def host_generation(self, evaluate, initialize: bool = False):
        """One host-evaluated generation: evaluate(keys) -> fitness rows for
        the rows that need it."""
        keys, fit, needs = self.host_begin(initialize)
        idx = np.nonzero(needs)[0]
        if idx.size:
            fit[idx] = np.asarray(evaluate(keys[idx]), np.float32).reshape(idx.size, -1)
        self.host_end(fit)

    def set_block(self, block: int):
        """Workgroup size of the generation kernel (results never depend on it)."""
        _check(lib().brkga_population_set_block(self.handle, int(block)))

    def set_children(self, children: int):
        """Children per workgroup of the ranking step (1, 2, 4); results never
        depend on it."""
        _check(lib().brkga_population_set_children(self.handle, int(children)))

    def step_shape(self):
        """(threads per child, children per workgroup) of the generation kernel."""
        b, c = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().brkga_population_step_shape(self.handle, ctypes.byref(b), ctypes.byref(c)))
        return b.value, c.value

    def autotune(self, reps: int = 20, stream: int = 0) -> int:
        """Fastest workgroup size of the generation kernel (results unchanged)."""
        b = ctypes.c_int32()
        _check(lib().brkga_population_autotune(self.handle, reps, ctypes.byref(b), stream or None))
        return b.value

    def profile(self, reps: int, stream: int = 0):
        """Mean (generation kernel, closing kernels) device time in
        microseconds over `reps` real generations (the population advances
        `reps` generations).  The closing time is 0 when the step ranks its
        parent island itself (islands no larger than the workgroup)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().brkga_population_profile(self.handle, reps, ctypes.byref(a), ctypes.byref(b),
                                              stream or None))
        return a.value, b.value

    def close(self):
        if self.handle and self._owned:
            lib().brkga_population_destroy(self.handle)
        self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_PLAIN = 0
BATCH_LAUNCH = 1
BATCH_GRAPH = 2
_MODES = {"plain": BATCH_PLAIN, "launch": BATCH_LAUNCH, "graph": BATCH_GRAPH}


class Batch:
    """Many independent populations evolved together (brkga_batch_*).

    mode "plain" runs the members one at a time through the single-population
    launches, "launch" runs every member per generation through the
    queue-driven batched kernels, "graph" replays those from a HIP graph.
    Each member evolves exactly as it would alone."""

    def __init__(self, engine: Engine, problems, params):
        problems = list(problems)
        params = list(params) if not isinstance(params, Params) else [params] * len(problems)
        assert len(problems) == len(params) and problems
        self.engine = engine
        self.problems = problems
        self.params = params
        hs = (_P * len(problems))(*[p.handle for p in problems])
        cps = (CParams * len(params))(*[q.to_c() for q in params])
        h = _P()
        _check(lib().brkga_batch_create(engine.handle, len(problems), hs, cps, ctypes.byref(h)))
        self.handle = h
        self._members = {}

    def __len__(self):
        return len(self.problems)

    @staticmethod
    def _mode(mode):
        return _MODES[mode] if isinstance(mode, str) else int(mode)

    def initialize(self, mode="launch", stream: int = 0):
        _check(lib().brkga_batch_initialize(self.handle, self._mode(mode), stream or None))

    def evolve(self, generations: int, mode="launch", stream: int = 0):
        _check(lib().brkga_batch_evolve(self.handle, generations, self._mode(mode), stream or None))

    def autotune(self, reps: int = 3, stream: int = 0):
        _check(lib().brkga_batch_autotune(self.handle, reps, stream or None))

    def member(self, i: int) -> Population:
        if i not in self._members:
            h = _P()
            _check(lib().brkga_batch_population(self.handle, i, ctypes.byref(h)))
            self._members[i] = Population._borrowed(self.problems[i], self.params[i], h, self)
        return self._members[i]

    def buckets(self):
        """Per size bucket: kind, n range, member count, workgroup size, slots per item."""
        c = ctypes.c_int32()
        _check(lib().brkga_batch_num_buckets(self.handle, ctypes.byref(c)))
        out = []
        for b in range(c.value):
            v = [ctypes.c_int32() for _ in range(6)]
            _check(lib().brkga_batch_bucket_info(self.handle, b, *[ctypes.byref(x) for x in v]))
            kind, n_min, n_max, count, block, chunk = (x.value for x in v)
            out.append({"kind": kind, "n_min": n_min, "n_max": n_max, "members": count, "block": block,
                        "chunk": chunk})
        return out

    def lanes(self):
        """Streams the batched launches are spread over (1 until the first batched step)."""
        c = ctypes.c_int32()
        _check(lib().brkga_batch_lanes(self.handle, ctypes.byref(c)))
        return c.value

    def close(self):
        if self.handle:
            for m in self._members.values():
                m.handle = _P()
            self._members = {}
            lib().brkga_batch_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (brkga_server_*).

    submit() queues a solve and returns its id; wait() returns
    (fitness, decoded, info) once the request reached its generation
    deadline.  Populations are admitted while they fit `hbm_budget_bytes`."""

    def __init__(self, engine: Engine, hbm_budget_bytes: int = 0, max_resident: int = 0,
                 ticks_per_round: int = 0, autotune_reps: int = 0):
        self.engine = engine
        cfg = CServerConfig(hbm_budget_bytes, max_resident, ticks_per_round, autotune_reps)
        h = _P()
        _check(lib().brkga_server_start(engine.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._shape = {}  # request id -> (num_obj, chrom_len)

    def submit(self, problem: Problem, params: Params, generations: int, own_problem: bool = False) -> int:
        rid = ctypes.c_uint64()
        cp = params.to_c()
        _check(lib().brkga_server_submit(self.handle, problem.handle, ctypes.byref(cp), generations,
                                         1 if own_problem else 0, ctypes.byref(rid)))
        if own_problem:
            problem.handle = _P()  # the server destroys it
        self._shape[rid.value] = (problem.num_obj, problem.chrom_len)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1):
        """(fitness, decoded, info dict); raises BrkgaError(ERR_BUSY) on timeout."""
        M, L = self._shape.get(rid, (2, 1))
        fit = np.empty(M, np.float32)
        dec = np.empty(L, np.int32)
        info = CRequestInfo()
        st = lib().brkga_server_wait(self.handle, rid, timeout_ms, _fptr(fit), _iptr(dec), ctypes.byref(info))
        if st == ERR_BUSY:
            raise BrkgaError(st, "request still running")
        _check(st)
        self._shape.pop(rid, None)
        d = {k: getattr(info, k) for k, _ in CRequestInfo._fields_}
        return (fit if M > 1 else float(fit[0])), dec, d

    def cancel(self, rid: int):
        """Cancels a request before its deadline; wait() then raises
        BrkgaError(ERR_CANCELLED).  No effect on a finished request."""
        _check(lib().brkga_server_cancel(self.handle, rid))

    def stats(self):
        st = CServerStats()
        _check(lib().brkga_server_stats_get(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in CServerStats._fields_}

    def stop(self):
        if self.handle:
            lib().brkga_server_stop(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
