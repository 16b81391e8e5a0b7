"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The engine never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref.so")

(TSP_COORDS, TSP_MATRIX, KNAP_THRESHOLD, KNAP_GREEDY, TSPJ, BITSP_COORDS, BITSP_MATRIX, CONTINUOUS, JOBSHOP,
 ZDT1, VRP_DAYS, MOVRP) = range(12)
SPHERE, RASTRIGIN, ROSENBROCK = range(3)
TAG_PARENTS, TAG_COINS, TAG_MUTANT, TAG_INIT, TAG_MUTCOIN = range(5)

_PF = ctypes.POINTER(ctypes.c_float)
_PI = ctypes.POINTER(ctypes.c_int32)
_PU = ctypes.POINTER(ctypes.c_uint32)


class OrcProblem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("n", ctypes.c_int),
        ("a", _PF),
        ("b", _PF),
        ("w", _PI),
        ("v", _PI),
        ("cap", ctypes.c_int64),
        ("fx0", ctypes.c_int),
        ("fx1", ctypes.c_int),
        ("ip0", ctypes.c_int),
        ("f0", ctypes.c_float),
        ("f1", ctypes.c_float),
        ("d0", ctypes.c_double),
    ]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("pop", ctypes.c_int),
        ("elite", ctypes.c_int),
        ("mutants", ctypes.c_int),
        ("rho", ctypes.c_double),
        ("islands", ctypes.c_int),
        ("mig_interval", ctypes.c_int),
        ("mig_count", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("maximize", ctypes.c_int),
        ("ls_interval", ctypes.c_int),
        ("ls_count", ctypes.c_int),
        ("ls_moves", ctypes.c_int),
        ("div_interval", ctypes.c_int),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_PU, _PU, _PU]
        L.orc_philox_word.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32]
        L.orc_philox_word.restype = ctypes.c_uint32
        L.orc_u32_to_unit.argtypes = [ctypes.c_uint32]
        L.orc_u32_to_unit.restype = ctypes.c_float
        L.orc_fixed_shift.argtypes = [ctypes.c_double]
        L.orc_fixed_shift.restype = ctypes.c_int
        L.orc_problem_fixed_shift.argtypes = [ctypes.POINTER(OrcProblem), ctypes.c_int]
        L.orc_problem_fixed_shift.restype = ctypes.c_int
        L.orc_argsort_keys.argtypes = [_PF, ctypes.c_int, _PI]
        L.orc_tsp_edge_coords.argtypes = [_PF, ctypes.c_int, ctypes.c_int]
        L.orc_tsp_edge_coords.restype = ctypes.c_float
        L.orc_cos_det.argtypes = [ctypes.c_float]
        L.orc_cos_det.restype = ctypes.c_float
        L.orc_evaluate.argtypes = [ctypes.POINTER(OrcProblem), _PF, ctypes.c_int, _PF, _PI]
        L.orc_init_keys.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _PF]
        L.orc_partition.argtypes = [_PF, ctypes.c_int, ctypes.c_int, ctypes.c_int, _PI]
        L.orc_brkga_init.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), _PF, _PF, _PI]
        L.orc_brkga_step.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), ctypes.c_int,
                                     _PF, _PI, _PF, _PF, _PI]
        L.orc_brkga_run.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), ctypes.c_int,
                                    _PF, _PF, _PI]
        L.orc_tspj_local_search.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), ctypes.c_int,
                                            _PF, _PF, _PI]
        L.orc_brkga_run.restype = ctypes.c_double
        L.orc_nds_ranks.argtypes = [_PF, ctypes.c_int, ctypes.c_int, _PI]
        L.orc_crowding.argtypes = [_PF, _PI, ctypes.c_int, ctypes.c_int, _PF]
        L.orc_nsga2_select.argtypes = [_PF, _PI, _PF, ctypes.c_int, ctypes.c_int, ctypes.c_int, _PI]
        L.orc_nsga2_init.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), _PF, _PF, _PI, _PF]
        L.orc_diversity_unique.argtypes = [_PF, ctypes.c_int, ctypes.c_int, _PI]
        L.orc_diversity_unique.restype = ctypes.c_int
        L.orc_nsga2_diversity.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), ctypes.c_int,
                                          _PF, _PF, _PI, _PF]
        L.orc_nsga2_diversity.restype = ctypes.c_int
        L.orc_nsga2_step.restype = ctypes.c_int
        L.orc_nsga2_step.argtypes = [ctypes.POINTER(OrcProblem), ctypes.POINTER(OrcParams), ctypes.c_int,
                                     _PF, _PF, _PI, _PF, _PF, _PF, _PI, _PF]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_PF)


def _i(a):
    return a.ctypes.data_as(_PI)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def philox_word(seed, stream, gen, slot, tag, word):
    return lib().orc_philox_word(seed, stream, gen, slot, tag, word)


def u32_to_unit(u):
    return lib().orc_u32_to_unit(u)


def fixed_shift(m):
    return lib().orc_fixed_shift(float(m))


def cos_det(y):
    return lib().orc_cos_det(float(np.float32(y)))


def argsort_keys(keys):
    keys = np.ascontiguousarray(keys, np.float32)
    perm = np.empty(keys.shape[0], np.int32)
    lib().orc_argsort_keys(_f(keys), keys.shape[0], _i(perm))
    return perm


class Problem:
    """Oracle-side problem: the same data and kinds as the engine's."""

    def __init__(self, kind, n, a=None, b=None, w=None, v=None, cap=0, ip0=0, f0=0.0, f1=0.0, d0=0.0):
        self.kind = kind
        self.n = n
        self._a = None if a is None else np.ascontiguousarray(a, np.float32)
        self._b = None if b is None else np.ascontiguousarray(b, np.float32)
        self._w = None if w is None else np.ascontiguousarray(w, np.int32)
        self._v = None if v is None else np.ascontiguousarray(v, np.int32)
        self.c = OrcProblem()
        self.c.kind = kind
        self.c.n = n
        self.c.a = _f(self._a) if self._a is not None else _PF()
        self.c.b = _f(self._b) if self._b is not None else _PF()
        self.c.w = _i(self._w) if self._w is not None else _PI()
        self.c.v = _i(self._v) if self._v is not None else _PI()
        self.c.cap = int(cap)
        self.c.ip0 = int(ip0)
        self.c.f0 = float(f0)
        self.c.f1 = float(f1)
        self.c.d0 = float(d0)
        self.c.fx0 = lib().orc_problem_fixed_shift(ctypes.byref(self.c), 0)
        self.c.fx1 = lib().orc_problem_fixed_shift(ctypes.byref(self.c), 1)

    @classmethod
    def tsp_coords(cls, xy):
        xy = np.asarray(xy, np.float32).reshape(-1, 2)
        return cls(TSP_COORDS, xy.shape[0], a=xy)

    @classmethod
    def tsp_matrix(cls, dist):
        dist = np.asarray(dist, np.float32)
        return cls(TSP_MATRIX, dist.shape[0], a=dist)

    @classmethod
    def knapsack(cls, w, v, cap, greedy=True):
        w = np.asarray(w, np.int32)
        return cls(KNAP_GREEDY if greedy else KNAP_THRESHOLD, w.shape[0], w=w, v=v, cap=cap)

    @classmethod
    def tspj(cls, travel, jobs):
        travel = np.asarray(travel, np.float32)
        return cls(TSPJ, travel.shape[0] - 1, a=travel, b=jobs)

    @classmethod
    def bitsp_coords(cls, xy0, xy1):
        xy0 = np.asarray(xy0, np.float32).reshape(-1, 2)
        return cls(BITSP_COORDS, xy0.shape[0] - 1, a=xy0, b=np.asarray(xy1, np.float32).reshape(-1, 2))

    @classmethod
    def bitsp_matrix(cls, m0, m1):
        m0 = np.asarray(m0, np.float32)
        return cls(BITSP_MATRIX, m0.shape[0] - 1, a=m0, b=m1)

    @classmethod
    def continuous(cls, function, n, lower, upper):
        return cls(CONTINUOUS, n, ip0=function, f0=np.float32(lower), f1=np.float32(upper))

    @classmethod
    def job_scheduling(cls, job_times, machines):
        t = np.asarray(job_times, np.float32)
        return cls(JOBSHOP, t.shape[0], a=t, ip0=machines)

    @classmethod
    def zdt1(cls, n):
        return cls(ZDT1, n)

    @classmethod
    def vrp_days(cls, time_matrix, max_stops_per_day, max_drive_minutes):
        t = np.asarray(time_matrix, np.float32)
        return cls(VRP_DAYS, t.shape[0] - 1, a=t, ip0=max_stops_per_day, d0=max_drive_minutes)

    @classmethod
    def movrp(cls, xy, demands, capacity, vehicles):
        xy = np.asarray(xy, np.float32).reshape(-1, 2)
        return cls(MOVRP, xy.shape[0] - 1, a=xy, w=demands, cap=capacity, ip0=vehicles)

    @property
    def chrom_len(self):
        return 2 * self.n if self.kind == TSPJ else self.n

    @property
    def num_obj(self):
        return 2 if self.kind in (BITSP_COORDS, BITSP_MATRIX, ZDT1, MOVRP) else 1

    @property
    def maximize(self):
        return 1 if self.kind in (KNAP_THRESHOLD, KNAP_GREEDY) else 0

    def evaluate(self, keys):
        keys = np.ascontiguousarray(keys, np.float32).reshape(-1, self.chrom_len)
        cnt = keys.shape[0]
        fit = np.empty((cnt, self.num_obj), np.float32)
        dec = np.empty((cnt, self.chrom_len), np.int32)
        lib().orc_evaluate(ctypes.byref(self.c), _f(keys), cnt, _f(fit), _i(dec))
        return (fit[:, 0] if self.num_obj == 1 else fit), dec


def params(pop, elite, mutants, rho, islands=1, mig_interval=0, mig_count=0, seed=0, maximize=0,
           ls_interval=0, ls_count=0, ls_moves=0, div_interval=0):
    q = OrcParams()
    q.pop, q.elite, q.mutants, q.rho = pop, elite, mutants, rho
    q.islands, q.mig_interval, q.mig_count = islands, mig_interval, mig_count
    q.seed, q.maximize = seed & 0xFFFFFFFFFFFFFFFF, maximize
    q.ls_interval, q.ls_count, q.ls_moves = ls_interval, ls_count, ls_moves
    q.div_interval = div_interval
    return q


def diversity_unique(obj):
    """Indices of the members ensure_minimum_diversity keeps as unique."""
    obj = np.ascontiguousarray(obj, np.float32)
    kept = np.empty(obj.shape[0], np.int32)
    u = lib().orc_diversity_unique(_f(obj), obj.shape[0], obj.shape[1], _i(kept))
    return kept[:u].copy()


def init_keys(seed, stream, slots, length):
    k = np.empty((slots, length), np.float32)
    lib().orc_init_keys(seed, stream, slots, length, _f(k))
    return k


def tspj_local_search(problem, q, g, keys, fitness, perm):
    """In place: the local search the oracle applies to generation g."""
    assert keys.dtype == np.float32 and fitness.dtype == np.float32 and perm.dtype == np.int32
    lib().orc_tspj_local_search(ctypes.byref(problem.c), ctypes.byref(q), g, _f(keys), _f(fitness), _i(perm))


def partition(fitness, segments, seg_len, maximize):
    fitness = np.ascontiguousarray(fitness, np.float32)
    perm = np.empty(segments * seg_len, np.int32)
    lib().orc_partition(_f(fitness), segments, seg_len, maximize, _i(perm))
    return perm


class BrkgaState:
    """Oracle population, stepped one generation at a time."""

    def __init__(self, problem: Problem, q: OrcParams):
        self.p, self.q = problem, q
        P, L = q.pop, problem.chrom_len
        self.keys = np.empty((P, L), np.float32)
        self.fitness = np.empty(P, np.float32)
        self.order = np.empty(P, np.int32)
        self.t = 0
        lib().orc_brkga_init(ctypes.byref(problem.c), ctypes.byref(q), _f(self.keys), _f(self.fitness),
                             _i(self.order))

    def step(self):
        P, L = self.q.pop, self.p.chrom_len
        k2 = np.empty((P, L), np.float32)
        f2 = np.empty(P, np.float32)
        o2 = np.empty(P, np.int32)
        lib().orc_brkga_step(ctypes.byref(self.p.c), ctypes.byref(self.q), self.t, _f(self.keys),
                             _i(self.order), _f(k2), _f(f2), _i(o2))
        self.keys, self.fitness, self.order = k2, f2, o2
        self.t += 1


def nds_ranks(obj):
    obj = np.ascontiguousarray(obj, np.float32)
    r = np.empty(obj.shape[0], np.int32)
    lib().orc_nds_ranks(_f(obj), obj.shape[0], obj.shape[1], _i(r))
    return r


def crowding(obj, rank):
    obj = np.ascontiguousarray(obj, np.float32)
    rank = np.ascontiguousarray(rank, np.int32)
    c = np.empty(obj.shape[0], np.float32)
    lib().orc_crowding(_f(obj), _i(rank), obj.shape[0], obj.shape[1], _f(c))
    return c


def nsga2_select(obj, rank, crowd, keep):
    obj = np.ascontiguousarray(obj, np.float32)
    rank = np.ascontiguousarray(rank, np.int32)
    crowd = np.ascontiguousarray(crowd, np.float32)
    order = np.empty(keep, np.int32)
    lib().orc_nsga2_select(_f(obj), _i(rank), _f(crowd), obj.shape[0], obj.shape[1], keep, _i(order))
    return order


class Nsga2State:
    def __init__(self, problem: Problem, q: OrcParams):
        self.p, self.q = problem, q
        N, L = q.pop, problem.chrom_len
        self.keys = np.empty((N, L), np.float32)
        self.obj = np.empty((N, 2), np.float32)
        self.rank = np.empty(N, np.int32)
        self.crowd = np.empty(N, np.float32)
        self.t = 0
        self.injected = 0      # 1 when the last step ended with a diversity injection
        self.injections = 0
        lib().orc_nsga2_init(ctypes.byref(problem.c), ctypes.byref(q), _f(self.keys), _f(self.obj),
                             _i(self.rank), _f(self.crowd))

    def step(self):
        N, L = self.q.pop, self.p.chrom_len
        k2 = np.empty((N, L), np.float32)
        o2 = np.empty((N, 2), np.float32)
        r2 = np.empty(N, np.int32)
        c2 = np.empty(N, np.float32)
        self.injected = lib().orc_nsga2_step(ctypes.byref(self.p.c), ctypes.byref(self.q), self.t, _f(self.keys),
                                             _f(self.obj), _i(self.rank), _f(self.crowd), _f(k2), _f(o2), _i(r2),
                                             _f(c2))
        self.injections += self.injected
        self.keys, self.obj, self.rank, self.crowd = k2, o2, r2, c2
        self.t += 1


_ref = None


def ref_lib():
    """The reference's own CUDA-free headers compiled into oracle/_ref (None if not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB_PATH):
            return None
        L = ctypes.CDLL(REF_LIB_PATH)
        L.ref_knapsack_threshold.argtypes = [_PF, _PF, ctypes.c_int, ctypes.c_float, _PF, ctypes.c_int, _PF, _PI]
        L.ref_multitsp.argtypes = [_PF, _PF, ctypes.c_int, _PF, ctypes.c_int, _PF, _PI]
        L.ref_dominates.argtypes = [_PF, ctypes.c_int, ctypes.c_int, _PI]
        L.ref_knapsack_load.argtypes = [ctypes.c_char_p, _PF, _PF, ctypes.c_int, _PF]
        L.ref_knapsack_load.restype = ctypes.c_int
        L.ref_multitsp_load_objectives.argtypes = [ctypes.c_char_p, _PF, ctypes.c_int, _PF]
        L.ref_multitsp_load_objectives.restype = ctypes.c_int
        L.ref_vrp_days.argtypes = [_PF, ctypes.c_int, ctypes.c_int, ctypes.c_double, _PF, ctypes.c_int, _PF, _PI]
        L.ref_movrp_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ref_movrp_info.restype = ctypes.c_int
        L.ref_movrp.argtypes = [ctypes.c_char_p, _PF, ctypes.c_int, _PF, _PI]
        _ref = L
    return _ref
