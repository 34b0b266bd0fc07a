"""ctypes binding of the C++/OpenMP restatement (oracle/d2q9_ref.cpp).

TEST INFRASTRUCTURE ONLY: used by the tests (fast parity at large step
counts) and by the CPU-baseline leg of ``bench.py`` (timed reference path).
Runs the same passes as ``oracle.d2q9_oracle`` on an ``OracleCase``, in
place, and agrees with it bit for bit.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_d2q9.so")
_TYPES = {"bounce_back": 0, "fixed_velocity": 1, "fixed_pressure": 2,
          "zero_gradient": 3, "periodic": 4}


class _Elem(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("n_nodes", ctypes.c_int),
                ("nodes", ctypes.c_void_p), ("out", ctypes.c_int * 3),
                ("inv", ctypes.c_int * 3), ("vec", ctypes.c_double * 2),
                ("scalar", ctypes.c_double), ("normal", ctypes.c_int * 2)]


class _Case(ctypes.Structure):
    _fields_ = [("nx_pad", ctypes.c_int), ("ny_pad", ctypes.c_int),
                ("solid", ctypes.c_void_p), ("ghost", ctypes.c_void_p),
                ("density", ctypes.c_void_p), ("velocity", ctypes.c_void_p),
                ("force", ctypes.c_void_p), ("pop", ctypes.c_void_p),
                ("pop_new", ctypes.c_void_p), ("omega", ctypes.c_double),
                ("inv_cs_2", ctypes.c_double), ("inv_cs_4", ctypes.c_double),
                ("float_min", ctypes.c_double), ("w", ctypes.c_double * 9),
                ("model", ctypes.c_int), ("forcing", ctypes.c_int),
                ("force_source", ctypes.c_int),
                ("gravity", ctypes.c_double * 2),
                ("x_periodic", ctypes.c_int), ("y_periodic", ctypes.c_int),
                ("n_elements", ctypes.c_int), ("elems", ctypes.c_void_p)]


def build():
    """Compile the library with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                "oracle/liboracle_d2q9.so is not built (run make -C oracle)")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_run.argtypes = [ctypes.POINTER(_Case), ctypes.c_int,
                                    ctypes.c_int]
        _lib.oracle_initialize.argtypes = [ctypes.POINTER(_Case)]
        _lib.oracle_max_threads.restype = ctypes.c_int
    return _lib


def max_threads():
    return int(load().oracle_max_threads())


class RefRunner:
    """Holds the ctypes view of one ``OracleCase``; arrays are shared, so
    ``case.pop`` etc. see every update."""

    def __init__(self, case):
        self.case = case
        self._solid = np.ascontiguousarray(case.solid.astype(np.uint8))
        self._ghost = np.ascontiguousarray(case.ghost.astype(np.uint8))
        for name in ("density", "velocity", "force_field", "pop", "pop_new"):
            arr = getattr(case, name)
            assert arr.flags.c_contiguous and arr.dtype == np.float64
        elems = [e for e in case.elements]
        self._nodes = [np.ascontiguousarray(e["nodes"], dtype=np.int64)
                       for e in elems]
        self._elems = (_Elem * max(1, len(elems)))()
        for n, e in enumerate(elems):
            el = self._elems[n]
            el.type = _TYPES[e["type"]]
            el.n_nodes = int(self._nodes[n].size)
            el.nodes = self._nodes[n].ctypes.data
            for q in range(3):
                el.out[q] = int(e["out"][q])
                el.inv[q] = int(e["inv"][q])
            vec = e.get("vector", (0.0, 0.0))
            el.vec[0], el.vec[1] = float(vec[0]), float(vec[1])
            el.scalar = float(e.get("scalar", 0.0))
            nrm = e.get("normal", (0, 0))
            el.normal[0], el.normal[1] = int(nrm[0]), int(nrm[1])
        c = _Case()
        c.nx_pad, c.ny_pad = case.shape
        c.solid = self._solid.ctypes.data
        c.ghost = self._ghost.ctypes.data
        c.density = case.density.ctypes.data
        c.velocity = case.velocity.ctypes.data
        c.force = case.force_field.ctypes.data
        c.pop = case.pop.ctypes.data
        c.pop_new = case.pop_new.ctypes.data
        c.omega = float(case.omega)
        c.inv_cs_2 = float(case.const["inv_cs_2"])
        c.inv_cs_4 = float(case.const["inv_cs_4"])
        c.float_min = float(case.const["float_min"])
        for k in range(9):
            c.w[k] = float(case.const["w"][k])
        c.model = {"BGK": 0, "MRT": 1}[case.model]
        c.forcing = {None: 0, "guo_linear": 1,
                     "guo_second_order": 2}[case.forcing]
        c.force_source = 0 if case.force_source == "gravity" else 1
        c.gravity[0], c.gravity[1] = (float(case.gravity[0]),
                                      float(case.gravity[1]))
        c.x_periodic = int(case.x_periodic)
        c.y_periodic = int(case.y_periodic)
        c.n_elements = len(elems)
        c.elems = ctypes.addressof(self._elems)
        self._c = c

    def initialize(self):
        load().oracle_initialize(ctypes.byref(self._c))

    def run(self, steps, n_threads=0):
        load().oracle_run(ctypes.byref(self._c), int(steps), int(n_threads))
        self.case.time_step += steps
