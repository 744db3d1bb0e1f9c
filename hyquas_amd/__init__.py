"""hyquas_amd -- MI355X state-vector simulator with the HyQuas method.

Python view of the C ABI in ``include/hyquas_amd.h`` (ctypes).  The API
mirrors the reference's C++ one (src/circuit.h, src/gate.h): ``Gate.H(0)``,
``Circuit(n).addGate(...)``, ``compile()``, ``run()``, ``ampAt(i)``,
``printState()``; ``State`` exposes the device layer for tests and tools.

Every call goes to the HIP library ``libhyquas_amd.so`` built in-tree by
``python -m hyquas_amd.build``; there is no CPU fallback.  Importing works
without a GPU; creating a state or running a circuit then raises ``HQError``
with status ``HQ_ERR_NO_DEVICE``.
"""
from __future__ import annotations

import ctypes
import enum
import os
import weakref
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhyquas_amd.so")

# status codes (hq_status)
HQ_OK, HQ_ERR_ARG, HQ_ERR_NO_DEVICE, HQ_ERR_HIP, HQ_ERR_OOM, HQ_ERR_PARSE, HQ_ERR_STATE, HQ_ERR_UNSUPPORTED = range(8)
HQ_C128, HQ_C64 = 0, 1
HQ_BACKEND_NONE, HQ_BACKEND_PERGATE, HQ_BACKEND_BLAS = 0, 1, 2
HQ_MODE_SERIAL, HQ_MODE_GROUP, HQ_MODE_BLAS, HQ_MODE_MIX = 0, 1, 3, 4
KIND_PLAIN, KIND_SHAREMEM, KIND_TRANSPOSE, KIND_DENSE = 0, 1, 2, 3
HQ_OUT_EXPECT, HQ_OUT_TOPK = 1, 2
HQ_JOB_QUEUED, HQ_JOB_COMPILED, HQ_JOB_ADMITTED, HQ_JOB_DONE, HQ_JOB_FAILED = range(5)


class GateType(enum.IntEnum):
    """Values of GateType (src/gate.h:7-9)."""
    CCX = 0; CNOT = 1; CY = 2; CZ = 3; CRX = 4; CRY = 5; CU1 = 6; CRZ = 7
    U1 = 8; U2 = 9; U3 = 10; H = 11; X = 12; Y = 13; Z = 14; S = 15; SDG = 16
    T = 17; TDG = 18; RX = 19; RY = 20; RZ = 21; TOTAL = 22; ID = 23; GII = 24
    GZZ = 25; GOC = 26; GCC = 27


DIAGONAL_TYPES = {GateType.CZ, GateType.CU1, GateType.CRZ, GateType.U1, GateType.Z, GateType.S,
                  GateType.SDG, GateType.T, GateType.TDG, GateType.RZ}


class HQGate(ctypes.Structure):
    _fields_ = [("type", c_int32), ("target", c_int32), ("control", c_int32), ("control2", c_int32),
                ("mat", c_double * 8)]


GATE_DTYPE = np.dtype([("type", "<i4"), ("target", "<i4"), ("control", "<i4"), ("control2", "<i4"),
                       ("mat", "<f8", (8,))])
assert GATE_DTYPE.itemsize == ctypes.sizeof(HQGate) == 80


class HQError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[hq status {status}] {message}")
        self.status = status


_lib = None


def _declare(L):
    S = c_int  # hq_status
    sigs = {
        "hq_last_error": (c_char_p, []),
        "hq_device_count": (S, [POINTER(c_int)]),
        "hq_gate_make": (S, [c_int, c_int, c_int, c_int, POINTER(c_double), c_int, POINTER(HQGate)]),
        "hq_state_create": (S, [c_int, c_int, c_int, POINTER(c_void_p)]),
        "hq_state_destroy": (S, [c_void_p]),
        "hq_state_reset": (S, [c_void_p]),
        "hq_state_download": (S, [c_void_p, c_void_p, c_uint64]),
        "hq_state_upload": (S, [c_void_p, c_void_p, c_uint64]),
        "hq_state_get_amp": (S, [c_void_p, c_uint64, POINTER(c_double), POINTER(c_double)]),
        "hq_state_sync": (S, [c_void_p]),
        "hq_state_info": (S, [c_void_p, POINTER(c_int), POINTER(c_int)]),
        "hq_apply_gates_plain": (S, [c_void_p, POINTER(HQGate), c_int]),
        "hq_program_create_sharemem": (S, [c_void_p, POINTER(HQGate), c_int, c_uint64, POINTER(c_void_p)]),
        "hq_program_create_blas": (S, [c_void_p, POINTER(c_int), POINTER(c_double), c_int, POINTER(c_void_p)]),
        "hq_program_create_dense": (S, [c_void_p, POINTER(c_int), POINTER(c_double), c_int, POINTER(c_void_p)]),
        "hq_program_launch": (S, [c_void_p, c_void_p]),
        "hq_program_destroy": (S, [c_void_p]),
        "hq_program_info": (S, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "hq_program_specialize": (S, [c_void_p]),
        "hq_sharemem_source": (S, [c_int, c_int, POINTER(HQGate), c_int, c_uint64, c_int, c_int, c_char_p, c_size_t,
                                   POINTER(c_size_t), POINTER(c_double), c_size_t, POINTER(c_size_t),
                                   POINTER(c_int)]),
        "hq_program_is_specialized": (S, [c_void_p, POINTER(c_int)]),
        "hq_apply_sharemem": (S, [c_void_p, POINTER(HQGate), c_int, c_uint64]),
        "hq_apply_transpose": (S, [c_void_p, POINTER(c_int)]),
        "hq_apply_dense": (S, [c_void_p, POINTER(c_double), c_int]),
        "hq_measure_zero": (S, [c_void_p, c_int, POINTER(c_double)]),
        "hq_state_norm": (S, [c_void_p, POINTER(c_double)]),
        "hq_expect_z": (S, [c_void_p, POINTER(c_double), POINTER(c_double)]),
        "hq_expect_z_noalloc": (S, [c_void_p, POINTER(c_double), POINTER(c_double)]),
        "hq_state_has_buffer": (S, [c_void_p, POINTER(c_int)]),
        "hq_circuit_use_state": (S, [c_void_p, c_void_p]),
        "hq_engine_kernels": (S, [c_void_p, POINTER(c_int64)]),
        "hq_state_inner": (S, [c_void_p, c_void_p, POINTER(c_double), POINTER(c_double)]),
        "hq_profile_enable": (S, [c_void_p, c_int]),
        "hq_profile_read": (S, [c_void_p, POINTER(c_int), POINTER(c_double), c_int, POINTER(c_int)]),
        "hq_circuit_create": (S, [c_int, POINTER(c_void_p)]),
        "hq_circuit_destroy": (S, [c_void_p]),
        "hq_circuit_add_gate": (S, [c_void_p, POINTER(HQGate)]),
        "hq_circuit_num_gates": (S, [c_void_p, POINTER(c_int)]),
        "hq_circuit_num_qubits": (S, [c_void_p, POINTER(c_int)]),
        "hq_circuit_load_qasm": (S, [c_char_p, POINTER(c_void_p)]),
        "hq_circuit_parse_qasm": (S, [c_char_p, POINTER(c_void_p)]),
        "hq_circuit_set_option": (S, [c_void_p, c_char_p, c_int64]),
        "hq_circuit_load_cost_model": (S, [c_void_p, c_char_p]),
        "hq_circuit_compile": (S, [c_void_p]),
        "hq_circuit_partition": (S, [c_void_p]),
        "hq_circuit_run": (S, [c_void_p, c_int, c_int, POINTER(c_int64)]),
        "hq_circuit_amp_at": (S, [c_void_p, c_uint64, POINTER(c_double), POINTER(c_double)]),
        "hq_circuit_amp_at_gpu": (S, [c_void_p, c_uint64, POINTER(c_double), POINTER(c_double)]),
        "hq_circuit_print_state": (S, [c_void_p, c_int, c_char_p, c_size_t, POINTER(c_size_t)]),
        "hq_circuit_state": (S, [c_void_p, POINTER(c_void_p)]),
        "hq_circuit_final_pos": (S, [c_void_p, POINTER(c_int)]),
        "hq_circuit_schedule": (S, [c_void_p, POINTER(c_int), c_int, POINTER(c_int), POINTER(c_int),
                                    POINTER(c_uint64), POINTER(c_int), POINTER(c_int)]),
        "hq_circuit_schedule_live": (S, [c_void_p, c_int, POINTER(c_double), POINTER(c_int)]),
        "hq_circuit_schedule_order": (S, [c_void_p, POINTER(c_int), c_int]),
        "hq_circuit_last_times": (S, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                                      POINTER(c_double)]),
        "hq_circuit_launch_times": (S, [c_void_p, POINTER(c_int), POINTER(c_double), c_int, POINTER(c_int)]),
        "hq_circuit_get_gates": (S, [c_void_p, POINTER(HQGate), c_int, POINTER(c_int)]),
        "hq_circuit_add_random_layers": (S, [c_void_p, c_int, c_uint64]),
        "hq_circuit_add_hea_layers": (S, [c_void_p, c_int, c_uint64]),
        "hq_circuit_save_schedule": (S, [c_void_p, c_char_p]),
        "hq_circuit_load_schedule": (S, [c_void_p, c_char_p]),
        "hq_engine_create": (S, [c_int, c_uint64, POINTER(c_void_p)]),
        "hq_engine_destroy": (S, [c_void_p]),
        "hq_engine_set_option": (S, [c_void_p, c_char_p, c_int64]),
        "hq_engine_set_tile": (S, [c_void_p, c_int, c_int]),
        "hq_engine_get_tile": (S, [c_void_p, c_int, POINTER(c_int)]),
        "hq_engine_submit": (S, [c_void_p, c_void_p, c_int, POINTER(c_int64)]),
        "hq_engine_step": (S, [c_void_p, POINTER(c_int)]),
        "hq_engine_drain": (S, [c_void_p]),
        "hq_engine_poll": (S, [c_void_p, POINTER(c_int64), c_int, POINTER(c_int)]),
        "hq_engine_job_status": (S, [c_void_p, c_int64, POINTER(c_int)]),
        "hq_engine_result": (S, [c_void_p, c_int64, POINTER(c_double), POINTER(c_uint64), POINTER(c_double),
                                 POINTER(c_double)]),
        "hq_engine_release": (S, [c_void_p, c_int64]),
        "hq_engine_stats": (S, [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64),
                                POINTER(c_int64)]),
        "hq_engine_now": (S, [c_void_p, POINTER(c_double)]),
        "hq_batch_create": (S, [c_void_p, POINTER(c_void_p), c_int, c_int, POINTER(c_void_p)]),
        "hq_batch_run": (S, [c_void_p, POINTER(c_double)]),
        "hq_batch_results": (S, [c_void_p, POINTER(c_double), POINTER(c_uint64), POINTER(c_double)]),
        "hq_batch_info": (S, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_uint64)]),
        "hq_batch_pass_bytes": (S, [c_void_p, POINTER(c_uint64)]),
        "hq_batch_destroy": (S, [c_void_p]),
        "hq_state_topk": (S, [c_void_p, c_int, POINTER(c_uint64), POINTER(c_double)]),
        "hq_state_sample": (S, [c_void_p, c_int, c_uint64, POINTER(c_uint64)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return list(sigs)


EXPORTED = None


def lib():
    """The loaded native library (raises if it was not built)."""
    global _lib, EXPORTED
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hyquas_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        EXPORTED = _declare(L)
        _lib = L
    return _lib


def _check(status: int):
    if status != HQ_OK:
        msg = lib().hq_last_error()
        raise HQError(status, msg.decode() if msg else "")


def device_count() -> int:
    n = c_int(0)
    st = lib().hq_device_count(byref(n))
    return n.value if st == HQ_OK else 0


# ----------------------------------------------------------------------------- gates

class Gate:
    """A gate (src/gate.h:11-66).  Factories build the matrices in the library."""

    __slots__ = ("rec",)

    def __init__(self, rec: HQGate):
        self.rec = rec

    @staticmethod
    def make(type_: int, control2: int, control: int, target: int, params=()):
        rec = HQGate()
        p = (c_double * max(1, len(params)))(*params)
        _check(lib().hq_gate_make(int(type_), control2, control, target, p, len(params), byref(rec)))
        return Gate(rec)

    # factories, same argument order as src/gate.h:30-57
    @staticmethod
    def CCX(c1, c2, t): return Gate.make(GateType.CCX, c2, c1, t)
    @staticmethod
    def CNOT(c, t): return Gate.make(GateType.CNOT, -1, c, t)
    @staticmethod
    def CY(c, t): return Gate.make(GateType.CY, -1, c, t)
    @staticmethod
    def CZ(c, t): return Gate.make(GateType.CZ, -1, c, t)
    @staticmethod
    def CRX(c, t, a): return Gate.make(GateType.CRX, -1, c, t, (a,))
    @staticmethod
    def CRY(c, t, a): return Gate.make(GateType.CRY, -1, c, t, (a,))
    @staticmethod
    def CU1(c, t, lam): return Gate.make(GateType.CU1, -1, c, t, (lam,))
    @staticmethod
    def CRZ(c, t, a): return Gate.make(GateType.CRZ, -1, c, t, (a,))
    @staticmethod
    def U1(t, lam): return Gate.make(GateType.U1, -1, -1, t, (lam,))
    @staticmethod
    def U2(t, phi, lam): return Gate.make(GateType.U2, -1, -1, t, (phi, lam))
    @staticmethod
    def U3(t, th, phi, lam): return Gate.make(GateType.U3, -1, -1, t, (th, phi, lam))
    @staticmethod
    def H(t): return Gate.make(GateType.H, -1, -1, t)
    @staticmethod
    def X(t): return Gate.make(GateType.X, -1, -1, t)
    @staticmethod
    def Y(t): return Gate.make(GateType.Y, -1, -1, t)
    @staticmethod
    def Z(t): return Gate.make(GateType.Z, -1, -1, t)
    @staticmethod
    def S(t): return Gate.make(GateType.S, -1, -1, t)
    @staticmethod
    def SDG(t): return Gate.make(GateType.SDG, -1, -1, t)
    @staticmethod
    def T(t): return Gate.make(GateType.T, -1, -1, t)
    @staticmethod
    def TDG(t): return Gate.make(GateType.TDG, -1, -1, t)
    @staticmethod
    def RX(t, a): return Gate.make(GateType.RX, -1, -1, t, (a,))
    @staticmethod
    def RY(t, a): return Gate.make(GateType.RY, -1, -1, t, (a,))
    @staticmethod
    def RZ(t, a): return Gate.make(GateType.RZ, -1, -1, t, (a,))
    @staticmethod
    def ID(t): return Gate.make(GateType.ID, -1, -1, t)

    @property
    def type(self): return GateType(self.rec.type)
    @property
    def targetQubit(self): return self.rec.target
    @property
    def controlQubit(self): return self.rec.control
    @property
    def controlQubit2(self): return self.rec.control2
    def isControlGate(self): return self.rec.control != -1
    def isC2Gate(self): return self.rec.control2 != -1
    def isDiagonal(self): return self.type in DIAGONAL_TYPES

    @property
    def mat(self) -> np.ndarray:
        m = np.array(self.rec.mat[:], dtype=np.float64)
        return (m[0::2] + 1j * m[1::2]).reshape(2, 2)

    def __repr__(self):
        return f"Gate({self.type.name}, t={self.targetQubit}, c={self.controlQubit}, c2={self.controlQubit2})"


def gates_to_array(gates) -> np.ndarray:
    """Gate list -> structured array with the hq_gate layout."""
    arr = np.zeros(len(gates), dtype=GATE_DTYPE)
    for i, g in enumerate(gates):
        r = g.rec if isinstance(g, Gate) else g
        arr[i] = (r.type, r.target, r.control, r.control2, tuple(r.mat))
    return arr


def _gate_ptr(arr: np.ndarray):
    arr = np.ascontiguousarray(arr, dtype=GATE_DTYPE)
    return arr, arr.ctypes.data_as(POINTER(HQGate))


def sharemem_source(num_qubits: int, gates, shared_mask: int, precision: int = HQ_C128, reg_qubits: int = 0,
                    variant: int = 0):
    """Host-only: (source, coefficients, (tile, reg, thread qubits)) of the
    specialised kernel for one OShareMem group."""
    arr, p = _gate_ptr(gates_to_array(gates) if not isinstance(gates, np.ndarray) else gates)
    ln, nc = c_size_t(), c_size_t()
    shape = (c_int * 3)()
    _check(lib().hq_sharemem_source(num_qubits, precision, p, len(arr), shared_mask, reg_qubits, variant, None, 0,
                                    byref(ln), None, 0, byref(nc), shape))
    buf = ctypes.create_string_buffer(ln.value + 1)
    coef = np.zeros(max(1, nc.value), dtype=np.float64)
    _check(lib().hq_sharemem_source(num_qubits, precision, p, len(arr), shared_mask, reg_qubits, variant, buf,
                                    ln.value + 1, byref(ln), coef.ctypes.data_as(POINTER(c_double)), nc.value,
                                    byref(nc), shape))
    return buf.value.decode(), coef[: nc.value], tuple(shape)


# ----------------------------------------------------------------------------- device layer

class Program:
    def __init__(self, handle):
        self.handle = handle

    def info(self):
        t, s, o = c_int(), c_int(), c_int()
        _check(lib().hq_program_info(self.handle, byref(t), byref(s), byref(o)))
        return t.value, s.value, o.value

    def specialize(self):
        """Compile a kernel specialised to this group's structure (hiprtc)."""
        _check(lib().hq_program_specialize(self.handle))
        return self

    def is_specialized(self) -> bool:
        y = c_int()
        _check(lib().hq_program_is_specialized(self.handle, byref(y)))
        return bool(y.value)

    def close(self):
        if self.handle:
            lib().hq_program_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class State:
    """A device state vector (hq_state)."""

    def __init__(self, num_qubits: int, precision: int = HQ_C128, with_buffer: bool = False):
        h = c_void_p()
        _check(lib().hq_state_create(num_qubits, precision, 1 if with_buffer else 0, byref(h)))
        self.handle = h
        self.n = num_qubits
        self.precision = precision
        self.owned = True

    @classmethod
    def borrow(cls, handle, num_qubits: int, precision: int) -> "State":
        """View of a state owned elsewhere (e.g. Circuit.device_state()); never freed here."""
        s = cls.__new__(cls)
        s.handle, s.n, s.precision, s.owned = handle, num_qubits, precision, False
        return s

    @property
    def dtype(self):
        return np.complex128 if self.precision == HQ_C128 else np.complex64

    def close(self):
        if getattr(self, "handle", None) and getattr(self, "owned", False):
            lib().hq_state_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(lib().hq_state_reset(self.handle))

    def sync(self):
        _check(lib().hq_state_sync(self.handle))

    def download(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=self.dtype)
        _check(lib().hq_state_download(self.handle, out.ctypes.data_as(c_void_p), 1 << self.n))
        return out

    def upload(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=self.dtype)
        assert a.size == 1 << self.n
        _check(lib().hq_state_upload(self.handle, a.ctypes.data_as(c_void_p), 1 << self.n))

    def get_amp(self, idx: int) -> complex:
        re, im = c_double(), c_double()
        _check(lib().hq_state_get_amp(self.handle, idx, byref(re), byref(im)))
        return complex(re.value, im.value)

    def apply_plain(self, gates):
        arr, p = _gate_ptr(gates_to_array(gates) if not isinstance(gates, np.ndarray) else gates)
        _check(lib().hq_apply_gates_plain(self.handle, p, len(arr)))

    def program_sharemem(self, gates, shared_mask: int) -> Program:
        arr, p = _gate_ptr(gates_to_array(gates) if not isinstance(gates, np.ndarray) else gates)
        h = c_void_p()
        _check(lib().hq_program_create_sharemem(self.handle, p, len(arr), shared_mask, byref(h)))
        return Program(h)

    def program_blas(self, perm=None, mat=None, mat_qubits=0) -> Program:
        h = c_void_p()
        pp = (c_int * self.n)(*perm) if perm is not None else None
        if mat is not None:
            m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(-1))
            mp = m.view(np.float64).ctypes.data_as(POINTER(c_double))
        else:
            m, mp = None, None
        _check(lib().hq_program_create_blas(self.handle, pp, mp, mat_qubits, byref(h)))
        return Program(h)

    def program_dense(self, positions, mat) -> Program:
        """U (column-major, K = 2^len(positions)) on the qubits at `positions`
        (matrix index bit j = positions[j]), in place (hq_program_create_dense)."""
        h = c_void_p()
        pp = (c_int * len(positions))(*positions)
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(-1))
        _check(lib().hq_program_create_dense(self.handle, pp, m.view(np.float64).ctypes.data_as(POINTER(c_double)),
                                             len(positions), byref(h)))
        return Program(h)

    def launch(self, prog: Program):
        _check(lib().hq_program_launch(self.handle, prog.handle))

    def apply_sharemem(self, gates, shared_mask: int):
        arr, p = _gate_ptr(gates_to_array(gates) if not isinstance(gates, np.ndarray) else gates)
        _check(lib().hq_apply_sharemem(self.handle, p, len(arr), shared_mask))

    def apply_transpose(self, perm):
        pp = (c_int * self.n)(*perm)
        _check(lib().hq_apply_transpose(self.handle, pp))

    def apply_dense(self, mat_colmajor, mat_qubits: int):
        m = np.ascontiguousarray(np.asarray(mat_colmajor, dtype=np.complex128).reshape(-1))
        _check(lib().hq_apply_dense(self.handle, m.view(np.float64).ctypes.data_as(POINTER(c_double)), mat_qubits))

    def measure_zero(self, target: int) -> float:
        p = c_double()
        _check(lib().hq_measure_zero(self.handle, target, byref(p)))
        return p.value

    def norm2(self) -> float:
        p = c_double()
        _check(lib().hq_state_norm(self.handle, byref(p)))
        return p.value

    def expect_z(self):
        z = (c_double * self.n)()
        zz = (c_double * max(1, self.n - 1))()
        _check(lib().hq_expect_z(self.handle, z, zz))
        return np.array(z[:]), np.array(zz[: self.n - 1])

    def inner(self, other: "State") -> complex:
        re, im = c_double(), c_double()
        _check(lib().hq_state_inner(self.handle, other.handle, byref(re), byref(im)))
        return complex(re.value, im.value)

    def topk(self, k: int = 16):
        """The k most probable physical basis states: (indices, probabilities)."""
        idx = (c_uint64 * k)()
        prob = (c_double * k)()
        _check(lib().hq_state_topk(self.handle, k, idx, prob))
        return list(idx[:]), list(prob[:])

    def sample(self, shots: int, seed: int = 0) -> np.ndarray:
        """`shots` measurement samples (physical basis indices) drawn with the portable generator."""
        out = np.zeros(max(1, shots), dtype=np.uint64)
        _check(lib().hq_state_sample(self.handle, shots, seed, out.ctypes.data_as(POINTER(c_uint64))))
        return out[:shots]

    def profile(self, enable: bool):
        _check(lib().hq_profile_enable(self.handle, 1 if enable else 0))

    def profile_read(self, cap: int = 65536):
        kinds = (c_int * cap)()
        ms = (c_double * cap)()
        cnt = c_int()
        _check(lib().hq_profile_read(self.handle, kinds, ms, cap, byref(cnt)))
        k = min(cnt.value, cap)
        return list(zip(kinds[:k], ms[:k]))


# ----------------------------------------------------------------------------- circuit layer

class Circuit:
    """Circuit API of src/circuit.h:20-46 over the C ABI."""

    def __init__(self, numQubits: int = 0, _handle=None):
        if _handle is None:
            h = c_void_p()
            _check(lib().hq_circuit_create(numQubits, byref(h)))
            self.handle = h
            self.numQubits = numQubits
        else:
            self.handle = _handle
            self.numQubits = 0
        self._gates = None

    @staticmethod
    def from_qasm(text: str) -> "Circuit":
        h = c_void_p()
        _check(lib().hq_circuit_parse_qasm(text.encode(), byref(h)))
        c = Circuit(_handle=h)
        n = c_int()
        _check(lib().hq_circuit_num_qubits(h, byref(n)))
        c.numQubits = n.value
        return c

    @staticmethod
    def load_qasm(path: str) -> "Circuit":
        with open(path) as f:
            return Circuit.from_qasm(f.read())

    def close(self):
        if getattr(self, "handle", None):
            lib().hq_circuit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def addGate(self, g: Gate):
        _check(lib().hq_circuit_add_gate(self.handle, byref(g.rec)))
        return self

    def add_random_layers(self, layers: int, seed: int):
        _check(lib().hq_circuit_add_random_layers(self.handle, layers, seed))
        return self

    def add_hea_layers(self, layers: int, seed: int):
        _check(lib().hq_circuit_add_hea_layers(self.handle, layers, seed))
        return self

    def gates(self):
        cnt = c_int()
        _check(lib().hq_circuit_get_gates(self.handle, None, 0, byref(cnt)))
        arr = (HQGate * max(1, cnt.value))()
        _check(lib().hq_circuit_get_gates(self.handle, arr, cnt.value, byref(cnt)))
        return [arr[i] for i in range(cnt.value)]

    def gate_array(self) -> np.ndarray:
        return gates_to_array(self.gates())

    def set_option(self, key: str, value: int):
        _check(lib().hq_circuit_set_option(self.handle, key.encode(), int(value)))
        return self

    def save_schedule(self, path: str):
        """Write the partitioned schedule (compile-once cache)."""
        _check(lib().hq_circuit_save_schedule(self.handle, path.encode()))

    def load_schedule(self, path: str):
        """Use a saved schedule instead of partitioning; builds the device programs."""
        _check(lib().hq_circuit_load_schedule(self.handle, path.encode()))
        return self

    def load_cost_model(self, path: str):
        _check(lib().hq_circuit_load_cost_model(self.handle, path.encode()))

    def compile(self):
        _check(lib().hq_circuit_compile(self.handle))

    def partition(self):
        """Partition only (host, no device): schedule() / schedule_order() then work."""
        _check(lib().hq_circuit_partition(self.handle))

    def run(self, copy_back: bool = True, destroy: bool = True) -> int:
        us = c_int64()
        _check(lib().hq_circuit_run(self.handle, 1 if copy_back else 0, 1 if destroy else 0, byref(us)))
        return us.value

    def ampAt(self, idx: int) -> complex:
        re, im = c_double(), c_double()
        _check(lib().hq_circuit_amp_at(self.handle, idx, byref(re), byref(im)))
        return complex(re.value, im.value)

    def ampAtGPU(self, idx: int) -> complex:
        re, im = c_double(), c_double()
        _check(lib().hq_circuit_amp_at_gpu(self.handle, idx, byref(re), byref(im)))
        return complex(re.value, im.value)

    def stateText(self, show_schedule: bool = False) -> str:
        ln = c_size_t()
        _check(lib().hq_circuit_print_state(self.handle, 1 if show_schedule else 0, None, 0, byref(ln)))
        buf = ctypes.create_string_buffer(ln.value + 1)
        _check(lib().hq_circuit_print_state(self.handle, 1 if show_schedule else 0, buf, ln.value + 1, byref(ln)))
        return buf.value.decode()

    def printState(self, show_schedule: bool = False):
        print(self.stateText(show_schedule), end="")

    def final_pos(self):
        pos = (c_int * self.numQubits)()
        _check(lib().hq_circuit_final_pos(self.handle, pos))
        return list(pos)

    def device_state(self) -> c_void_p:
        h = c_void_p()
        _check(lib().hq_circuit_state(self.handle, byref(h)))
        return h

    def use_state(self, state: "State | None"):
        """Run on `state` (owned by the caller, kept allocated across runs and
        after close()) instead of a state of the circuit's own; None undoes it."""
        _check(lib().hq_circuit_use_state(self.handle, state.handle if state is not None else None))
        self._borrowed = state
        return self

    def state_view(self, precision: int = HQ_C128) -> State:
        """The device state after run(destroy=False), physical order."""
        return State.borrow(self.device_state(), self.numQubits, precision)

    def logical_state(self, precision: int = HQ_C128) -> np.ndarray:
        """Whole state in logical order (final layout undone), from the device."""
        phys = self.state_view(precision).download()
        pos = self.final_pos()
        idx = np.arange(1 << self.numQubits, dtype=np.int64)
        p = np.zeros_like(idx)
        for q in range(self.numQubits):
            p |= ((idx >> q) & 1) << pos[q]
        return phys[p]

    def schedule(self):
        n = c_int()
        _check(lib().hq_circuit_schedule(self.handle, byref(n), 0, None, None, None, None, None))
        k = n.value
        be, ng, mq, st = (c_int * max(1, k))(), (c_int * max(1, k))(), (c_int * max(1, k))(), (c_int * max(1, k))()
        rel = (c_uint64 * max(1, k))()
        _check(lib().hq_circuit_schedule(self.handle, byref(n), k, be, ng, rel, mq, st))
        live = (c_double * max(1, k))()
        m = c_int()
        _check(lib().hq_circuit_schedule_live(self.handle, k, live, byref(m)))
        return [dict(backend=be[i], num_gates=ng[i], related=rel[i], mat_qubit=mq[i], num_stages=st[i],
                     live_fraction=live[i]) for i in range(k)]

    def schedule_order(self):
        cnt = c_int()
        _check(lib().hq_circuit_num_gates(self.handle, byref(cnt)))
        order = (c_int * max(1, cnt.value))()
        _check(lib().hq_circuit_schedule_order(self.handle, order, cnt.value))
        return list(order[: cnt.value])

    def last_times(self):
        v = [c_double() for _ in range(4)]
        _check(lib().hq_circuit_last_times(self.handle, *[byref(x) for x in v]))
        return [x.value for x in v]

    def launch_times(self, cap: int = 65536):
        kinds = (c_int * cap)()
        ms = (c_double * cap)()
        cnt = c_int()
        _check(lib().hq_circuit_launch_times(self.handle, kinds, ms, cap, byref(cnt)))
        k = min(cnt.value, cap)
        return list(zip(kinds[:k], ms[:k]))


# ----------------------------------------------------------------------------- many circuits

class Engine:
    """Batching / serving engine (hq_engine_*): circuits are submitted, admitted
    while their device memory fits the budget, and advanced one group per
    round; ``step()`` launches a round, ``poll()`` returns finished jobs."""

    def __init__(self, precision: int = HQ_C128, hbm_budget: int = 1 << 34):
        h = c_void_p()
        _check(lib().hq_engine_create(precision, hbm_budget, byref(h)))
        self.handle = h
        self.precision = precision
        self._batches = weakref.WeakSet()

    def close(self):
        if getattr(self, "handle", None):
            for b in list(self._batches):
                b.close()
            lib().hq_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, key: str, value: int):
        _check(lib().hq_engine_set_option(self.handle, key.encode(), int(value)))
        return self

    def set_tile(self, num_qubits: int, tile: int):
        _check(lib().hq_engine_set_tile(self.handle, num_qubits, tile))
        return self

    def tile(self, num_qubits: int) -> int:
        t = c_int()
        _check(lib().hq_engine_get_tile(self.handle, num_qubits, byref(t)))
        return t.value

    def submit(self, circuit: "Circuit", outputs: int = HQ_OUT_EXPECT) -> int:
        j = c_int64()
        _check(lib().hq_engine_submit(self.handle, circuit.handle, outputs, byref(j)))
        return j.value

    def step(self) -> int:
        p = c_int()
        _check(lib().hq_engine_step(self.handle, byref(p)))
        return p.value

    def drain(self):
        _check(lib().hq_engine_drain(self.handle))

    def poll(self, cap: int = 4096):
        jobs = (c_int64 * cap)()
        cnt = c_int()
        _check(lib().hq_engine_poll(self.handle, jobs, cap, byref(cnt)))
        return list(jobs[: cnt.value])

    def status(self, job: int) -> int:
        st = c_int()
        _check(lib().hq_engine_job_status(self.handle, job, byref(st)))
        return st.value

    def result(self, job: int, topk: bool = False):
        """dict(norm, z, zz, times=(submitted, admitted, done)[, topk=(idx, prob)])."""
        e3 = (c_double * 3)()
        t3 = (c_double * 3)()
        idx = (c_uint64 * 16)() if topk else None
        prob = (c_double * 16)() if topk else None
        _check(lib().hq_engine_result(self.handle, job, e3, idx, prob, t3))
        r = dict(norm=e3[0], z=e3[1], zz=e3[2], times=tuple(t3))
        if topk:
            r["topk"] = (list(idx[:]), list(prob[:]))
        return r

    def release(self, job: int):
        _check(lib().hq_engine_release(self.handle, job))

    def stats(self):
        u, p, r = c_uint64(), c_uint64(), c_uint64()
        n = c_int64()
        _check(lib().hq_engine_stats(self.handle, byref(u), byref(p), byref(r), byref(n)))
        return dict(bytes_used=u.value, bytes_peak=p.value, bytes_reserved=r.value, rounds=n.value)

    def now(self) -> float:
        t = c_double()
        _check(lib().hq_engine_now(self.handle, byref(t)))
        return t.value

    def kernels(self) -> int:
        """Specialised batched kernels compiled so far (option "jit")."""
        k = c_int64()
        _check(lib().hq_engine_kernels(self.handle, byref(k)))
        return k.value

    def batch(self, circuits, outputs: int = HQ_OUT_EXPECT) -> "Batch":
        return Batch(self, circuits, outputs)


class Batch:
    """A fixed set of circuits captured as one hipGraph (hq_batch_*)."""

    def __init__(self, engine: Engine, circuits, outputs: int = HQ_OUT_EXPECT):
        self.engine = engine
        self.circuits = list(circuits)
        arr = (c_void_p * len(self.circuits))(*[c.handle.value for c in self.circuits])
        h = c_void_p()
        _check(lib().hq_batch_create(engine.handle, arr, len(self.circuits), outputs, byref(h)))
        self.handle = h
        self.outputs = outputs
        engine._batches.add(self)

    def close(self):
        if getattr(self, "handle", None):
            lib().hq_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self) -> float:
        ms = c_double()
        _check(lib().hq_batch_run(self.handle, byref(ms)))
        return ms.value

    def info(self):
        c, r, l = c_int(), c_int(), c_int()
        b = c_uint64()
        _check(lib().hq_batch_info(self.handle, byref(c), byref(r), byref(l), byref(b)))
        p = c_uint64()
        _check(lib().hq_batch_pass_bytes(self.handle, byref(p)))
        return dict(count=c.value, rounds=r.value, launches=l.value, device_bytes=b.value, pass_bytes=p.value)

    def results(self):
        """(expect[count, 3] = norm, sum Z, sum ZZ; topk (idx[count, 16], prob[count, 16]) or None)."""
        cnt = len(self.circuits)
        e = np.zeros((cnt, 3))
        if self.outputs & HQ_OUT_TOPK:
            idx = np.zeros((cnt, 16), dtype=np.uint64)
            prob = np.zeros((cnt, 16))
            _check(lib().hq_batch_results(self.handle, e.ctypes.data_as(POINTER(c_double)),
                                          idx.ctypes.data_as(POINTER(c_uint64)),
                                          prob.ctypes.data_as(POINTER(c_double))))
            return e, (idx, prob)
        _check(lib().hq_batch_results(self.handle, e.ctypes.data_as(POINTER(c_double)), None, None))
        return e, None
