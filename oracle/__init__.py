"""CPU FP64 oracle of the reference simulator's hot path -- TEST INFRASTRUCTURE.

ctypes view of ``oracle/liboracle.so`` (built from ``hq_oracle.c`` by
``oracle/Makefile``).  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use it, always as the checker or the timed
CPU baseline; the product (hyquas_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_size_t, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PERGATE, BLAS = 1, 2
MODE_GROUP, MODE_BLAS, MODE_MIX = 1, 3, 4


class OraGate(ctypes.Structure):
    _fields_ = [("type", c_int32), ("target", c_int32), ("control", c_int32), ("control2", c_int32),
                ("mat", c_double * 8)]


GATE_DTYPE = np.dtype([("type", "<i4"), ("target", "<i4"), ("control", "<i4"), ("control2", "<i4"),
                       ("mat", "<f8", (8,))])


class PartitionParams(ctypes.Structure):
    _fields_ = [("mode", c_int), ("local_size", c_int), ("coalesce", c_int), ("min_mat", c_int),
                ("blas_mat_limit", c_int), ("mat_lo", c_int), ("mat_hi", c_int), ("window", c_int),
                ("single_cost", c_double * 22), ("ctr_cost", c_double * 22), ("group_overhead", c_double),
                ("blas_cost", c_double * 64), ("cutt_cost", c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        gp = POINTER(OraGate)
        dp = POINTER(c_double)
        ip = POINTER(c_int)
        L.ora_gate_make.argtypes = [c_int, c_int, c_int, c_int, dp, c_int, gp]
        L.ora_is_diagonal.argtypes = [c_int]
        L.ora_state_zero.argtypes = [dp, c_int]
        L.ora_state_zero.restype = None
        L.ora_apply_gate.argtypes = [dp, c_int, gp]
        L.ora_run_plain.argtypes = [dp, c_int, gp, c_int, c_int]
        L.ora_max_threads.argtypes = []
        L.ora_dense_unitary.argtypes = [c_int, gp, c_int, dp]
        L.ora_sharemem_group.argtypes = [dp, c_int, gp, c_int, c_uint64]
        L.ora_fuse_group.argtypes = [gp, c_int, ip, c_int, dp]
        L.ora_transpose.argtypes = [dp, dp, c_int, ip]
        L.ora_apply_dense_runs.argtypes = [dp, dp, c_int, dp, c_int]
        L.ora_partition_defaults.argtypes = [POINTER(PartitionParams), c_int, c_int]
        L.ora_partition_defaults.restype = None
        L.ora_partition.argtypes = [c_int, gp, c_int, POINTER(PartitionParams), c_int, ip, ip, ip,
                                    POINTER(c_uint64), ip, ip, ip, ip]
        L.ora_print_state.argtypes = [dp, c_int, ip, c_int, c_char_p, c_size_t]
        L.ora_print_state.restype = c_size_t
        L.ora_zero_wrapper.argtypes = [c_double]
        L.ora_zero_wrapper.restype = c_double
        L.ora_parse_qasm.argtypes = [c_char_p, ip, gp, c_int]
        _lib = L
    return _lib


def _gp(arr):
    arr = np.ascontiguousarray(arr, dtype=GATE_DTYPE)
    return arr, arr.ctypes.data_as(POINTER(OraGate))


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def gate(type_, target, control=-1, control2=-1, params=()):
    """One gate record with the oracle's matrices (gate.cpp factories)."""
    g = OraGate()
    p = (c_double * max(1, len(params)))(*params)
    rc = lib().ora_gate_make(int(type_), control2, control, target, p, len(params), byref(g))
    if rc:
        raise ValueError(f"invalid gate {type_} t={target} c={control} c2={control2} params={params}")
    arr = np.zeros(1, dtype=GATE_DTYPE)
    arr[0] = (g.type, g.target, g.control, g.control2, tuple(g.mat))
    return arr[0]


def zero_state(n):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    return psi


def run_plain(gates, n, psi=None, threads=0, copy=True):
    """Plain per-gate FP64 simulation (kernelSimple.cu semantics).  With
    copy=False a contiguous complex128 `psi` is updated in place (large states)."""
    if psi is None:
        psi = zero_state(n)
    if copy or psi.dtype != np.complex128 or not psi.flags.c_contiguous:
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    arr, p = _gp(gates)
    rc = lib().ora_run_plain(_dp(psi.view(np.float64)), n, p, len(arr), threads)
    if rc:
        raise ValueError(f"ora_run_plain failed ({rc})")
    return psi


def max_threads():
    return lib().ora_max_threads()


def dense_unitary(gates, n):
    N = 1 << n
    U = np.zeros(N * N, dtype=np.complex128)
    arr, p = _gp(gates)
    rc = lib().ora_dense_unitary(n, p, len(arr), _dp(U.view(np.float64)))
    if rc:
        raise ValueError("ora_dense_unitary failed")
    return U.reshape(N, N).T  # column-major storage -> U[row, col]


def sharemem_group(psi, n, gates, shared_mask):
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    arr, p = _gp(gates)
    rc = lib().ora_sharemem_group(_dp(psi.view(np.float64)), n, p, len(arr), shared_mask)
    if rc:
        raise ValueError(f"ora_sharemem_group failed ({rc})")
    return psi


def fuse_group(gates, pos, mat_qubit):
    K = 1 << mat_qubit
    m = np.zeros(K * K, dtype=np.complex128)
    arr, p = _gp(gates)
    posa = (c_int * len(pos))(*pos)
    rc = lib().ora_fuse_group(p, len(arr), posa, mat_qubit, _dp(m.view(np.float64)))
    if rc:
        raise ValueError("ora_fuse_group failed")
    return m  # column-major: m[col*K + row]


def transpose(psi, n, perm):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.empty_like(psi)
    pa = (c_int * n)(*perm)
    rc = lib().ora_transpose(_dp(psi.view(np.float64)), _dp(out.view(np.float64)), n, pa)
    if rc:
        raise ValueError("ora_transpose failed")
    return out


def dense_runs(psi, n, mat_colmajor, mat_qubit):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    m = np.ascontiguousarray(mat_colmajor, dtype=np.complex128)
    out = np.empty_like(psi)
    rc = lib().ora_apply_dense_runs(_dp(psi.view(np.float64)), _dp(out.view(np.float64)), n,
                                    _dp(m.view(np.float64)), mat_qubit)
    if rc:
        raise ValueError("ora_apply_dense_runs failed")
    return out


def partition_params(n, mode=MODE_MIX):
    p = PartitionParams()
    lib().ora_partition_defaults(byref(p), n, mode)
    return p


def partition(gates, n, params):
    """Reference partition: list of dicts (gate indices, backend, related, mat_qubit, perm),
    plus the final pos/layout."""
    arr, gp = _gp(gates)
    m = len(arr)
    cap = max(1, m)
    order = (c_int * cap)()
    start = (c_int * (cap + 1))()
    backend = (c_int * cap)()
    related = (c_uint64 * cap)()
    matq = (c_int * cap)()
    perms = (c_int * (cap * n))()
    fpos = (c_int * n)()
    flay = (c_int * n)()
    ng = lib().ora_partition(n, gp, m, byref(params), cap, order, start, backend, related, matq, perms, fpos, flay)
    if ng < 0:
        raise ValueError(f"ora_partition failed ({ng})")
    groups = []
    for g in range(ng):
        groups.append(dict(gates=list(order[start[g]:start[g + 1]]), backend=backend[g], related=related[g],
                           mat_qubit=matq[g], perm=list(perms[g * n:(g + 1) * n])))
    return groups, list(fpos), list(flay)


def print_state(psi, n, pos, show_schedule=False):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    pa = (c_int * n)(*pos)
    ln = lib().ora_print_state(_dp(psi.view(np.float64)), n, pa, 1 if show_schedule else 0, None, 0)
    buf = ctypes.create_string_buffer(ln + 1)
    lib().ora_print_state(_dp(psi.view(np.float64)), n, pa, 1 if show_schedule else 0, buf, ln + 1)
    return buf.value.decode()


def parse_qasm(text):
    n = c_int(-1)
    cnt = lib().ora_parse_qasm(text.encode(), byref(n), None, 0)
    if cnt < 0:
        raise ValueError("oracle: qasm parse error")
    arr = np.zeros(max(1, cnt), dtype=GATE_DTYPE)
    cnt2 = lib().ora_parse_qasm(text.encode(), byref(n), arr.ctypes.data_as(POINTER(OraGate)), cnt)
    assert cnt2 == cnt
    return n.value, arr[:cnt]
