"""CPU check of the specialised OShareMem kernel generator: the HIP source the
library emits for a group is compiled with g++ and executed with one host
thread per GPU thread (tests/cpp/jit_harness.cpp), then compared with the FP64
oracle.  No GPU involved."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from helpers import group_for_mask as _group_for_mask, rand_state as _rand_state, rel_err, to_oracle

HERE = os.path.dirname(os.path.abspath(__file__))


CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("n,k,seed,prec,rb,variant", [
    (10, 9, 31, 0, 0, 0), (11, 10, 32, 0, 3, 0), (12, 11, 33, 1, 0, 0), (11, 11, 34, 0, 5, 0),
    (14, 13, 35, 0, 4, 0), (15, 13, 36, 0, 5, 0), (14, 13, 37, 1, 5, 0),
    (11, 10, 32, 0, 3, 1), (14, 13, 35, 0, 4, 1), (15, 13, 36, 0, 5, 1), (14, 13, 37, 1, 5, 1),
    (15, 14, 38, 1, 0, 0), (16, 14, 39, 1, 6, 1), (15, 14, 40, 1, 6, 1), (13, 12, 41, 1, 4, 1),
    # first group of a run (bit 1): builds |0...0>, other tiles only store zeros
    (15, 13, 42, 0, 5, 2), (15, 13, 43, 0, 5, 3), (16, 14, 44, 1, 6, 3),
    # non-temporal stores only (bit 5), on the plain and the whole-line stages
    (14, 13, 45, 0, 5, 32), (15, 13, 46, 0, 5, 33), (16, 14, 47, 1, 6, 33), (15, 14, 48, 1, 6, 32)])
def test_generated_kernel_matches_oracle(hq, tmp_path, n, k, seed, prec, rb, variant):
    rs = np.random.default_rng(seed)
    rest = rs.choice(np.arange(3, n), k - 3, replace=False)
    mask = 0
    for q in [0, 1, 2] + [int(x) for x in rest]:
        mask |= 1 << q
    g = _group_for_mask(n, mask, 90, seed)
    src, coef, (tile, rbits, tbits) = hq.sharemem_source(n, g, mask, prec, rb, variant)
    assert tile == k and (rb == 0 or rbits == rb)
    if variant & 32:
        assert "__builtin_nontemporal_store" in src and "__builtin_nontemporal_load" not in src
    elif variant & 1:
        assert "__builtin_nontemporal_store" in src
        assert "__builtin_nontemporal_load" in src or variant & 2
    # non-temporal builtins and vector types (16-byte complex64 pairs) are
    # clang extensions: the host clang of ROCm runs them
    cxx = CLANG if os.path.exists(CLANG) else "g++"
    if cxx == "g++" and "ext_vector_type" in src:
        pytest.skip("needs ROCm clang++ for clang vector extensions")
    ksrc = tmp_path / "kernel.hip"
    ksrc.write_text(src)
    exe = tmp_path / "harness"
    subprocess.run([cxx, "-O1", "-std=c++20", "-pthread", "-Wno-unknown-attributes", f'-DGROUP_SRC="{ksrc}"',
                    os.path.join(HERE, "cpp", "jit_harness.cpp"), "-o", str(exe)], check=True)
    psi0 = _rand_state(n, seed)
    dt = np.complex128 if prec == 0 else np.complex64
    psi0.astype(dt).tofile(tmp_path / "in.bin")
    np.asarray(coef, dtype=np.float64).tofile(tmp_path / "coef.bin")
    subprocess.run([str(exe), str(n), str(k), str(1 << tbits), str(tmp_path / "in.bin"),
                    str(tmp_path / "coef.bin"), str(tmp_path / "out.bin")], check=True, timeout=120)
    got = np.fromfile(tmp_path / "out.bin", dtype=dt)
    if variant & 2:
        psi0 = np.zeros(1 << n, dtype=np.complex128)
        psi0[0] = 1
        assert "if (base != 0)" in src
    want = oracle.run_plain(g, n, psi0)
    assert rel_err(got, want) <= (1e-13 if prec == 0 else 2e-6)


def test_exchange_barriers(hq):
    # every LDS exchange after the first is preceded by a barrier that orders the
    # previous exchange's reads before its writes: 2 * (stages - 1) - 1 barriers
    n, mask = 12, (1 << 12) - 1
    g = _group_for_mask(n, mask, 120, 7)
    src, _, _ = hq.sharemem_source(n, g, mask)
    exchanges = src.count("lds[sw ^")
    stages = src.count("r0 = lds[") + 1
    assert stages >= 3
    assert src.count("__syncthreads") == 2 * (stages - 1) - 1
    assert exchanges > 0


def test_split_exchange_for_large_tiles(hq):
    # a 13-qubit complex128 tile (128 KiB) is exchanged in two 64 KiB halves
    n, mask = 15, (1 << 13) - 1
    g = _group_for_mask(n, mask, 150, 8)
    src, _, (tile, rb, tb) = hq.sharemem_source(n, g, mask, 0, 5)
    assert (tile, rb, tb) == (13, 5, 8)
    assert "__shared__ C lds[4096]" in src and " q0 = lds[" in src
    src64, _, _ = hq.sharemem_source(n, g, mask, 1, 5)    # complex64: 64 KiB, one exchange phase
    assert "__shared__ C lds[8192]" in src64 and " q0 = lds[" not in src64


def _zsums(n):
    idx = np.arange(1 << n)
    z = np.zeros(1 << n)
    zz = np.zeros(1 << n)
    for q in range(n):
        z += 1 - 2 * ((idx >> q) & 1)
    for q in range(n - 1):
        zz += 1 - 2 * (((idx >> q) ^ (idx >> (q + 1))) & 1)
    return z, zz


@pytest.mark.parametrize("n,k,seed,prec,rb,variant", [
    (12, 10, 51, 0, 0, 4 | 16), (13, 11, 52, 1, 0, 4 | 8 | 16), (14, 13, 53, 0, 5, 4 | 8 | 16),
    (14, 13, 54, 0, 5, 2 | 4 | 8 | 16), (15, 14, 55, 1, 6, 1 | 4 | 8 | 16), (12, 12, 56, 0, 4, 2 | 4 | 8 | 16)])
def test_batched_kernel_matches_oracle(hq, tmp_path, n, k, seed, prec, rb, variant):
    """The engine's batched form of a generated group kernel (item table,
    shared code for same-structure groups, fused readout, first-group form),
    run on the CPU for three circuits of one structure with their own angles:
    one source for all three, each state equal to the oracle's, each tile's
    partials summing to the circuit's norm, sum <Z_i> and sum <Z_i Z_i+1>."""
    from helpers import group_spec_for_mask, reangle
    rs = np.random.default_rng(seed)
    rest = rs.choice(np.arange(3, n), k - 3, replace=False) if k < n else np.arange(3, n)
    mask = 0
    for q in [0, 1, 2] + [int(x) for x in rest]:
        mask |= 1 << q
    base_spec = group_spec_for_mask(n, mask, 90, seed)
    specs = [reangle(base_spec, seed + 100 * i) for i in range(3)]
    srcs, coefs = [], []
    for sp in specs:
        src, coef, (tile, rbits, tbits) = hq.sharemem_source(n, to_oracle(sp), mask, prec, rb, variant)
        srcs.append(src)
        coefs.append(np.asarray(coef, dtype=np.float64))
    assert srcs[0] == srcs[1] == srcs[2], "same structure must give one kernel source"
    assert "hq_group_batch" in srcs[0]
    init, readout = bool(variant & 2), bool(variant & 8)
    cxx = CLANG if os.path.exists(CLANG) else "g++"
    if cxx == "g++" and "ext_vector_type" in srcs[0]:
        pytest.skip("needs ROCm clang++ for clang vector extensions")
    ksrc = tmp_path / "kernel.hip"
    ksrc.write_text(srcs[0])
    exe = tmp_path / "harness"
    subprocess.run([cxx, "-O1", "-std=c++20", "-pthread", "-Wno-unknown-attributes", f'-DGROUP_SRC="{ksrc}"',
                    os.path.join(HERE, "cpp", "jit_batch_harness.cpp"), "-o", str(exe)], check=True)
    dt = np.complex128 if prec == 0 else np.complex64
    psi0 = [rand_state_or_zero(n, seed + i, init) for i in range(3)]
    np.concatenate(psi0).astype(dt).tofile(tmp_path / "in.bin")
    np.concatenate(coefs).tofile(tmp_path / "coef.bin")
    subprocess.run([str(exe), str(n), str(k), str(1 << tbits), "3", "0", str(tmp_path / "in.bin"),
                    str(tmp_path / "coef.bin"), str(tmp_path / "out.bin"), str(tmp_path / "part.bin")],
                   check=True, timeout=300)
    got = np.fromfile(tmp_path / "out.bin", dtype=dt).reshape(3, -1)
    part = np.fromfile(tmp_path / "part.bin", dtype=np.float64).reshape(3, -1, 3)
    z, zz = _zsums(n)
    for i in range(3):
        start = np.zeros(1 << n, dtype=np.complex128)
        if init:
            start[0] = 1
        else:
            start = psi0[i]
        want = oracle.run_plain(to_oracle(specs[i]), n, start)
        assert rel_err(got[i], want) <= (1e-13 if prec == 0 else 2e-6)
        if readout:
            p = np.abs(want) ** 2
            tot = part[i].sum(axis=0)
            tol = 1e-12 if prec == 0 else 1e-5
            assert abs(tot[0] - p.sum()) <= tol
            assert abs(tot[1] - np.dot(p, z)) <= tol * n
            assert abs(tot[2] - np.dot(p, zz)) <= tol * n
    # mode bit 1: no store (the partials still come out)
    if readout:
        np.concatenate(psi0).astype(dt).tofile(tmp_path / "in.bin")
        subprocess.run([str(exe), str(n), str(k), str(1 << tbits), "3", "2", str(tmp_path / "in.bin"),
                        str(tmp_path / "coef.bin"), str(tmp_path / "out2.bin"), str(tmp_path / "part2.bin")],
                       check=True, timeout=300)
        same = np.fromfile(tmp_path / "out2.bin", dtype=dt).reshape(3, -1)
        assert np.array_equal(same, np.concatenate(psi0).astype(dt).reshape(3, -1))
        assert np.array_equal(np.fromfile(tmp_path / "part2.bin").reshape(3, -1, 3), part)


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_rotations_as_shears(hq, tmp_path, prec):
    """Shared-code kernels apply RY and RX as three shears whose scalar goes
    to the final scale op: one source for every angle, including the ones
    where the rotation's cosine is negative (|theta| > pi: the sign moves
    into the scalar), pi and 2 pi, and angles near zero (exactly zero makes
    the gate diagonal, another structure); each circuit's state equal to the
    oracle's."""
    from helpers import CZ, H, RX, RY, RZ
    n, k = 12, 10
    mask = (1 << k) - 1
    angle_sets = [
        [3.5, -3.9, 6.0, -6.2, 4.4, 0.3],
        [math.pi, -math.pi, 2 * math.pi, -2 * math.pi, 1e-12, 1e-9],
        [0.1, -0.7, 1.3, 2.9, -3.1, 3.1416],
    ]
    specs = []
    for ang in angle_sets:
        sp = [(H, 0, -1, ())]
        for i, th in enumerate(ang):
            sp.append((RY if i % 2 == 0 else RX, i % k, -1, (th,)))
            sp.append((RZ, (i + 3) % k, -1, (0.5 * th + 0.1,)))
            sp.append((CZ, (i + 1) % k, i % k, ()))
        sp += [(RY, 9, -1, (ang[0],)), (RX, 8, -1, (ang[1],))]
        specs.append(sp)
    srcs, coefs = [], []
    for sp in specs:
        src, coef, (tile, rbits, tbits) = hq.sharemem_source(n, to_oracle(sp), mask, prec, 0, 4 | 16)
        srcs.append(src)
        coefs.append(np.asarray(coef, dtype=np.float64))
    assert srcs[0] == srcs[1] == srcs[2], "rotations at any angle must give one kernel source"
    assert "lo.x += " in srcs[0] and "lo.x -= " in srcs[0], "RY and RX as shears"
    cxx = CLANG if os.path.exists(CLANG) else "g++"
    if cxx == "g++" and "ext_vector_type" in srcs[0]:
        pytest.skip("needs ROCm clang++ for clang vector extensions")
    ksrc = tmp_path / "kernel.hip"
    ksrc.write_text(srcs[0])
    exe = tmp_path / "harness"
    subprocess.run([cxx, "-O1", "-std=c++20", "-pthread", "-Wno-unknown-attributes", f'-DGROUP_SRC="{ksrc}"',
                    os.path.join(HERE, "cpp", "jit_batch_harness.cpp"), "-o", str(exe)], check=True)
    dt = np.complex128 if prec == 0 else np.complex64
    psi0 = [_rand_state(n, 70 + i) for i in range(3)]
    np.concatenate(psi0).astype(dt).tofile(tmp_path / "in.bin")
    np.concatenate(coefs).tofile(tmp_path / "coef.bin")
    subprocess.run([str(exe), str(n), str(k), str(1 << tbits), "3", "0", str(tmp_path / "in.bin"),
                    str(tmp_path / "coef.bin"), str(tmp_path / "out.bin"), str(tmp_path / "part.bin")],
                   check=True, timeout=300)
    got = np.fromfile(tmp_path / "out.bin", dtype=dt).reshape(3, -1)
    for i in range(3):
        want = oracle.run_plain(to_oracle(specs[i]), n, psi0[i].astype(dt).astype(np.complex128))
        assert rel_err(got[i], want) <= (1e-13 if prec == 0 else 2e-6)


def rand_state_or_zero(n, seed, init):
    # a first-group kernel never reads the state: hand it garbage
    if init:
        return np.full(1 << n, 7.0 + 3.0j)
    return _rand_state(n, seed)


RTC = os.path.join(os.path.dirname(HERE), "hyquas_amd", "hq_rtc")


@pytest.mark.skipif(not os.path.exists(RTC), reason="compile helper not built")
def test_compile_helper(hq, tmp_path):
    """The compile helper (what the library runs, one child process per
    concurrent compile) turns a generated group kernel into a gfx950 code
    object without a GPU, and reports a bad source with the compiler log."""
    n, mask = 14, (1 << 12) - 1
    g = _group_for_mask(n, mask, 60, 3)
    src, _, _ = hq.sharemem_source(n, g, mask, 1, 0, 4 | 8 | 16)
    (tmp_path / "k.hip").write_text(src)
    r = subprocess.run([RTC, str(tmp_path / "k.hip"), str(tmp_path / "k.co"), str(tmp_path / "k.log")], timeout=300)
    assert r.returncode == 0
    co = (tmp_path / "k.co").read_bytes()
    assert co[:4] == b"\x7fELF" and int.from_bytes(co[18:20], "little") == 224   # EM_AMDGPU
    assert b"hq_group_batch" in co
    (tmp_path / "bad.hip").write_text(src.replace("__syncthreads();", "__syncthreads(;", 1))
    r = subprocess.run([RTC, str(tmp_path / "bad.hip"), str(tmp_path / "bad.co"), str(tmp_path / "bad.log")],
                       timeout=300)
    assert r.returncode == 1 and not (tmp_path / "bad.co").exists()
    assert "error" in (tmp_path / "bad.log").read_text()
