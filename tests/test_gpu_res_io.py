"""Result files and vector I/O against the reference's formats.

The expected bytes are composed here from the reference's format strings
(Dof_WriteFileRES0 / Dof_WriteFileRES, DofData.cpp:631-642, 741-767;
formatted_write_vector, Sparskit.cpp:1353-1359; LinAlg_WriteVector,
LinAlg_SPARSKIT.cpp:225-229) with Python's %-formatting, which rounds
correctly like C's printf; values read from text must equal a correctly
rounded conversion (fscanf "%lf", Python float()).  The vectors live on the
device, hence the gpu mark."""
import ctypes as C
import struct

import numpy as np
import pytest

from getdp_amd import linalg as LA
from getdp_amd._lib import GdpError, lib

pytestmark = pytest.mark.gpu

HEADER = "$ResFormat /* getdp_amd, {} */\n1.1 {}\n$EndResFormat\n"


def _values(n, seed=0):
    v = np.random.default_rng(seed).standard_normal(n) * 10.0 ** np.random.default_rng(seed + 1).integers(-300, 300, n)
    v[:8] = [0.0, -0.0, 1.0, -2.5, 0.1, 1.0 / 3.0, 123456789012345678.0, 5e-324]
    return v


def _block(num, t, ti, step, v, binary):
    head = "$Solution  /* DofData #%d */\n%d %.16g %.16g %d\n" % (num, num, t, ti, step)
    if binary:
        body = v.astype("<f8").tobytes() + b"\n"
    else:
        body = "".join("%.16g\n" % x for x in v).encode()
    return head.encode() + body + b"$EndSolution\n"


@pytest.mark.parametrize("binary", [False, True])
def test_res_write_bytes_and_read_back(gpu, tmp_path, binary):
    n = 5000
    v = _values(n)
    x = gpu.Vector(n, v)
    path = str(tmp_path / "u.res")
    gpu.res_write(path, x, create=True, binary=binary, num=0, time=0.25, time_imag=0.0, step=3)
    w = _values(n, 7)
    gpu.res_write(path, gpu.Vector(n, w), create=False, binary=binary, num=1, time=1e-3, step=4)
    gpu.res_write(path, gpu.Vector(n, -v), create=False, binary=binary, num=0, time=0.5, time_imag=2.0, step=5)
    want = (HEADER.format("binary" if binary else "ascii", int(binary)).encode() + _block(0, 0.25, 0.0, 3, v, binary)
            + _block(1, 1e-3, 0.0, 4, w, binary) + _block(0, 0.5, 2.0, 5, -v, binary))
    assert open(path, "rb").read() == want
    # what a reader gets: the text rounded to 16 digits, or the bits
    conv = (lambda a: a) if binary else (lambda a: np.array([float("%.16g" % t) for t in a]))
    y = gpu.Vector(n)
    info = gpu.res_read(path, y, num=0)  # the last solution of DofData 0
    assert info == {"time": 0.5, "time_imag": 2.0, "step": 5, "count": 2}
    assert np.array_equal(y.get().view(np.uint64), conv(-v).view(np.uint64))
    info = gpu.res_read(path, y, num=0, which=0)
    assert (info["step"], info["time"]) == (3, 0.25)
    assert np.array_equal(y.get().view(np.uint64), conv(v).view(np.uint64))
    info = gpu.res_read(path, y, num=1)
    assert info["count"] == 1 and info["step"] == 4
    assert np.array_equal(y.get().view(np.uint64), conv(w).view(np.uint64))
    assert gpu.res_read(path, y, num=-1)["count"] == 3
    with pytest.raises(GdpError):
        gpu.res_read(path, y, num=0, which=2)
    with pytest.raises(GdpError):
        gpu.res_read(path, y, num=7)


def test_res_read_version_1_0_and_free_layout(gpu, tmp_path):
    """Version 1.0 solution headers have no imaginary time; values may share
    lines (fscanf semantics)."""
    path = tmp_path / "old.res"
    path.write_text("$ResFormat /* GetDP 1.0, ascii */\n1.0 0\n$EndResFormat\n"
                    "$Solution  /* DofData #0 */\n0 1.5 7\n1 2 3\n  4.25e1\n-5\n$EndSolution\n")
    y = gpu.Vector(5)
    info = gpu.res_read(str(path), y, num=0)
    assert info == {"time": 1.5, "time_imag": 0.0, "step": 7, "count": 1}
    assert y.get().tolist() == [1.0, 2.0, 3.0, 42.5, -5.0]


def test_res_read_errors(gpu, tmp_path):
    path = tmp_path / "short.res"
    path.write_text("$ResFormat /* x */\n1.1 0\n$EndResFormat\n$Solution  /* DofData #0 */\n0 0 0 0\n1\n2\n")
    with pytest.raises(GdpError, match="premature"):
        gpu.res_read(str(path), gpu.Vector(5), num=0)
    with pytest.raises(GdpError, match="cannot open"):
        gpu.res_read(str(tmp_path / "missing.res"), gpu.Vector(5), num=0)


def test_vector_file_io(gpu, tmp_path):
    """gdp_vector_print / write / scan / read on C FILE streams."""
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    n = 3000
    v = _values(n, 3)
    x = gpu.Vector(n, v)
    for name, fn, binary in (("p.txt", "gdp_vector_print", False), ("w.bin", "gdp_vector_write", True)):
        path = str(tmp_path / name)
        f = libc.fopen(path.encode(), b"wb")
        assert getattr(lib(), fn)(x.h, f) == 0
        libc.fclose(f)
        data = open(path, "rb").read()
        want = v.astype("<f8").tobytes() + b"\n" if binary else "".join("%.16g\n" % t for t in v).encode()
        assert data == want
        y = gpu.Vector(n)
        f = libc.fopen(path.encode(), b"rb")
        assert getattr(lib(), "gdp_vector_read" if binary else "gdp_vector_scan")(y.h, f) == 0
        libc.fclose(f)
        got = y.get()
        ref = v if binary else np.array([float("%.16g" % t) for t in v])
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_res_write_read_at_scale(gpu, tmp_path):
    """811,200 values (the elasticity 64^3 solution size), ASCII and binary."""
    import time
    n = 811_200
    v = np.random.default_rng(9).standard_normal(n)
    x = gpu.Vector(n, v)
    for binary in (False, True):
        path = str(tmp_path / f"big{int(binary)}.res")
        t0 = time.perf_counter()
        gpu.res_write(path, x, binary=binary)
        t1 = time.perf_counter()
        y = gpu.Vector(n)
        gpu.res_read(path, y)
        t2 = time.perf_counter()
        print(f"res {'binary' if binary else 'ascii'}: write {1e3 * (t1 - t0):.0f} ms, read {1e3 * (t2 - t1):.0f} ms")
        if binary:
            assert np.array_equal(y.get(), v)
        else:
            assert np.max(np.abs(y.get() - v) / np.abs(v)) < 1e-15
