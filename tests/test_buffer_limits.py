"""Which systems the matvec kernel may address through buffer descriptors (orctm_amd/csrc/nh_kernels.h,
compiled for the host): every inner array (nz planes of the padded pitch x ny in the inner precision)
must hold at most 2^31 bytes, so that a lane's 32-bit byte offset stays below the out-of-range offset
2^31 that masked lanes take. Larger systems run the 64-bit-pointer form (same bits)."""
import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    so = str(tmp_path_factory.mktemp("buffers") / "libbuffers.so")
    subprocess.run([hipcc, "-std=c++17", "-O1", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "buffers_host.cpp")],
                   check=True, capture_output=True)
    L = ctypes.CDLL(so)
    L.fits.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int]
    L.fits.restype = ctypes.c_int
    L.max_bytes.restype = ctypes.c_ulonglong
    return L


def pitch(nx):
    return (nx + 31) // 32 * 32


def test_limit_is_2_pow_31(lib):
    assert lib.max_bytes() == 1 << 31


@pytest.mark.parametrize("fp32, elem", [(1, 4), (0, 8)])
def test_boundary(lib, fp32, elem):
    # nz planes of `plane` elements: exactly 2^31 bytes fits, one more plane element per level does not
    nz = 64
    plane = (1 << 31) // (nz * elem)
    assert plane * nz * elem == 1 << 31
    assert lib.fits(plane, nz, fp32) == 1
    assert lib.fits(plane + 1, nz, fp32) == 0


def test_workload_sizes(lib):
    # the largest serving request, 2048x1536x80: 2.01 GB per FP64 array, 1.01 GB per FP32 array
    plane = pitch(2048) * 1536
    assert lib.fits(plane, 80, 0) == 1 and lib.fits(plane, 80, 1) == 1
    # 2048x2048x65 (tests/test_gpu_configs.py): 2.18 GB per FP64 array (pointers), 1.09 GB in FP32
    plane = pitch(2048) * 2048
    assert lib.fits(plane, 65, 0) == 0 and lib.fits(plane, 65, 1) == 1
    # the regional tile and the transect
    assert lib.fits(pitch(1024) * 768, 60, 0) == 1
    assert lib.fits(pitch(512) * 4, 100, 1) == 1
