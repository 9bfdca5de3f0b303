"""C-ABI checks that need no GPU: the library loads, exports every function
include/gpufft/gpufft.h declares, and its host-side planner validates
geometries and picks kernels as documented."""
import ctypes
import math
import os
import re

import pytest

import gpu_fft_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfft_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = declared_functions("gpufft/gpufft.h")
    assert "gfft_exec_r2c" in names and "gfft_plan_create" in names
    for n in names:
        assert hasattr(L, n), n
    assert L.gfft_version() == 100


def test_strerror_covers_codes():
    for code in range(0, 14):
        s = G.lib().gfft_strerror(code).decode()
        assert s and s != "unknown status"
    assert G.lib().gfft_strerror(99).decode() == "unknown status"


@pytest.mark.parametrize("args,status", [
    ((0, 8, 8, 1, 0, 0), 2),     # edge < 1
    ((8, 8, -4, 1, 0, 0), 2),
    ((8, 8, 8, 0, 0, 0), 2),     # procs < 1
    ((8, 8, 8, 1, 1, 0), 5),     # rank outside [0, procs)
    ((8, 8, 8, 2, -1, 0), 5),
    ((12, 8, 8, 8, 0, 0), 2),    # procs does not divide Nx
    ((8, 12, 8, 8, 0, 0), 2),    # ... nor Ny
    ((8, 8, 8, 1, 0, 7), 1),     # bad precision flag
    ((8, 8, 34, 1, 0, 0), 3),    # prime factor 17
    ((19, 8, 8, 1, 0, 1), 3),    # prime factor 19
    ((8, 8, 14, 1, 0, 0), 0),    # 7, 11, 13 are radices
    ((11, 13, 8, 1, 0, 1), 0),
    ((8, 8, 8, 1, 0, 0), 0),
    ((768, 768, 768, 1, 0, 0), 0),
    ((128, 1024, 1024, 8, 3, 0), 0),
])
def test_query_plan_validation(args, status):
    buf = ctypes.create_string_buffer(8192)
    assert G.lib().gfft_query_plan(*args, buf, len(buf)) == status


def test_query_plan_buffer_too_small():
    buf = ctypes.create_string_buffer(8)
    assert G.lib().gfft_query_plan(8, 8, 8, 1, 0, 0, buf, len(buf)) == 9


def test_null_arguments_rejected_without_gpu():
    L = G.lib()
    assert L.gfft_plan_create(None, 8, 8, 8, 1, 0, 0) == 1
    assert L.gfft_plan_destroy(None) == 1
    assert L.gfft_exec_r2c(None, None, None) == 1
    assert L.gfft_exec_c2r(None, None, None) == 1
    v = ctypes.c_size_t()
    assert L.gfft_plan_device_bytes(None, ctypes.byref(v)) == 1
    assert L.gfft_server_set_small_lane(None, ctypes.c_size_t(1)) == 1
    h = ctypes.c_void_p()
    assert L.gfft_server_create(ctypes.byref(h), ctypes.c_size_t(1 << 20), 2, 4) == 1  # unknown flag


def test_plan_choice_512_uses_register_kernels():
    d = G.query_plan(512, 512, 512, precision="f32")
    assert d["K"] == 257
    for ax in ("z", "y", "x"):
        assert d[ax]["engine"] == "register" and d[ax]["radices"] == [8, 8, 8]
    assert d["z"]["pass"] == "rows_two_for_one" and d["x"]["pass"] == "strided_columns"


def test_plan_choice_generic_and_identity():
    d = G.query_plan(1, 45, 27, precision="f64")
    assert d["x"]["pass"] == "identity"
    assert d["y"]["engine"] == "generic" and sorted(d["y"]["radices"]) == [3, 3, 5]
    assert d["z"]["engine"] == "generic" and d["z"]["radices"] == [3, 3, 3]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_every_serving_size_has_compiled_kernels(prec):
    # 2^a 3^b 5^c edges in [64, 512]: all three passes use the register engine
    sizes = [n for n in range(64, 513) if _smooth(n)]
    assert len(sizes) == 42
    for n in sizes:
        d = G.query_plan(n, n, n, precision=prec)
        for ax in ("z", "y", "x"):
            assert d[ax]["engine"] == "register", (n, prec, ax)
            assert math.prod(d[ax]["radices"]) == n


def _smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def test_twiddle_table_sizes():
    # stage s >= 1 stores p_s (R_s - 1) complex entries: 512 = 8*8*8 -> 8*7 + 64*7
    d = G.query_plan(512, 512, 512, precision="f32")
    assert d["z"]["twiddle_bytes"] == (8 * 7 + 64 * 7) * 8
    d = G.query_plan(512, 512, 512, precision="f64")
    assert d["z"]["twiddle_bytes"] == (8 * 7 + 64 * 7) * 16


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 25, 27, 45, 75, 81, 100, 125, 135,
                               225, 243, 250, 375, 405, 480, 500, 640, 720, 1000])
def test_all_235_smooth_sizes_plan(n):
    d = G.query_plan(n, n, n, precision="f32")
    for ax in ("z", "y", "x"):
        r = d[ax]["radices"]
        prod = 1
        for v in r:
            prod *= v
        assert prod == n and set(r) <= {2, 3, 4, 5, 8}


@pytest.mark.parametrize("n", [7, 11, 13, 14, 49, 77, 91, 143, 169, 343, 448, 1001, 1183])
def test_13_smooth_sizes_plan(n):
    d = G.query_plan(n, n, n, precision="f64")
    for ax in ("z", "y", "x"):
        r = d[ax]["radices"]
        assert math.prod(r) == n and set(r) <= {2, 3, 4, 5, 7, 8, 11, 13}


def test_tune_cache_text_round_trip_and_validation():
    """The tuning cache imports, exports and rejects malformed text without a
    GPU (entries are only consulted by gfft_plan_tune)."""
    G.clear_tune_cache()
    assert G.tune_cache_text() == ""
    text = ("# comment\n"
            "gfx950/256|p0:rows:512:262144 P8:T4:W256:K2:O1:A0:R8.8.8\n"
            "gfx950/256|p0:plane:128:128:128 fused:1\n")
    G.load_tune_cache_text(text)
    got = G.tune_cache_text().splitlines()
    assert sorted(got) == sorted(l for l in text.splitlines() if not l.startswith("#"))
    for bad in ("no-separator-or-value\n", "gfx950/256|key a b\n", "nodevicekey value\n"):
        with pytest.raises(G.GfftError) as e:
            G.load_tune_cache_text(bad)
        assert e.value.status == 1
    assert len(G.tune_cache_text().splitlines()) == 2  # a rejected text adds nothing
    need = ctypes.c_size_t(0)
    assert G.lib().gfft_tune_cache_export(None, 0, ctypes.byref(need)) == 0
    small = ctypes.create_string_buffer(4)
    assert G.lib().gfft_tune_cache_export(small, 4, None) == 9
    G.clear_tune_cache()
    assert G.tune_cache_text() == ""


def test_header_compiles_as_c_and_links_from_c(tmp_path):
    """include/gpufft/gpufft.h as plain C11 (gcc), linked against the library,
    running host-only entry points."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "gpu_fft_amd")
    exe = str(tmp_path / "abi_c_consumer")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c_abi", "abi_c_consumer.c"), "-L", libdir, "-lgpufft",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout and out.stdout.count("ok ") == 9


REF_TEST_CU = "/root/reference/GPU_FFT_TEST_SCRIPT/test.cu"


@pytest.mark.skipif(not os.path.exists(REF_TEST_CU), reason="reference sources not present")
def test_reference_caller_compiles_and_links_unchanged(tmp_path):
    """The reference's own caller (GPU_FFT_TEST_SCRIPT/test.cu, with its
    header.h) compiled in place, unmodified, against include/GPU_FFT/GPU_FFT.h
    and linked against libgpufft.so and MPI. Nothing is copied; the object and
    executable go to a temporary directory."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    mpi_inc = next((d for d in ("/opt/conda/include", "/usr/include/mpich", "/usr/include/openmpi")
                    if os.path.exists(os.path.join(d, "mpi.h"))), None)
    if mpi_inc is None:
        pytest.skip("no mpi.h")
    obj = str(tmp_path / "test.o")
    r = subprocess.run([hipcc, "-std=c++17", "-w", "--offload-arch=gfx950", "-c", REF_TEST_CU,
                        "-I", os.path.join(ROOT, "include"), "-I", mpi_inc, "-o", obj],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    mpi_lib = os.path.join(os.path.dirname(mpi_inc), "lib")
    if not os.path.exists(os.path.join(mpi_lib, "libmpi.so")):
        return
    libdir = os.path.join(ROOT, "gpu_fft_amd")
    exe = str(tmp_path / "GPU_FFT_test")
    # libmpi by full path: a -L on MPI's lib directory would also pick up any
    # older libstdc++ installed beside it
    libmpi = os.path.join(mpi_lib, "libmpi.so")
    r = subprocess.run([hipcc, "--offload-arch=gfx950", obj, "-L", libdir, "-lgpufft", "-Wl,-rpath," + libdir,
                        libmpi, "-Wl,-rpath-link," + mpi_lib, "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.getsize(exe) > 0
