"""The C ABI library: it builds, loads, exports every symbol include/lbm_hip.h
declares, and its host-only functions (the seeded generator) match independent
restatements.  CPU only: no GPU compute is called here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


@pytest.fixture(scope="module")
def lib():
    from lbm_amd import _native, build
    if not os.path.exists(_native.LIB_PATH):
        build.build()
    return _native.load()


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(lbm\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_functions():
    names = declared_functions()
    assert "lbm_create" in names and "lbm_step" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(lib):
    from lbm_amd import _native
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (\w+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    # and the ctypes binding covers all of them
    unbound = [n for n in declared_functions() if n not in _native.SIGNATURES]
    assert not unbound, unbound


def test_config_defaults(lib):
    from lbm_amd import _native as N
    cfg = N.lbm_config()
    assert lib.lbm_config_default(ctypes.byref(cfg), 400, 160) == 0
    assert cfg.nx == 400 and cfg.ny == 160
    assert abs(cfg.omega - 1.0 / 0.8) < 1e-7   # lib.rs:39
    assert abs(cfg.u_init - 0.1) < 1e-7         # lbm.rs:740
    assert cfg.precision == 0 and cfg.boundary == 0


def test_create_rejects_bad_config_without_gpu_work(lib):
    from lbm_amd import _native as N
    cfg = N.lbm_config()
    lib.lbm_config_default(ctypes.byref(cfg), 2, 160)
    h = ctypes.c_void_p()
    assert lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == N.LBM_EINVAL
    assert b"nx and ny" in lib.lbm_last_error(None)
    lib.lbm_config_default(ctypes.byref(cfg), 64, 64)
    cfg.precision = 7
    assert lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == N.LBM_EINVAL


def test_null_handle_is_an_error(lib):
    from lbm_amd import _native as N
    assert lib.lbm_step(None, 1) == N.LBM_EINVAL
    assert lib.lbm_destroy(None) == N.LBM_EINVAL


def test_splitmix64_published_vectors(lib):
    """SplitMix64 seeded with 0 (Vigna's reference splitmix64.c)."""
    from lbm_amd import splitmix64
    assert splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def _py_gen_circles(nx, ny, seed, n, rmin, rmax):
    M = (1 << 64) - 1
    s = seed

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def below(m):
        return ((nxt() >> 32) * m) >> 32

    m = np.zeros((ny, nx), np.uint8)
    m[0, :] = m[-1, :] = 1
    yy, xx = np.mgrid[0:ny, 0:nx]
    for _ in range(n):
        r = rmin + below(rmax - rmin + 1)
        cx = r + 1 + below(nx - r - 1)
        cy = below(ny)
        m[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 1
    return m


@pytest.mark.parametrize("nx,ny,seed,n,rmin,rmax", [(400, 160, 1, 10, 4, 20), (1024, 512, 42, 200, 8, 64),
                                                    (300, 100, 7, 5, 2, 3)])
def test_gen_circles_matches_restatement(lib, nx, ny, seed, n, rmin, rmax):
    from lbm_amd import gen_circles
    got = gen_circles(nx, ny, seed, n, rmin, rmax)
    exp = _py_gen_circles(nx, ny, seed, n, rmin, rmax)
    assert np.array_equal(got, exp)
    assert got[:, 0].sum() == 2  # column 0 stays free apart from the two wall rows


def test_gen_circles_solid_fraction_config2(lib):
    """Config 2 obstacle field: 200 circles, r ~ U{8..64}.  Solid fraction of the
    interior is a few percent at 8192x4096 and ~10% at 4096x2048."""
    from lbm_amd import gen_circles
    for (nx, ny), lo, hi in [((4096, 2048), 0.06, 0.16), ((8192, 4096), 0.015, 0.05)]:
        m = gen_circles(nx, ny, 2024, 200, 8, 64, walls=False)
        frac = m.mean()
        assert lo < frac < hi, frac


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with status 2 and says why."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_save_ppm_writes_a_binary_ppm(tmp_path):
    from lbm_amd import save_ppm
    rgb = np.zeros((2, 3, 3), np.float32)
    rgb[0, 0] = (1.0, 0.5, 0.0)
    path = tmp_path / "f.ppm"
    save_ppm(path, rgb)
    data = path.read_bytes()
    assert data.startswith(b"P6 3 2 255\n")
    body = data[len(b"P6 3 2 255\n"):]
    assert len(body) == 18 and body[:3] == bytes([255, 128, 0])


CLI = os.path.join(os.path.dirname(__file__), "..", "lbm_amd", "lbm_cli")


def test_cpp_cli_builds_and_parses_arguments_without_a_device():
    """The C++ driver over include/lbm_hip.hpp: usage, argument errors reported
    as errors (exit 1 / 2), no crash; nothing here touches a GPU."""
    assert os.path.exists(CLI), "lbm_amd/lbm_cli is built with the library"
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage: lbm_cli" in r.stdout
    r = subprocess.run([CLI, "--precision", "f8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--precision is f32, f16s or f16e" in r.stderr
    r = subprocess.run([CLI, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown option" in r.stderr
