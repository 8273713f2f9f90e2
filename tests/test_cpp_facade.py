"""The C++ facade (include/sr/Renderer.h) compiles against the C ABI and
drives the library with reference-style host code (tests/cpp/facade_scene.cpp)."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _build(tmp_path):
    exe = tmp_path / "facade_scene"
    lib = ROOT / "softrast_amd"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
           str(ROOT / "tests/cpp/facade_scene.cpp"), f"-L{lib}", "-lsrgpu", f"-Wl,-rpath,{lib}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_facade_compiles_and_fails_cleanly_without_gpu(tmp_path):
    exe = _build(tmp_path)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.strip() == "no-device" or res.stdout.startswith("drawn=")


@pytest.mark.gpu
def test_facade_renders_on_gpu(tmp_path):
    exe = _build(tmp_path)
    print("[facade] running reference-style C++ host program", flush=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("drawn=")
