"""The C++ drop-in claim, compiled: a unit making every call the reference's
UI makes on its Renderer (glwidget.cpp, the picking tools), in the UI's own
spelling, builds against voxeltoy_amd/host/renderer.h with `using namespace
voxeltoy;`, links libvoxeltoy.so and runs on a renderer without a device
(no GPU calls)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_glwidget_calls_compile_and_run(vtapi, tmp_path):
    lib_dir = os.path.dirname(vtapi.library_path())
    exe = str(tmp_path / "glwidget_calls")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "voxeltoy_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "glwidget_calls.cpp"), "-o", exe,
                    "-L", lib_dir, "-lvoxeltoy", f"-Wl,-rpath,{lib_dir}"], check=True, capture_output=True,
                   text=True)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "scene_fall.vox"),
                        os.path.join(ROOT, "tests", "golden", "bunny.obj"), str(tmp_path / "out.pfm")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "OK"


def test_environment_image_formats(vtapi, tmp_path):
    """readImage (the library's loadImage): Radiance RGBE with run-length
    and flat scanlines in both row orders, 8-bit PPM and 16-bit PGM, decoded
    to the expected floats; other formats and truncated files refused."""
    lib_dir = os.path.dirname(vtapi.library_path())
    exe = str(tmp_path / "image_io")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "voxeltoy_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "image_io.cpp"), "-o", exe,
                    "-L", lib_dir, "-lvoxeltoy", f"-Wl,-rpath,{lib_dir}"], check=True, capture_output=True,
                   text=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "OK"
