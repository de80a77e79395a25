"""Link compatibility with the reference's caller: a C++ program compiled against headers declared the
way the reference's build declares them (Halide AOT headers with c_plus_plus_name_mangling for
backprojection / backprojection_cuda, C linkage for img_output_*, dft.h's C++ declarations) links
against libsarbp_hip.so and binds the mangled names (CMakeLists.txt:73-84, dft.h:5-6)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "halide-sar-app_amd")
LIB = os.path.join(PKG, "libsarbp_hip.so")
SRC = os.path.join(ROOT, "tests", "link", "reference_caller.cpp")

MANGLED = {
    "_Z14backprojectionP15halide_buffer_tS0_iidS0_S0_S0_S0_S0_",
    "_Z19backprojection_cudaP15halide_buffer_tS0_iidS0_S0_S0_S0_S0_",
    "_Z19backprojection_argvPPv",
    "_Z24backprojection_cuda_argvPPv",
    "_Z13dft_init_fftwm",
    "_Z16dft_destroy_fftwv",
}
C_NAMES = {"img_output_to_dB", "img_output_u8"}


def _need_tools():
    if not os.path.exists(LIB):
        pytest.skip("libsarbp_hip.so is not built")
    for tool in ("g++", "nm"):
        if shutil.which(tool) is None:
            pytest.skip(f"{tool} not found")


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    _need_tools()
    d = tmp_path_factory.mktemp("link")
    obj, exe = str(d / "caller.o"), str(d / "caller")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-c", SRC, "-o", obj], check=True)
    subprocess.run(["g++", obj, "-o", exe, "-L", PKG, "-lsarbp_hip", f"-Wl,-rpath,{PKG}"],
                   check=True)
    return obj, exe


def test_caller_binds_reference_linkage(caller):
    obj, _ = caller
    undef = set(subprocess.run(["nm", "-u", obj], check=True, capture_output=True,
                               text=True).stdout.split())
    assert MANGLED | C_NAMES <= undef, sorted((MANGLED | C_NAMES) - undef)
    # the C-linkage names of sarbp_hip.h are not what this caller binds
    assert "backprojection" not in undef and "dft_init_fftw" not in undef


def test_library_exports_both_linkages():
    _need_tools()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    want = MANGLED | C_NAMES | {"backprojection", "backprojection_cuda", "dft_init_fftw",
                                "dft_destroy_fftw"}
    assert want <= names, sorted(want - names)


def test_caller_runs_bounds_queries(caller):
    """Bounds queries through the C++ names and the argv form fill the required input regions
    (k_r {ns}, pos {3, np}, r {nu*nv, 3}) and return 0 without touching a device."""
    _, exe = caller
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.split("\n")
    for w in range(3):
        assert f"query {w} rv 0 k_r 24 pos 3,15 r 512,3" in lines, res.stdout
    assert "OK" in lines


def test_headers_coexist_with_c_header(tmp_path):
    """A translation unit may include the C++-linkage pipeline headers and sarbp_hip.h in either
    order; whichever comes first declares the name."""
    _need_tools()
    for order in (("backprojection.h", "sarbp_hip.h"), ("sarbp_hip.h", "backprojection.h")):
        src = tmp_path / "both.cpp"
        src.write_text("".join(f'#include "{h}"\n' for h in order) +
                       '#include "backprojection_cuda.h"\nint main() { return 0; }\n')
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], check=True)
