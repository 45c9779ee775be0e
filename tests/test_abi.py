"""CPU tests of the C ABI library: it loads, exports every entry point that
include/icbc_gpu.h and include/icbc.h declare, and rejects bad arguments
before touching a device.  No encode runs here (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

import icbc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    text = open(os.path.join(ROOT, "include", "icbc_gpu.h")).read()
    return sorted(set(re.findall(r"\b(icbc_(?:gpu|server|container)_\w+)\s*\(", text)))


def test_library_exports_every_declared_c_function():
    lib = ctypes.CDLL(icbc_amd.LIB_PATH)
    names = _declared_c_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"{name} declared in icbc_gpu.h but not exported"
    assert set(names) == set(icbc_amd._SIGNATURES), "python binding out of sync with the header"


def test_library_exports_reference_cpp_api():
    out = subprocess.check_output(["nm", "-D", "--defined-only", icbc_amd.LIB_PATH], text=True)
    # mangled names of the drop-in namespace icbc API (include/icbc.h)
    for sym in ("_ZN4icbc4initENS_7DecoderE",
                "_ZN4icbc12compress_bc1ENS_7QualityEPKfS2_S2_bbPv",
                "_ZN4icbc12compress_bc3ENS_7QualityEPKfS2_S2_bPv",
                "_ZN4icbc10decode_bc1EPKvPhNS_7DecoderE",
                "_ZN4icbc10decode_bc3EPKvPhNS_7DecoderE",
                "_ZN4icbc18evaluate_bc1_errorEPKhPKvNS_7DecoderE",
                "_ZN4icbc18evaluate_bc3_errorEPKhPKvbNS_7DecoderE"):
        assert sym in out, sym


def test_no_internal_symbols_leak():
    # only the C ABI and the namespace icbc API are exported: kernel launchers,
    # device stubs and server internals stay local to the library
    out = subprocess.check_output(["nm", "-D", "--defined-only", icbc_amd.LIB_PATH], text=True)
    exported = [l.split()[-1] for l in out.splitlines() if " T " in l or " W " in l]
    for s in exported:
        assert s.startswith(("icbc_gpu_", "icbc_server_", "icbc_container_", "_ZN4icbc")), s
    assert sum(s.startswith("_ZN4icbc") for s in exported) == 7


def test_programs_link_against_library(tmp_path):
    # a C caller of the batch ABI and a C++ caller of the namespace icbc API,
    # written as the reference's users write them, link against the library
    # (linked only: running them needs a GPU)
    libdir = os.path.dirname(icbc_amd.LIB_PATH)
    c = tmp_path / "caller.c"
    c.write_text('#include "icbc_gpu.h"\n#include <stddef.h>\n'
                 'int main(void){icbc_bc1_job j = {0}; if (icbc_gpu_init(ICBC_DECODER_D3D10)) return 1;\n'
                 'return icbc_gpu_compress_bc1_batch(&j, 1, NULL);}\n')
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(c),
                           "-L", libdir, "-licbc_amd", "-o", str(tmp_path / "caller_c")])
    cpp = tmp_path / "caller.cpp"
    cpp.write_text('#include "icbc.h"\nint main(){icbc::init();\n'
                   'float c[64]={0},w[16]={1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1},cw[3]={1,1,1}; unsigned char o[8];\n'
                   'float e = icbc::compress_bc1(icbc::Quality_Default, c, w, cw, false, false, o);\n'
                   'unsigned char rgba[64]; icbc::decode_bc1(o, rgba);\n'
                   'unsigned char b3[16] = {0}; icbc::decode_bc3(b3, rgba, icbc::Decoder_AMD);\n'
                   'return (int)(e + icbc::evaluate_bc1_error(rgba, o) + icbc::evaluate_bc3_error(rgba, b3, true));}\n')
    subprocess.check_call(["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(cpp),
                           "-L", libdir, "-licbc_amd", "-o", str(tmp_path / "caller_cpp")])


def test_library_links_hip_runtime_and_has_gfx950_code():
    out = subprocess.check_output(["ldd", icbc_amd.LIB_PATH], text=True)
    assert "libamdhip64" in out
    data = open(icbc_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_invalid_arguments_are_rejected_without_device():
    lib = icbc_amd.lib()
    assert lib.icbc_gpu_init(7) == icbc_amd.ERR_INVALID_ARGUMENT
    assert "decoder" in icbc_amd.last_error()
    assert lib.icbc_gpu_decode_bc1_batch(None, 0, None, 9, None) == icbc_amd.ERR_INVALID_ARGUMENT
    assert lib.icbc_gpu_evaluate_bc1_error_batch(None, None, 0, None, -1, None) == icbc_amd.ERR_INVALID_ARGUMENT
    assert lib.icbc_gpu_decode_bc3_batch(None, 0, None, 4, None) == icbc_amd.ERR_INVALID_ARGUMENT
    assert lib.icbc_gpu_evaluate_bc3_error_batch(None, None, 0, None, 1, -1, None) == icbc_amd.ERR_INVALID_ARGUMENT
    assert lib.icbc_gpu_version() == 10500


def test_job_struct_layout_matches_header():
    # icbc_bc1_job: 4 int32, 3 float, 1 uint32, uint64, 6 pointers
    assert ctypes.sizeof(icbc_amd.Job) == 16 + 16 + 8 + 6 * 8
    assert icbc_amd.Job.block_count.offset == 32
    assert icbc_amd.Job.colors.offset == 40
    assert icbc_amd.Job.image.offset == 80
    # icbc_image: pointer, 4 uint32, uint64
    assert ctypes.sizeof(icbc_amd.Image) == 8 + 16 + 8
    assert ctypes.sizeof(icbc_amd.PlanOptions) == 16


def test_c_header_compiles_as_c_and_cpp_header_as_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "icbc_gpu.h"\nint main(void){icbc_bc1_job j; (void)j; return 0;}\n')
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           "-c", str(src), "-o", str(tmp_path / "t.o")])
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "icbc.h"\nint main(){float c[64]={0},w[16]={0},cw[3]={1,1,1};'
                   'unsigned char o[8];(void)&icbc::compress_bc1;(void)c;(void)w;(void)cw;(void)o;return 0;}\n')
    subprocess.check_call(["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           "-c", str(cpp), "-o", str(tmp_path / "t2.o")])


def test_benchmark_programs_report_usage():
    # the C++ benchmark callers are built beside the library and check their
    # arguments before touching a device
    import subprocess
    from icbc_amd import build as hip_build
    for prog in (hip_build.BENCH_BIN, hip_build.SERVE_BIN):
        res = subprocess.run([prog], capture_output=True, text=True, timeout=60)
        assert res.returncode == 2 and "usage" in res.stderr, (prog, res.returncode, res.stderr)
