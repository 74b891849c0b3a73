"""The C ABI libraries load without a GPU and export every entry point the
public headers declare."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

import lambda_amd._abi as abi


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lambda_[a-z0-9_]+)\s*\(", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("header,lib", [("lambda_hip.h", abi.HIP_LIB), ("lambda_host.h", abi.HOST_LIB)])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert names, header
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing


def test_libraries_load_without_gpu():
    C.CDLL(abi.HOST_LIB)
    C.CDLL(abi.HIP_LIB)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors agree with the C compiler's layout of the header."""
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lambda_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu", sizeof(lambda_hit), sizeof(lambda_render_params),'
                   ' sizeof(lambda_scene_view), offsetof(lambda_scene_view, camera), offsetof(lambda_render_params, seed),'
                   ' sizeof(lambda_progressive_stats), offsetof(lambda_progressive_stats, shadow_rays));'
                   'return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(abi.Hit), C.sizeof(abi.RenderParams), C.sizeof(abi.SceneView),
            abi.SceneView.camera.offset, abi.RenderParams.seed.offset, C.sizeof(abi.ProgressiveStats),
            abi.ProgressiveStats.shadow_rays.offset]
    assert got == want
