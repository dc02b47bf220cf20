"""The C-ABI boundary on the CPU: both libraries load, export every function that include/*.h
declares, and the numpy / ctypes mirrors of the POD structs have the C layout (sizes and field
offsets as gcc lays out the header). No GPU call is made here."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import sw_abi
import sw_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "square-wheel_amd")


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"^[A-Za-z_][\w \t\*]*?\b(sw\w*)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("sw_gpu.h", "libsw_gpu.so"), ("sw_server.h", "libsw_gpu.so"),
                                        ("sw_host.h", "libsw_host.so")])
def test_library_exports_every_declared_function(header, lib):
    names = declared_functions(header)
    assert len(names) >= 8, names
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), "build the libraries first (python square-wheel_amd/build.py)"
    so = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_gpu_library_has_no_oracle_dependency():
    # the product library must not link the CPU oracle (it is test infrastructure only)
    out = subprocess.run(["ldd", os.path.join(PKG, "libsw_gpu.so")], capture_output=True, text=True).stdout
    assert "oracle" not in out
    out = subprocess.run(["ldd", os.path.join(PKG, "libsw_host.so")], capture_output=True, text=True).stdout
    assert "oracle" not in out


STRUCTS = {
    "sw_texel": sw_abi.TEXEL_DTYPE,
    "sw_texture_level": sw_abi.LEVEL_DTYPE,
    "sw_directional_lightmap_element": sw_abi.DIRLM_DTYPE,
    "sw_light": sw_abi.LIGHT_DTYPE,
    "sw_surface_desc": sw_abi.SURFACE_DTYPE,
    "sw_draw": sw_abi.DRAW_DTYPE,
    "sw_view": sw_abi.VIEW_DTYPE,
    "sw_model_instance": sw_abi.MODEL_INSTANCE_DTYPE,
    "sw_model_vertex": sw_abi.MODEL_VERTEX_DTYPE,
    "sw_texture_op": sw_abi.TEXTURE_OP_DTYPE,
    "sw_light_shadow": sw_abi.LIGHT_SHADOW_DTYPE,
    "sw_depth_draw": sw_abi.DEPTH_DRAW_DTYPE,
    "sw_depth_target": sw_abi.DEPTH_TARGET_DTYPE,
    "sw_wave_point": sw_abi.WAVE_POINT_DTYPE,
    "sw_heat_point": sw_abi.HEAT_POINT_DTYPE,
    "sw_water_texture_params": np.dtype([(n, "<u4" if t is not ctypes.c_float else "<f4") for n, t in sw_abi.WaterTextureParams._fields_]),
    "sw_portal_mix": sw_abi.PORTAL_MIX_DTYPE,
    "swh_portal_info": sw_host.PORTAL_INFO_DTYPE,
    "swh_view_portal": sw_host.VIEW_PORTAL_DTYPE,
}


CTYPES_STRUCTS = [(sw_abi.MapDesc, "sw_map_desc"), (sw_abi.Batch, "sw_batch"), (sw_abi.Stats, "sw_stats"),
                  (sw_abi.ViewC, "sw_view"), (sw_abi.FrameRequest, "sw_frame_request"), (sw_abi.Completion, "sw_completion"),
                  (sw_abi.ServerStats, "sw_server_stats"), (sw_abi.ModelMesh, "sw_model_mesh"),
                  (sw_abi.BloomParams, "sw_bloom_params"), (sw_abi.ExposureParams, "sw_exposure_params")]


@pytest.fixture(scope="module")
def c_layout():
    """sizeof / offsetof of every struct field, printed by a C program compiled from the header."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sw_server.h"', '#include "sw_host.h"', "int main(void) {"]
    for name, dt in STRUCTS.items():
        lines.append(f'printf("{name} @sizeof %zu\\n", sizeof({name}));')
        for field in dt.names:
            lines.append(f'printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    for cls, cname in CTYPES_STRUCTS:
        lines.append(f'printf("{cname} @sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-I", INCLUDE, "-o", exe, src], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    layout = {}
    for line in out.splitlines():
        struct, field, value = line.split()
        layout[(struct, field)] = int(value)
    return layout


@pytest.mark.parametrize("name", sorted(STRUCTS))
def test_numpy_dtype_matches_c_layout(c_layout, name):
    dt = STRUCTS[name]
    assert dt.itemsize == c_layout[(name, "@sizeof")]
    for field in dt.names:
        assert dt.fields[field][1] == c_layout[(name, field)], field


@pytest.mark.parametrize("cls,cname", CTYPES_STRUCTS)
def test_ctypes_struct_matches_c_layout(c_layout, cls, cname):
    assert ctypes.sizeof(cls) == c_layout[(cname, "@sizeof")]
    for field, _ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[(cname, field)], field


def test_enums_match_header():
    text = open(os.path.join(INCLUDE, "sw_gpu.h")).read()

    def value(name):
        m = re.search(r"\b" + name + r"\s*=\s*([^,/\n]+)", text) or re.search(r"#define\s+" + name + r"\s+([^/\n]+)", text)
        assert m, name
        expr = re.sub(r"(\d+)u\b", r"\1", m.group(1).strip())
        assert re.fullmatch(r"[\d\s<|()]+", expr), expr
        return int(eval(expr))

    for name in dir(sw_abi):
        if re.match(r"^(BLEND|TC|DRAW|TEXSRC|SURF)_[A-Z_]+$", name) and not name.endswith("_DTYPE"):
            assert getattr(sw_abi, name) == value("SW_" + name), name


def test_product_path_fails_loudly_without_library(tmp_path, monkeypatch):
    # a missing HIP library is an error, never a silent CPU fallback
    monkeypatch.setattr(sw_abi, "GPU_LIB_PATH", str(tmp_path / "libsw_gpu.so"))
    monkeypatch.setattr(sw_abi, "_gpu_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        sw_abi.load_gpu()
    with pytest.raises(RuntimeError):
        sw_abi.Context(0)


def test_gpu_symbol_table_covers_header():
    # every declared entry point has a typed ctypes binding
    assert sorted(sw_abi.GPU_SYMBOLS) == sorted(declared_functions("sw_gpu.h") + declared_functions("sw_server.h"))
