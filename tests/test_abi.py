"""CPU tests of the C ABI: the library loads, exports every entry point that
include/snkpt.h declares, and rejects bad arguments without touching a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(snkpt_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    import snkpt
    lib = snkpt.load()
    declared = _declared()
    assert "snkpt_trace" in declared and "snkpt_render" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(snkpt.EXPORTS) <= declared


def test_struct_sizes_match_reference_layouts():
    import snkpt
    assert C.sizeof(snkpt.Instance) == 20  # InstanceData, RaytracingBufferSystem.h:14-28
    assert C.sizeof(snkpt.Material) == 64  # Material, TexMatBuffers.glsl:5-20
    assert C.sizeof(snkpt.Camera) == 240   # CommonUBO, CommonUBO.glsl:1-13 (std140 minus tail pad)


def test_settings_layout_matches_header(tmp_path):
    """The binding's snkpt_settings has the header's field offsets (gcc on
    include/snkpt.h prints them)."""
    import shutil
    import subprocess

    import snkpt
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    fields = [f[0] for f in snkpt.Settings._fields_]
    src = tmp_path / "off.c"
    src.write_text('#include <stdio.h>\n#include "snkpt.h"\nint main(void) {\n'
                   + "".join(f'printf("%zu\\n", offsetof(snkpt_settings, {f}));\n' for f in fields)
                   + 'printf("%zu\\n", sizeof(snkpt_settings));\nreturn 0;\n}\n')
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(snkpt.Settings, f).offset for f in fields] + [C.sizeof(snkpt.Settings)]
    assert got == want


def test_null_arguments_rejected_without_gpu():
    import snkpt
    lib = snkpt.load()
    assert lib.snkpt_create(0, None) == snkpt.ERR_INVALID_ARG
    assert lib.snkpt_build_accel(None, None) == snkpt.ERR_INVALID_ARG
    assert lib.snkpt_trace(None, None, 0, None, 0, None) == snkpt.ERR_INVALID_ARG
    assert lib.snkpt_scene_generate(9, 0, C.byref(C.POINTER(snkpt.Scene)())) == snkpt.ERR_INVALID_ARG
    assert lib.snkpt_scene_generate(5, 16, C.byref(C.POINTER(snkpt.Scene)())) == snkpt.ERR_INVALID_ARG


def test_oracle_library_is_separate_from_product():
    """The product library must not depend on the oracle."""
    so = open(os.path.join(ROOT, "snake-vk_amd", "libsnkpt.so"), "rb").read()
    assert b"liboracle" not in so and b"o_trace_brute" not in so
