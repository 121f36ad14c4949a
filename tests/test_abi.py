"""The C-ABI library loads and exports every symbol include/lvn_mesher.h declares (CPU only)."""
import ctypes as C
import os
import re

import pytest

from leven_amd import abi, mesher

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lvn_mesher.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lvn_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    assert declared_functions() == sorted(mesher.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(mesher.LIB_PATH):
        pytest.skip("library not built (python -m leven_amd.build)")
    L = C.CDLL(mesher.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_struct_layouts_match_reference_sizes():
    # CSGOperationInfo, MeshVertex, SeamNodeInfo are 48 bytes in the reference
    assert C.sizeof(abi.CsgOp) == 48
    assert C.sizeof(abi.Vertex) == 48
    assert C.sizeof(abi.SeamNode) == 48
    assert C.sizeof(abi.FieldDesc) == 48


def test_calls_before_init_fail_cleanly():
    if not os.path.exists(mesher.LIB_PATH):
        pytest.skip("library not built")
    L = mesher.lib()
    h = C.c_void_p()
    # no GPU work: the state check comes first
    if L.lvn_get_perm_table(None) != abi.LVN_ERR_INVALID:
        pytest.fail("null output must be rejected")
    state = L.lvn_ctx_create(64, 0, C.byref(h))
    assert state in (abi.LVN_ERR_STATE, abi.LVN_OK)
    if state == abi.LVN_OK:
        L.lvn_ctx_destroy(h)


def test_header_constants_match_binding():
    text = open(os.path.join(ROOT, "include", "lvn_mesher.h")).read()
    defines = {k: int(v) for k, v in re.findall(r"#define\s+(LVN_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", text)}
    assert defines["LVN_NUM_STAGES"] == mesher.NUM_STAGES
    for name in ["LVN_OK", "LVN_ERR_INVALID", "LVN_ERR_NOMEM", "LVN_ERR_CAPACITY", "LVN_ERR_DEVICE",
                 "LVN_ERR_STATE", "LVN_ERR_OVERFLOW", "LVN_MATERIAL_NONE", "LVN_MATERIAL_AIR",
                 "LVN_CTX_FP16_DENSITY", "LVN_CTX_PRECISE_QEF", "LVN_REQ_MESH", "LVN_REQ_EDIT",
                 "LVN_NOALLOC_MAX_EDGES", "LVN_NOALLOC_MAX_VOXELS", "LVN_NOALLOC_MAX_OPS"]:
        short = name[len("LVN_"):]
        assert defines[name] == getattr(abi, short if hasattr(abi, short) else name), name


def test_stage_names_cover_every_stage():
    if not os.path.exists(mesher.LIB_PATH):
        pytest.skip("library not built (python -m leven_amd.build)")
    names = [mesher.lib().lvn_batch_stage_name(i) for i in range(mesher.NUM_STAGES + 1)]
    assert all(n for n in names[:-1]) and names[-1] is None
    assert len(set(names[:-1])) == mesher.NUM_STAGES


def test_init_rejects_unsupported_fields():
    """Invalid field descriptors fail before any device work (no GPU needed)."""
    if not os.path.exists(mesher.LIB_PATH):
        pytest.skip("library not built (python -m leven_amd.build)")
    L = mesher.lib()
    for kw in [dict(height_octaves=-1), dict(cave_octaves=-1), dict(kind=3)]:
        d = abi.field_desc()
        for k, v in kw.items():
            setattr(d, k, v)
        assert L.lvn_init(C.byref(d), 1, 0) == abi.LVN_ERR_INVALID, kw
    assert L.lvn_init(C.byref(abi.field_desc()), 1, 256) == abi.LVN_ERR_INVALID  # material > 255
