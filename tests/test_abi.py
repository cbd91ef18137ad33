"""The C-ABI library loads, exports every symbol include/phsph.h declares,
and its host-side logic (defaults, argument validation) works without a GPU.
No compute entry point is called here."""
import ctypes as C
import os
import re

import pytest

from phantom_amd import abi, sph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "phsph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(phsph_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    L = sph.load_library()
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(sph.EXPORTED)


def test_params_struct_matches_header():
    # every field of phsph_params, in order
    txt = open(os.path.join(ROOT, "include", "phsph.h")).read()
    body = txt[txt.index("typedef struct phsph_params"):txt.index("} phsph_params;")]
    fields = []
    for line in body.splitlines()[1:]:
        line = re.sub(r"/\*.*?\*/", "", line).strip()
        m = re.match(r"(int32_t|double)\s+(.*);", line)
        if m:
            fields += [f.strip() for f in m.group(2).split(",")]
    assert fields == [n for n, _ in abi.Params._fields_]
    assert C.sizeof(abi.Params) == 16 * 4 + 26 * 8


def _struct_fields(txt, name):
    body = re.sub(r"/\*.*?\*/", "", txt[txt.index(f"typedef struct {name}"):txt.index(f"}} {name};")], flags=re.S)
    fields = []
    for line in body.splitlines()[1:]:
        m = re.match(r"\s*(?:int8_t|int32_t|float|double)\s*(.*);", line)
        if m:
            fields += [f.strip().lstrip("*").strip() for f in m.group(1).split(",")]
    return fields


@pytest.mark.parametrize("name,mirror", [("phsph_step_state", abi.StepState), ("phsph_step_clock", abi.StepClock),
                                         ("phsph_step_summary", abi.StepSummary)])
def test_step_structs_match_header(name, mirror):
    txt = open(os.path.join(ROOT, "include", "phsph.h")).read()
    assert _struct_fields(txt, name) == [n for n, _ in mirror._fields_]


@pytest.mark.parametrize("kernel", [abi.KERNEL_CUBIC, abi.KERNEL_QUINTIC])
def test_default_params_agree(kernel):
    L = sph.load_library()
    c = abi.Params()
    L.phsph_default_params(C.byref(c), kernel)
    py = abi.default_params(kernel)
    assert abi.params_to_dict(c) == abi.params_to_dict(py)


@pytest.mark.parametrize("field,value", [("kernel", 7), ("ieos", 3), ("massoftype_gas", 0.0), ("maxvxyzu", 5),
                                         ("tolh", -1.0)])
def test_invalid_params_rejected(field, value):
    p = abi.default_params()
    setattr(p, field, value)
    with pytest.raises(sph.PhsphError) as e:
        sph.Context(p)
    assert e.value.code == abi.ERR_INPUT


@pytest.mark.parametrize("acc", [-0.5, 1.5])
def test_tree_accuracy_range(acc):
    """tree_accuracy is read with gravity on, in 0..1 (neigh_kdtree.f90:345)."""
    p = abi.default_params()
    p.gravity = 1
    p.tree_accuracy = acc
    with pytest.raises(sph.PhsphError) as e:
        sph.Context(p)
    assert e.value.code == abi.ERR_INPUT


@pytest.mark.parametrize("extra", ["mhd", "dust"])
def test_gravity_only_with_hydro(extra):
    """Self-gravity is built for the hydro force; with MHD or one-fluid dust
    the context refuses it rather than computing without gravity."""
    p = abi.default_params()
    p.gravity = 1
    if extra == "mhd":
        p.mhd = 1
    else:
        p.use_dustfrac = 1
        p.ndustsmall = 1
    with pytest.raises(sph.PhsphError) as e:
        sph.Context(p)
    assert e.value.code == abi.ERR_INPUT


def test_oracle_library_loads():
    from oracle import oracle
    L = oracle.lib()
    for n in ("orc_create", "orc_maketree", "orc_densityiterate", "orc_cons2prim", "orc_force"):
        assert hasattr(L, n)
