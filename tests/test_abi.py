"""The C ABI library: it loads without a GPU, exports every entry point the header declares, its records match
the reference's layout, and its host-side setup helpers equal the reference's and the oracle's bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

from beamlab_amd import abi, scenes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "beamlab_gpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(bl_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    lib = abi.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in abi.SIGNATURES, f"{n} has no ctypes signature"


def test_version_and_errors():
    lib = abi.load()
    assert lib.bl_abi_version() == 6
    assert lib.bl_error_string(abi.BL_ERR_ARG)
    s = abi.Settings()
    lib.bl_default_settings(ctypes.byref(s))
    d = abi.default_settings()
    assert bytes(s) == bytes(d)
    o = abi.BodyOpts()
    lib.bl_default_body_opts(ctypes.byref(o))
    assert bytes(o) == bytes(abi.default_body_opts())


def test_null_arguments_rejected():
    lib = abi.load()
    assert lib.bl_create(None, None) == abi.BL_ERR_ARG
    assert lib.bl_finalize(None, 3, None, 2, None, 0) == abi.BL_ERR_ARG
    assert lib.bl_box_from_quat(None, None, None, 0, None) == abi.BL_ERR_ARG


def test_layouts_match_reference_sizes():
    # static_asserts in oracle/ref_capi.cpp pin the same against the reference headers at build time
    assert abi.NODE.itemsize == 32 and abi.BEAM.itemsize == 52 and abi.NODE_INFO.itemsize == 20
    assert abi.BEAM.fields["type"][1] == 48 and abi.BEAM.fields["group"][1] == 50


def _raw_lattice(seed):
    b = scenes.lattice_body(9, 5, 4, 0.25, 2.0, (0.0, 1.0, 0.0), seed=seed)
    return b


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_finalize_stabilize_match(seed, ref_mod):
    b = _raw_lattice(seed)
    rng = np.random.default_rng(seed)
    beams = b.beams.copy()
    beams["k"] = rng.uniform(1e4, 1e7, len(beams)).astype(np.float32)
    beams["d"] = rng.uniform(10, 3000, len(beams)).astype(np.float32)
    nodes = b.nodes.copy()
    nodes["mass"][::7] = np.float32(0.0)
    ref_b, ref_ch = ref_mod.ref_finalize_stabilize(nodes, beams, 1, 0.0005)
    mine = beams.copy()
    lib = abi.load()
    assert lib.bl_finalize(abi.ptr(nodes), len(nodes), abi.ptr(mine), len(mine), None, 0) == 0
    ch = ctypes.c_int32(0)
    assert lib.bl_stabilize(abi.ptr(nodes), len(nodes), abi.ptr(mine), len(mine), 0.0005, 0.6, 0.5, ctypes.byref(ch)) == 0
    assert mine.tobytes() == ref_b.tobytes()
    assert ch.value == ref_ch > 0
    orc = ref_mod.oracle_stabilize(nodes, ref_mod.oracle_finalize(nodes, beams), 0.0005)
    assert orc[0].tobytes() == ref_b.tobytes() and orc[1] == ref_ch


def test_stabilize_budget_closed_form():
    """After stabilize every node's sum k dt^2/m is within the 0.6 budget (src/phys/softbody.cpp:85-110)."""
    b = scenes.config1_body()
    dt = 0.0005
    ksum = np.zeros(len(b.nodes))
    np.add.at(ksum, b.beams["a"], b.beams["k"].astype(np.float64))
    np.add.at(ksum, b.beams["b"], b.beams["k"].astype(np.float64))
    budget = ksum * dt * dt / b.nodes["mass"]
    assert budget.max() <= 0.6 * (1 + 1e-5)
    assert budget.max() > 0.5  # the interior nodes sit at the budget


@pytest.mark.parametrize("seed", range(6))
def test_box_from_quat_matches(seed, ref_mod):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    c, h = rng.uniform(-5, 5, 3), rng.uniform(0.1, 3, 3)
    mine = abi.np.zeros(1, abi.BOX)
    lib = abi.load()
    cc, hh, qq = (np.asarray(x, np.float32) for x in (c, h, q))
    assert lib.bl_box_from_quat(abi.ptr(cc), abi.ptr(hh), abi.ptr(qq), 4, abi.ptr(mine)) == 0
    ref = ref_mod.box_from_quat(c, h, q, 4, which="ref")
    orc = ref_mod.box_from_quat(c, h, q, 4, which="oracle")
    assert mine[0].tobytes() == ref.tobytes() == orc.tobytes()
