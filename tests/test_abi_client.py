"""The drop-in called from a plain C program (tests/abi/client.c): compiled by the C compiler against
include/beamlab_gpu.h alone and linked against libbeamlab_gpu.so, as the reference's host would use it. Without a
GPU: the library's defaults and the body preparation (finalize, stabilize) against the oracle, bit for bit. On the
GPU: a world stepped from C against the oracle, bit for bit."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from _cases import GROUND, _boxes
from beamlab_amd import abi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def client(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    abi.load()  # (the library is built)
    out = str(tmp_path_factory.mktemp("abi") / "client")
    libdir = os.path.join(ROOT, "beamlab_amd")
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "abi", "client.c"), "-L", libdir, "-lbeamlab_gpu", f"-Wl,-rpath,{libdir}",
                    "-o", out], check=True)
    return out


def _write(path, nodes, info, beams):
    with open(path, "wb") as f:
        f.write(np.array([len(nodes), len(beams)], np.uint32).tobytes())
        f.write(nodes.tobytes()), f.write(info.tobytes()), f.write(beams.tobytes())


def _read(path):
    raw = open(path, "rb").read()
    n, nb = np.frombuffer(raw[:8], np.uint32)
    o = 8
    nodes = np.frombuffer(raw[o:o + n * abi.NODE.itemsize], abi.NODE).copy()
    o += n * abi.NODE.itemsize
    info = np.frombuffer(raw[o:o + n * abi.NODE_INFO.itemsize], abi.NODE_INFO).copy()
    o += n * abi.NODE_INFO.itemsize
    beams = np.frombuffer(raw[o:o + nb * abi.BEAM.itemsize], abi.BEAM).copy()
    return nodes, info, beams


def test_client_defaults(client):
    out = subprocess.run([client, "info"], check=True, capture_output=True, text=True).stdout
    assert f"abi {abi.load().bl_abi_version()}" in out
    assert "dt 0.0005 collision_hz 120" in out and "hull_depth 0.3 face_skin 0.002" in out


def test_client_prepares_a_body_as_the_oracle(client, oracle_mod, tmp_path):
    b = scenes.lattice_body(5, 4, 3, 0.25, 2.0, (0.0, 0.3, 0.0), seed=12)
    beams = b.beams.copy()
    beams["L"] = beams["L0"] = 0
    _write(tmp_path / "in.bin", b.nodes, b.info, beams)
    res = subprocess.run([client, "prep", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True,
                         capture_output=True, text=True)
    _, _, got = _read(tmp_path / "out.bin")
    want, changed = oracle_mod.oracle_stabilize(b.nodes, oracle_mod.oracle_finalize(b.nodes, beams), 0.0005)
    assert got.tobytes() == want.tobytes()
    assert res.stdout.strip() == f"stabilized {changed}"


@pytest.mark.gpu
def test_client_steps_a_world_as_the_oracle(client, oracle_mod, tmp_path):
    b = scenes.lattice_body(6, 4, 3, 0.25, 2.0, (0.0, 0.4, 0.0), seed=13, eps=0.05)
    _write(tmp_path / "in.bin", b.nodes, b.info, b.beams)
    subprocess.run([client, "step", str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "600"], check=True,
                   capture_output=True, text=True, timeout=120)
    nodes, _, beams = _read(tmp_path / "out.bin")
    ow = oracle_mod.OracleWorld(abi.default_settings(), _boxes(GROUND))
    ow.add_body(b)
    ow.step_substeps(600)
    wn, wb = ow.read_body(0)
    assert nodes.tobytes() == wn.tobytes()
    assert beams.tobytes() == wb.tobytes()


def test_reference_side_adapter_compiles_against_the_reference_headers():
    """INTEGRATION.md's adapter (tests/abi/gpu_step_adapter.cpp), compiled against the reference's own headers where
    they lie: the types and fields it reads are the reference's."""
    ref = "/root/reference/src"
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(ref):
        pytest.skip("the reference's sources (or g++) are not here")
    tp = os.path.join(os.path.dirname(ref), "third_party", "imgui_perfmon")
    subprocess.run([gxx, "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-I", ref,
                    "-I", os.path.join(os.path.dirname(ref), "tests"), "-I", tp,
                    os.path.join(HERE, "abi", "gpu_step_adapter.cpp")], check=True)
