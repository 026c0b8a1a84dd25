"""bl_fem_solve_noalloc (the substitutions with the library's static device scratch, device pointers in and out)
against bl_fem_solve on the same factorization, bit for bit: two calls in a row, the second on a smaller system, then
the first system again (the scratch keeps the last call's flags), and the size check."""
import ctypes

import numpy as np
import pytest

from beamlab_amd import abi, fem

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself links (libamdhip64.so.7, already loaded with it): device buffers for the
    test's inputs and outputs."""
    abi.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(G, b):
    hip = _hip()
    lib = abi.load()
    b = np.ascontiguousarray(b, np.float64)
    x = np.full_like(b, np.nan)
    bd, xd = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(bd), b.nbytes) == 0 and hip.hipMalloc(ctypes.byref(xd), b.nbytes) == 0
    assert hip.hipMemcpy(bd, b.ctypes.data, b.nbytes, 1) == 0  # host to device
    assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, 1) == 0
    need = ctypes.c_uint64(0)
    rc = lib.bl_fem_solve_noalloc(G._h, bd, xd, ctypes.byref(need))
    abi.check(rc, "bl_fem_solve_noalloc")
    assert hip.hipMemcpy(x.ctypes.data, xd, x.nbytes, 2) == 0  # device to host
    hip.hipFree(bd), hip.hipFree(xd)
    return x, int(need.value)


def test_noalloc_solve_matches_solve():
    rng = np.random.default_rng(9)
    bodies = []
    for n in (24, 12):
        print(f"no-alloc solve {n}x{n}", flush=True)
        G = fem.ShellBody.plate(n, n)
        G.step(rng.normal(0, 20, (G.nn, 3)).astype(np.float32), 0.0005)
        bodies.append(G)
        b = rng.normal(0, 1, 6 * int((G.fixed == 0).sum()))
        x_ref = G.solve(b)
        x, need = _noalloc(G, b)
        assert need > 0
        assert x.tobytes() == x_ref.tobytes(), f"{n}x{n}: the no-allocation solve differs from bl_fem_solve"
    # the larger system again, after the smaller one left its flags in the scratch
    G = bodies[0]
    b = rng.normal(0, 1, 6 * int((G.fixed == 0).sum()))
    x_ref = G.solve(b)
    x, _ = _noalloc(G, b)
    assert x.tobytes() == x_ref.tobytes()
    for G in bodies:
        G.close()


def test_noalloc_solve_rejects_oversized_system():
    """The 256 x 256 plate fits the 32 MiB scratch; a 320 x 320 plate does not and is refused with BL_ERR_ARG (the
    size is checked before anything runs)."""
    lib = abi.load()
    need = ctypes.c_uint64(0)
    G = fem.ShellBody.plate(256, 256)
    assert lib.bl_fem_solve_noalloc(G._h, None, None, ctypes.byref(need)) == abi.BL_OK
    print(f"256x256 plate: {need.value} bytes of scratch", flush=True)
    assert 0 < need.value <= (32 << 20)
    G.close()
    G = fem.ShellBody.plate(320, 320)
    assert lib.bl_fem_solve_noalloc(G._h, None, None, ctypes.byref(need)) == abi.BL_OK
    print(f"320x320 plate: {need.value} bytes of scratch", flush=True)
    assert need.value > (32 << 20)
    hip = _hip()
    n = 8 * 6 * int((G.fixed == 0).sum())
    bd, xd = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(bd), n) == 0 and hip.hipMalloc(ctypes.byref(xd), n) == 0
    rc = lib.bl_fem_solve_noalloc(G._h, bd, xd, None)
    hip.hipFree(bd), hip.hipFree(xd)
    assert rc == abi.BL_ERR_ARG
    G.close()
