"""The implicit shell step on the GPU (bl_fem_*) against the oracle (oracle/fem_oracle.c): the element stiffness bit
for bit, the assembled system and its right side within FP64 rounding, the multifrontal Cholesky's solution by its
residual (<= 1e-8, the reference factorization's bar) and against the oracle's block Cholesky, trajectories within
1e-5 of the displacement, and the clamped plate against Kirchhoff plate theory."""
import numpy as np
import pytest

from beamlab_amd import fem

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, nu, nv, section=None):
    sec = section or oracle_mod.SECTION_C3
    P = oracle_mod.FemPlate(nu, nv, section=sec)
    G = fem.ShellBody.plate(nu, nv, section=sec)
    return P, G


def test_fem_mesh_and_element_bitwise(oracle_mod):
    print("fem element 6x5", flush=True)
    P, G = _pair(oracle_mod, 6, 5)
    tn, xd, fixed, mass, inertia = P.topology()
    assert np.array_equal(tn, G.tris) and np.array_equal(fixed, G.fixed)
    assert np.array_equal(xd, G.p.astype(np.float64))
    for k in range(P.nt):
        p = xd[tn[k]].astype(np.float32)
        K = oracle_mod.fem_tri_local_k(oracle_mod.fem_rest_shape(p))
        assert G.tri_k(k).tobytes() == K.tobytes(), f"element {k}"


@pytest.mark.parametrize("shape", [(6, 5), (9, 7), (16, 16)])
def test_fem_system_and_solution(oracle_mod, shape):
    print(f"fem system {shape}", flush=True)
    P, G = _pair(oracle_mod, *shape)
    rng = np.random.default_rng(sum(shape))
    for it in range(3):
        loads = rng.normal(0, 40, (P.nn, 3)).astype(np.float32)
        Ao, ro, xo, fm = P.step(loads, 0.002, system=True)
        rep = G.step(loads, 0.002)
        Ag, rg, xg = G.last_system()
        # the element frames and corner moments are float quantities (atan2f, sinf of the GPU against the C
        # library's): the system and the right side agree to float rounding, not bit for bit
        scale = np.abs(Ao).max()
        assert np.abs(Ag - Ao).max() <= 1e-6 * scale, f"step {it}: system"
        assert np.linalg.norm(rg - ro) <= 1e-6 * np.linalg.norm(ro), f"step {it}: right side"
        assert rep["residual"] <= 1e-8 and rep["failures"] == 0, rep
        # the GPU's solution solves its system (checked on the host too) and agrees with the oracle's block Cholesky
        assert np.linalg.norm(Ag @ xg - rg) <= 1e-8 * np.linalg.norm(rg)
        assert np.linalg.norm(xg - xo) <= 1e-5 * np.linalg.norm(xo)
        xs, vs, ws, qs = P.state()
        xg_, vg, wg, qg = G.state()
        assert np.abs(xg_ - xs).max() <= 1e-8, f"step {it}: positions"
        # carry on from the oracle's state so the next step's comparison starts equal
        G.set_state(xs, vs, ws, qs)


def test_fem_solve_random_right_sides(oracle_mod):
    P, G = _pair(oracle_mod, 12, 10)
    G.step(None, 0.002)
    A, _, _ = G.last_system()
    rng = np.random.default_rng(5)
    for _ in range(4):
        b = rng.normal(0, 1, A.shape[0])
        x = G.solve(b)
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("amp", [3.0, 30.0])
def test_fem_trajectory(oracle_mod, amp):
    """40 steps of a loaded plate under gravity against the oracle. The GPU's float transcendentals (atan2f, sinf,
    cosf) round differently from the C library's, and the step amplifies such differences; the bar is 1e-5 of the
    largest displacement, or 4x the oracle's own divergence from a copy of itself nudged by one ulp, if larger."""
    print(f"fem trajectory 12x12 amp {amp}", flush=True)
    P, G = _pair(oracle_mod, 12, 12)
    Q = oracle_mod.FemPlate(12, 12)
    xd, v, w, q = Q.state()
    q[:, 3] = np.nextafter(q[:, 3], np.float32(2))
    Q.set_state(xd, v, w, q)
    x0 = P.state()[0].copy()
    rng = np.random.default_rng(11)
    loads = rng.normal(0, amp, (P.nn, 3)).astype(np.float32)
    disp = gpu_div = self_div = 0.0
    for s in range(40):
        P.step(loads, 0.002)
        Q.step(loads, 0.002)
        rep = G.step(loads, 0.002)
        assert rep["residual"] <= 1e-8
        xs, xq, xg = P.state()[0], Q.state()[0], G.state()[0]
        disp = max(disp, np.abs(xs - x0).max())
        gpu_div = max(gpu_div, np.abs(xg - xs).max())
        self_div = max(self_div, np.abs(xq - xs).max())
    print(f"max displacement {disp:.3e}, gpu vs oracle {gpu_div:.3e}, oracle vs nudged oracle {self_div:.3e}")
    assert disp > 1e-4
    assert gpu_div <= max(1e-5 * disp, 4 * self_div)


def test_fem_kirchhoff_plate():
    """Clamped square plate, central load: w = 0.00560 P a^2 / D (Timoshenko & Woinowsky-Krieger, table 55)."""
    n = 24
    sec = list(fem.SECTION_PLATE)
    sec[5] = 0.0
    G = fem.ShellBody.plate(n, n, section=sec)
    loads = np.zeros((G.nn, 3), np.float32)
    c = (n // 2) * (n + 1) + n // 2
    loads[c, 1] = -1.0
    x0 = G.state()[0].copy()
    rep = G.step(loads, 1e3, theta=1.0, dissipation=0.0, g=(0, 0, 0))
    assert rep["residual"] <= 1e-8
    w = (G.state()[0] - x0)[c, 1]
    E, nu, t = 70e9, 0.33, 1.5e-3
    D = E * t ** 3 / (12 * (1 - nu * nu))
    assert w == pytest.approx(-0.0056 / D, rel=0.01)


def test_fem_large_plate_residual():
    """The 256 x 256 plate (65025 free nodes, 390150 unknowns): the factorization's residual stays <= 1e-8."""
    print("fem plate 256", flush=True)
    G = fem.ShellBody.plate(256, 256)
    rng = np.random.default_rng(3)
    for _ in range(3):
        rep = G.step(rng.normal(0, 5, (G.nn, 3)).astype(np.float32), 0.002)
        assert rep["residual"] <= 1e-8 and rep["failures"] == 0, rep
    print(rep, flush=True)
