"""Hull (one-sided) triangles in the oracle against the reference's own expectations (tests/test_physics.cpp:
2754-2802): a ball pushed 15 cm behind a hull triangle comes back out in front, no faster than the push-out speed
(2 m/s); behind a two-sided triangle the same ball stays where it is. Bit-for-bit agreement with the compiled
reference on the same scenes is tests/test_oracle_vs_ref.py's."""
import numpy as np

from _contact_cases import case_hull_pushout


def _run(oracle_mod, hull):
    settings, boxes, bodies, _ = case_hull_pushout(hull)
    w = oracle_mod.OracleWorld(settings, boxes)
    for b in bodies:
        w.add_body(b)
    speed = 0.0
    for _ in range(60):
        w.step_substeps(33)
        nodes, _ = w.read_body(1)
        m = nodes["mass"].astype(np.float64)
        v = (nodes["v"].astype(np.float64) * m[:, None]).sum(0) / m.sum()
        speed = max(speed, float(np.linalg.norm(v)))
    nodes, _ = w.read_body(1)
    m = nodes["mass"].astype(np.float64)
    centre = float((nodes["p"][:, 1].astype(np.float64) * m).sum() / m.sum()) - 5.0
    return centre, speed


def test_ball_behind_hull_comes_out(oracle_mod):
    centre, speed = _run(oracle_mod, True)
    assert centre > 0.05, centre
    assert speed < 2.1, speed


def test_ball_behind_two_sided_triangle_stays(oracle_mod):
    centre, _ = _run(oracle_mod, False)
    assert centre < -0.1, centre
