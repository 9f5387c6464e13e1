"""Host-side helpers of bench.py checked on the CPU: the state aggregates the FP64 comparison
reports, against closed forms."""
import math

import numpy as np

import bench
from physxcoopa_amd import _abi


def one_body(v, w, q, iil, mass=2.0, awake=1):
    b = np.zeros(1, _abi.BODY_DTYPE)
    b["type"], b["alive"], b["awake"], b["mass"] = _abi.PXG_DYNAMIC, 1, awake, mass
    b["linear_velocity"], b["angular_velocity"], b["orientation"], b["inv_inertia_local"] = v, w, q, iil
    return b


def test_kinetic_energy_linear_and_rotational_in_the_body_frame():
    # 2 kg at 1 m/s: 1 J; spinning at 1 rad/s about world x while yawed 90 degrees about z, so the
    # spin is about the body's -y axis, whose moment is 1 / 2: 0.25 J (0.5 J if the frame were ignored)
    s = math.sin(math.pi / 4)
    b = one_body((1, 0, 0), (1, 0, 0), (0, 0, s, s), (1.0, 2.0, 4.0))
    agg = bench.state_aggregates(b, np.zeros(0, _abi.MANIFOLD_DTYPE))
    assert abs(agg["kinetic_energy"] - 1.25) < 1e-6
    assert agg["asleep"] == 0 and agg["max_penetration"] == 0.0


def test_penetration_and_sleep_counts():
    b = np.concatenate([one_body((0, 0, 0), (0, 0, 0), (0, 0, 0, 1), (1, 1, 1), awake=0),
                        one_body((0, 0, 0), (0, 0, 0), (0, 0, 0, 1), (1, 1, 1))])
    m = np.zeros(3, _abi.MANIFOLD_DTYPE)
    m["valid"], m["count"] = 1, 2
    m["points"]["penetration"][0, :2] = (0.01, 0.03)
    m["points"]["penetration"][1, :3] = (0.02, 0.0, 0.5)  # the third point is past the count
    m["points"]["penetration"][2, :2] = (0.9, 0.9)
    m["is_trigger"][2] = 1  # trigger manifolds do not push
    agg = bench.state_aggregates(b, m)
    assert agg["asleep"] == 1
    assert abs(agg["max_penetration"] - 0.03) < 1e-7
