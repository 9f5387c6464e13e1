"""The reference's PhysicsWorld tests run through the C++ facade (include/pxg_world.hpp) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "physxcoopa_amd", "pxg_world_test")


def test_cpp_physics_world_suite():
    assert os.path.exists(BIN), "build first: python -m physxcoopa_amd.build"
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=300)
    out = r.stdout.decode()
    print(out)
    assert r.returncode == 0, out
    assert out.count("PASS ") >= 33 and "FAIL " not in out
