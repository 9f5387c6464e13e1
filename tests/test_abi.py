"""The C ABI library loads and exports every function include/pxg.h declares, and the Python
layouts match the header's struct sizes. No compute calls: these run without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from physxcoopa_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pxg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pxg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("pxg_create", "pxg_destroy", "pxg_set_bodies", "pxg_set_shapes", "pxg_step_fixed",
                     "pxg_step_phase", "pxg_step", "pxg_get_bodies", "pxg_get_manifolds", "pxg_get_pairs",
                     "pxg_state_hash", "pxg_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpxg.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    for n in declared_functions():
        assert n in _lib._SIGNATURES, n


def test_struct_sizes_match_the_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "pxg.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(pxg_body), sizeof(pxg_shape), sizeof(pxg_contact),
         sizeof(pxg_manifold), sizeof(pxg_config), sizeof(pxg_stats), sizeof(pxg_profile), sizeof(pxg_joint));
  return 0;
}
"""
    import tempfile
    d = tempfile.mkdtemp()
    c = os.path.join(d, "s.c")
    exe = os.path.join(d, "s")
    open(c, "w").write(src)
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
    sizes = [int(x) for x in subprocess.check_output([exe]).split()]
    assert sizes == [_abi.BODY_DTYPE.itemsize, _abi.SHAPE_DTYPE.itemsize, _abi.CONTACT_DTYPE.itemsize,
                     _abi.MANIFOLD_DTYPE.itemsize, ctypes.sizeof(_abi.Config), ctypes.sizeof(_abi.Stats),
                     ctypes.sizeof(_abi.Profile), _abi.JOINT_DTYPE.itemsize]


def test_config_default_matches_reference_defaults():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpxg.so not built")
    c = _abi.Config()
    _lib.lib().pxg_config_default(ctypes.byref(c))
    p = _abi.default_config()
    for f, _ in _abi.Config._fields_:
        a, b = getattr(c, f), getattr(p, f)
        if f in ("gravity", "reserved"):
            assert list(a) == pytest.approx(list(b))
        else:
            assert a == pytest.approx(b), f
    # physxcoopa/util/config.h:22-145
    assert (c.velocity_iterations, c.relax_iterations, c.position_iterations) == (16, 2, 4)
    assert c.aabb_margin == pytest.approx(0.1) and c.max_substeps == 8


def test_no_gpu_create_fails_cleanly():
    """Without a GPU the library reports an error instead of crashing (no CPU fallback)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpxg.so not built")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = _lib.lib().pxg_create(None, ctypes.byref(h))
    assert rc != 0
    assert b"device" in _lib.lib().pxg_last_error()


def test_cpp_facade_header_compiles(tmp_path):
    """include/pxg_world.hpp (the PhysicsWorld facade) is self-contained C++17 over pxg.h."""
    src = tmp_path / "facade.cpp"
    src.write_text('#include "pxg_world.hpp"\nint main() { pxg::BodyId id; return id.is_valid() ? 1 : 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout.decode()


def test_device_error_bits_are_distinct():
    """The device's error word (Counters::err) packs one bit per condition: the capacity limits
    and the grid barrier's expired wait. A shared bit would make a capacity error abort every
    later barrier wait (the barriers poll the word for the timeout bit) and report both."""
    text = open(os.path.join(ROOT, "physxcoopa_amd", "csrc", "pxg_state.h")).read()
    bits = {name: int(val) for name, val in re.findall(r"\b(ERR_[A-Z_]+)\s*=\s*(\d+)u", text)}
    assert {"ERR_PAIRS", "ERR_MESH_VERTICES", "ERR_BARRIER_BIT"} <= set(bits)
    assert all(v and v & (v - 1) == 0 for v in bits.values())
    assert len(set(bits.values())) == len(bits)
