"""The facade's cloth and hair builders (include/pxg_soft.hpp: make_grid_cloth, pin_static,
pin_to_body, build_tethers, make_strands; cloth/cloth_builder.h, hair/strand_builder.h) produce
the oracle builders' records bit for bit: a turned, shear, pinned and tethered sheet, a default
sheet without shear, and a groom with turned roots and a one-point descriptor that is skipped."""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from physxcoopa_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONSTRAINT = _abi.CONSTRAINT_DTYPE


@pytest.fixture(scope="module")
def dump():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "dump")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "soft_builders_dump.cpp"), "-o", exe], check=True)
        out = os.path.join(tmp, "out.bin")
        subprocess.run([exe, out], check=True)
        with open(out, "rb") as f:
            yield f.read()


class Reader:
    def __init__(self, data):
        self.data, self.pos = data, 0

    def array(self, dtype):
        n = int(np.frombuffer(self.data, np.uint32, 1, self.pos)[0])
        self.pos += 4
        a = np.frombuffer(self.data, dtype, n, self.pos)
        self.pos += n * np.dtype(dtype).itemsize
        return a

    def raw(self, nbytes):
        b = self.data[self.pos:self.pos + nbytes]
        self.pos += nbytes
        return b


def same(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape, what
    assert a.tobytes() == b.tobytes(), what


def check_cloth(r, c):
    same(r.array(_abi.PARTICLE_DTYPE), c.particles, "particles")
    same(r.array(CONSTRAINT), c.stretch, "stretch")
    same(r.array(CONSTRAINT), c.bend, "bend")
    same(r.array(np.uint32), c.stretch_batches.reshape(-1), "stretch batches")
    same(r.array(np.uint32), c.bend_batches.reshape(-1), "bend batches")
    same(r.array(_abi.TETHER_DTYPE), c.tethers, "tethers")
    same(r.array(_abi.ANCHOR_DTYPE), c.anchors, "anchors")
    same(r.array(np.uint32), c.triangles, "triangles")
    assert r.raw(len(bytes(c.params))) == bytes(c.params), "params (self_distance)"


def test_builders_match_the_oracle(oracle, dump):
    r = Reader(dump)
    c = oracle.grid_cloth(9, 7, 2.0, 1.5, (0.3, -0.2, 2.0), (0.0, 0.38268343, 0.0, 0.92387953), total_mass=1.7)
    p0 = tuple(float(x) for x in c.particles["position"][0])
    p1 = tuple(float(x) for x in c.particles["position"][-1])
    pose = ((1.0, 0.5, 2.5), (0.1, 0.2, -0.3, 0.9273618))
    c = oracle.grid_cloth(9, 7, 2.0, 1.5, (0.3, -0.2, 2.0), (0.0, 0.38268343, 0.0, 0.92387953), total_mass=1.7,
                          pins=[(p0, 0.3), (p1, 0.4, (3, 7), pose)], tethers=True)
    assert len(c.anchors) >= 2 and len(c.tethers) > 0
    check_cloth(r, c)
    check_cloth(r, oracle.grid_cloth(shear=False))
    strands = [[(np.float32(0.1) * k, np.float32(0.02) * k * k, 0.5) for k in range(5)], [(1.0, 1.0, 1.0)],
               [(-0.3, np.float32(0.05) * k, np.float32(1.0) - np.float32(0.03) * k) for k in range(9)]]
    s = oracle.make_strands(strands, _abi.StrandParams(thickness=0.02), masses=[0.001, 0.001, 0.002],
                            root_orientations=[(0, 0, 0.38268343, 0.92387953), (0, 0, 0, 1), (0, 0, 0, 1)])
    same(r.array(_abi.PARTICLE_DTYPE), s.particles, "strand particles")
    same(r.array(CONSTRAINT), s.stretch, "strand stretch")
    same(r.array(np.uint32), s.stretch_batches.reshape(-1), "strand batches")
    same(r.array(np.float32), s.rest_local.reshape(-1), "rest local")
    same(r.array(np.float32), s.rest_offset.reshape(-1), "rest offset")
    same(r.array(_abi.STRAND_DTYPE), s.strands, "strands")
    assert r.pos == len(dump)
