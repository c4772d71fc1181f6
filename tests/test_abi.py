"""The C ABI library loads and exports every symbol include/biocfd_hip.h
declares; host-only entry points (body kinematics) run without a GPU and
match the reference fixtures bit for bit."""
import ctypes as C
import math
import os
import re

import numpy as np

from biocfd_amd import hip as H
from parity import ROOT

import sys
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden  # noqa: E402

HEADER = os.path.join(ROOT, "include", "biocfd_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(biocfd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(H.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    syms = set(declared_symbols()) - {"biocfd_last_error"}
    assert syms == set(H.SIGNATURES), syms ^ set(H.SIGNATURES)


def test_no_oracle_in_product_library():
    """The product library links nothing from oracle/."""
    data = open(H.LIB_PATH, "rb").read()
    assert b"orc_" not in data


def test_host_kinematics_match_reference():
    """shiftSurfaceNodesInitial + computeSurfaceVariables (Search.F90:49-208)
    through the C ABI reproduce the reference's moved nodes and body
    velocities (nested case: heaving sphere; wing case: flapping about x)."""
    lib = H.load()
    for name in ("nested", "wing"):
        case = make_golden.golden_cases()[name]
        fx = np.load(make_golden.fixture_path(name))
        g = len(case.blocks)
        spec = case.blocks[g - 1]
        b = spec.body.body
        xn, yn, zn = (np.ascontiguousarray(b.nodes[:, q]) for q in range(3))
        nid = np.ascontiguousarray(b.node_id)
        st = np.zeros(H.BKS_N)
        st[:6] = [spec.body.centre[0], spec.body.centre[1], spec.body.centre[2], spec.body.a0_deg,
                  spec.body.yamp, spec.body.bfreq]
        x1, y1, z1 = np.zeros(len(xn)), np.zeros(len(xn)), np.zeros(len(xn))
        p = H.ptr
        rc = lib.biocfd_shift_surface_nodes_initial(len(xn), p(xn), p(yn), p(zn), p(nid), case.aoa, case.piv_pt,
                                                    0, case.dt, case.dxmin, 0.0, p(st), p(x1), p(y1), p(z1))
        assert rc == 0
        totime = case.dt
        phase = case.phase_angle_deg * math.pi / 180.0
        for ita in range(1, max(make_golden.STEPS) + 1):
            totime = totime + case.dt
            rc = lib.biocfd_compute_surface_variables(len(xn), p(xn), p(yn), p(zn), p(nid), phase, case.piv_pt,
                                                      ita, case.dt, case.dxmin, case.freq, totime, p(st), p(x1),
                                                      p(y1), p(z1))
            assert rc == 0
        s = max(make_golden.STEPS)
        pre = f"s{s}_b{g}_"
        assert np.array_equal(y1, fx[pre + "ynode1"]), name
        assert st[12] == fx[pre + "ydot"]            # ydot
        assert st[18] == fx[pre + "thetaDot1"]       # thetaDot1
        assert st[24] == fx[pre + "piv_y"]           # piv_y
