"""The Fortran side of the boundary: the iso_c_binding module declares every
entry point of include/biocfd_hip.h with the same name and argument count,
and the Fortran host driver is built (no GPU calls here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "biocfd_hip.h")
BINDING = os.path.join(ROOT, "bio-cfd_amd", "fortran", "biocfd_hip.f90")
RUN = os.path.join(ROOT, "bio-cfd_amd", "lib", "biocfd_run")


def header_protos():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(?:int|const char \*)\s*(biocfd_\w+)\s*\(([^)]*)\);", src):
        args = args.strip()
        out[name] = 0 if args in ("", "void") else len(args.split(","))
    return out


def binding_protos():
    src = open(BINDING).read().replace("&\n", "")
    out = {}
    for dummies, name in re.findall(r"function\s+\w+\(([^)]*)\)\s*bind\(C,\s*name=\"(biocfd_\w+)\"\)", src):
        dummies = dummies.strip()
        out[name] = 0 if not dummies else len(dummies.split(","))
    return out


def test_binding_covers_the_header():
    h = header_protos()
    b = binding_protos()
    assert len(h) > 50
    assert set(h) == set(b), (sorted(set(h) - set(b)), sorted(set(b) - set(h)))
    bad = {k: (h[k], b[k]) for k in h if h[k] != b[k]}
    assert not bad, bad


@pytest.mark.skipif(not os.path.exists(RUN), reason="Fortran driver not built")
def test_fortran_driver_usage():
    r = subprocess.run([RUN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage: biocfd_run" in r.stdout
