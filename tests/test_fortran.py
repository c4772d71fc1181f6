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


# the module procedures the reference driver imports from the hot-path
# modules (app/main.F90:4-27), which the drop-in must define
MAIN_USES = {
    "biocfd_search": ["findDistnode", "shiftSurfaceNodesInitial", "computeSurfaceNorm", "tagging_th",
                      "tagging_th_move", "block_move_check", "cellcount_solid", "cellcount_solid_coarse",
                      "cellcount_solid_coarse_mv", "change_block_coords", "change_block_interface",
                      "computenormdistance", "computesurfacevariables", "findtscells", "fine_block_cell",
                      "selectiveretagging_th", "change_block_coords_interfaces"],
    "biocfd_pcor_vcor": ["poissonSolver", "updateVelocity_newv"],
    "biocfd_boundary_conditions": ["velocityBC", "solidcellbc", "solidcellbc_move"],
    "biocfd_coefficient_matrix": ["coefficientMatrix"],
    "biocfd_navier_stokes": ["non_uni_coeff", "nsmomentum2order"],
    "biocfd_write_output_corner1": ["body_plot", "writeresult", "write_output_ascii"],
    "biocfd_forcing": ["pressureForcing1", "pressureforcingfield", "pressureforcingghost", "velocityforcing1",
                       "velocityforcingfield", "velocityforcingghost"],
    "biocfd_gpu_helpers": ["set_gpu"],
    # called inside the replaced modules in the reference
    "biocfd_coarse_update": ["coarseUpdate_newv", "coarseUpdate_pc", "coarseUpdate"],
    "biocfd_fine_interp_bound": ["fineUpdate_pc_bd", "fineUpdate_newv_bd", "fineUpdate_bd_mv"],
    "biocfd_fine_interp": ["fineUpdate_mv"],
    "biocfd_stress_calculation": ["stress_calculation"],
}


def test_dropin_modules_define_what_main_uses():
    """Every hot-path module and procedure the reference driver uses has a
    replacement with that name, public and defined as a subroutine."""
    d = os.path.join(ROOT, "bio-cfd_amd", "fortran", "dropin")
    src = "\n".join(open(os.path.join(d, f)).read() for f in sorted(os.listdir(d)) if f.endswith(".f90"))
    src_l = src.lower()
    modules = dict(re.findall(r"(?ms)^module (\w+)\n(.*?)^end module \1", src_l))
    missing = []
    for mod, procs in MAIN_USES.items():
        body = modules.get(mod)
        if body is None:
            missing.append(f"module {mod}")
            continue
        public = set(re.findall(r"\w+", " ".join(re.findall(r"public ::([^\n]*)", body))))
        for p in procs:
            if p.lower() not in public or not re.search(rf"subroutine {p.lower()}\(", body):
                missing.append(f"{mod}::{p}")
    assert not missing, missing
