"""The drop-in translation units of assemble-them-all_amd/host/route_a compile in
place of the reference's simulation/redmax/Body/BodySDFObj.cpp and
Force/ForceGeneralSDFContact.cpp against the reference's own, unchanged headers
(its include set: simulation/CMakeLists.txt, externals/CMakeLists.txt:10-20), and
each defines every member function its header declares (Body/BodySDFObj.h:18-90,
Force/ForceGeneralSDFContact.h:22-53), so the reference builds and links with
them. Nothing of the reference is copied: the test reads its tree where it lies
and is skipped where that tree is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/simulation"
ROUTE_A = os.path.join(ROOT, "assemble-them-all_amd", "host", "route_a")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "redmax")) or shutil.which("g++") is None,
                                reason="the reference tree (or g++) is not present")

INCLUDES = ["redmax", "externals/eigen", "externals/glew/include", "externals/glfw/include/GLFW", "externals/imgui",
            "externals/libigl/include", "externals/stb", "externals/pugixml/src", "externals/tiny_obj_loader",
            "externals/sdfgen", "opengl_viewer/include"]

# member functions each header declares (demangled prefixes)
MEMBERS = {
    "BodySDFObj.cpp": ["BodySDFObj::BodySDFObj(", "BodySDFObj::distance(", "BodySDFObj::surface_contact_pt(",
                       "BodySDFObj::collision(", "BodySDFObj::test_collision_derivatives()",
                       "BodySDFObj::test_collision_derivatives_runtime(", "BodySDFObj::clear_saved_SDF(",
                       "BodySDFObj::precompute_SDF(", "BodySDFObj::load_SDF(", "BodySDFObj::save_SDF(",
                       "BodySDFObj::query_SDF(", "BodySDFObj::query_dSDF(", "BodySDFObj::query_ddSDF(",
                       "BodySDFObj::query_grid_location("],
    "ForceGeneralSDFContact.cpp": [
        "ForceGeneralSDFContact::ForceGeneralSDFContact(", "ForceGeneralSDFContact::set_stiffness(",
        "ForceGeneralSDFContact::set_friction(", "ForceGeneralSDFContact::set_damping(",
        "ForceGeneralSDFContact::set_scale(", "ForceGeneralSDFContact::computeForce(",
        "ForceGeneralSDFContact::computeForceWithDerivative(", "ForceGeneralSDFContact::test_derivatives()",
        "ForceGeneralSDFContact::test_derivatives_runtime()",
        "ForceGeneralSDFContact::test_design_derivatives_runtime()"],
}
# overload counts: collision x2; computeForce x2 (globals, contact list);
# computeForceWithDerivative x4 (two global, two contact-list)
OVERLOADS = {"BodySDFObj::collision(": 2, "ForceGeneralSDFContact::computeForce(": 2,
             "ForceGeneralSDFContact::computeForceWithDerivative(": 4}


def _cmd(tu, *extra):
    return (["g++", "-std=c++14", "-DGLEW_NO_GLU", "-DGLEW_STATIC", '-DPROJECT_DIR="."']
            + ["-I" + os.path.join(REF, d) for d in INCLUDES]
            + ["-I" + os.path.join(ROOT, "include"), "-I" + ROUTE_A, os.path.join(ROUTE_A, tu)] + list(extra))


@pytest.mark.parametrize("tu", sorted(MEMBERS))
def test_route_a_defines_the_reference_interface(tu, tmp_path):
    obj = str(tmp_path / (tu + ".o"))
    r = subprocess.run(_cmd(tu, "-c", "-O0", "-o", obj), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    nm = subprocess.run(["nm", "-C", "--defined-only", obj], capture_output=True, text=True, check=True).stdout
    defined = [ln.split(" ", 2)[2] for ln in nm.splitlines() if " T " in ln or " W " in ln]
    for m in MEMBERS[tu]:
        hits = {d for d in defined if ("redmax::" + m) in d}
        assert hits, f"{tu} does not define {m}"
        if m in OVERLOADS:
            assert len(hits) >= OVERLOADS[m], (m, sorted(hits))
