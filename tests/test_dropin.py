"""The HIP backend as a drop-in behind the reference's unchanged engine (CPU; the reference is only in the build
container).

HipSolver.cpp and HipCodegen.cpp are compiled against the reference's own headers (ahead of this repository's), with
the runtime (fe_hip.cpp) and the MFMA net library, and linked with the reference's Law, Solver, IR and Net objects
(built from its sources by oracle/Makefile.ref). The reference's Solver.cpp gets the two-line hook INTEGRATION.md shows
(an include and a FIELD_BACKEND=hip branch in Solver::use_gpu), compiled from a patched copy in a temporary folder.
With no device here, use_gpu must reach HipSolver::create and come back false with the runtime's reason; the
generator must still produce the law's translation unit.
"""
import os
import shutil
import subprocess

import pytest

REFERENCE = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "online-physics-cuda_amd")
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not (os.path.isdir(os.path.join(REFERENCE, "src", "field")) and os.path.isdir(OBJ)),
                                reason="the reference's sources and objects are only in the build container")

# The hook a maintainer adds to the reference's src/field/Solver.cpp (INTEGRATION.md shows it as a diff).
INCLUDE_ANCHOR = '#include "field/cuda/CudaSolver.hpp"\n'
BRANCH_ANCHOR = '    std::string cuda_why;\n'
HOOK_INCLUDE = '#include "field/hip/HipSolver.hpp"\n'
HOOK_BRANCH = ('    if (want == "hip") {\n'
               '        gpu_ = HipSolver::create(law_, fields_.shape(), why);\n'
               '        return gpu_ != nullptr;\n'
               '    }\n')


def patched_solver(tmp):
    text = open(os.path.join(REFERENCE, "src", "field", "Solver.cpp")).read()
    assert text.count(INCLUDE_ANCHOR) == 1 and text.count(BRANCH_ANCHOR) == 1
    text = text.replace(INCLUDE_ANCHOR, INCLUDE_ANCHOR + HOOK_INCLUDE)
    text = text.replace(BRANCH_ANCHOR, HOOK_BRANCH + BRANCH_ANCHOR)
    path = os.path.join(tmp, "Solver_with_hip.cpp")
    open(path, "w").write(text)
    return path


def test_hip_backend_builds_behind_the_references_solver(tmp_path):
    tmp = str(tmp_path)
    inc = ["-I" + os.path.join(REFERENCE, "src"), "-I" + os.path.join(PKG, "csrc"), "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    cxx = ["g++", "-std=c++20", "-O1", "-fopenmp", "-w"] + inc
    sources = {
        "HipSolver.o": os.path.join(PKG, "csrc", "field", "hip", "HipSolver.cpp"),
        "HipCodegen.o": os.path.join(PKG, "csrc", "field", "hip", "HipCodegen.cpp"),
        "fe_hip.o": os.path.join(PKG, "csrc", "runtime", "fe_hip.cpp"),
        "Solver_with_hip.o": patched_solver(tmp),
        "main.o": os.path.join(ROOT, "tests", "dropin", "main.cpp"),
    }
    objs = []
    for obj, src in sources.items():
        out = os.path.join(tmp, obj)
        r = subprocess.run(cxx + ["-c", src, "-o", out], capture_output=True, text=True)
        assert r.returncode == 0, f"{src} against the reference's headers:\n{r.stderr[-4000:]}"
        objs.append(out)
    net = os.path.join(tmp, "net_mfma.o")
    r = subprocess.run([HIPCC, "-std=c++20", "-O2", "--offload-arch=gfx950", "-w", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"] + inc[:3] + ["-c", os.path.join(PKG, "csrc", "kernels", "net_mfma.hip"), "-o", net],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    objs.append(net)
    ref_objs = []
    for rel in ("field/Law.o", "field/Grid.o", "field/LawBounds.o", "field/Scene.o", "field/GpuSolver.o", "field/nn/Net.o",
                "field/ir/LawIR.o", "field/ir/Adjoint.o", "field/cuda/CudaApi.o", "field/cuda/CudaCodegen.o",
                "field/cuda/CudaNet.o", "field/cuda/CudaSolver.o"):
        ref_objs.append(os.path.join(OBJ, rel))
    exe = os.path.join(tmp, "dropin")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-fopenmp", "-o", exe] + objs + ref_objs +
                       ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-ldl", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, FIELD_HIP_CACHE=os.path.join(tmp, "cache"))
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    seen = dict(line.split("=", 1) for line in run.stdout.splitlines() if "=" in line)
    assert seen["gather"] == "1" and int(seen["source_bytes"]) > 1000
    if seen["unavailable"]:
        # No device: the reference's Solver reached HipSolver::create, which refused with the runtime's reason.
        assert seen["use_gpu"] == "0"
        assert seen["why"] == seen["unavailable"]
        assert seen["api"] == "CPU"
    else:
        assert seen["use_gpu"] == "1" and seen["api"] == "HIP" and seen["same_bits"] == "1"
