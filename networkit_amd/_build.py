"""Build recipe of the native parts of networkit_amd (run by __graft_entry__.build()).

- libnkpr.so      hipcc, gfx950 only: HIP kernels + the C ABI of include/nkpr.h
- _core*.so       Cython + g++: Graph, the PageRank drop-in (C++) and their Python classes,
                  linked against libnkpr.so
The objects go to networkit_amd/_build/ and the libraries next to this file (in-tree, so the
snapshot that reaches a GPU machine carries them).
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_INCLUDE = "/opt/rocm/include"
ROCM_LIB = "/opt/rocm/lib"

HIP_SOURCES = ["nkpr_sweep.hip", "nkpr_graph.hip", "nkpr_krylov.hip", "nkpr_batch.hip",
               "nkpr_adjoint.hip", "nkpr_groups.hip", "nkpr_binned.hip", "nkpr_power.hip",
               "nkpr_csr.hip", "nkpr_algebraic.hip", "nkpr_nkbg.hip", "nkpr_apr.hip", "nkpr_capi.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay off (some of their forms are unreliable on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I" + INCLUDE, "-I" + CSRC,
]
LIB_NKPR = os.path.join(HERE, "libnkpr.so")
LIB_NKGEN = os.path.join(HERE, "libnkgen.so")
CXX_SOURCES = ["Graph.cpp", "PageRank.cpp", "EigenvectorCentrality.cpp", "AlgebraicPageRank.cpp",
               "NetworkitBinaryIO.cpp", "SelectiveCommunity.cpp"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def build_hip(jobs=4):
    os.makedirs(BUILD, exist_ok=True)

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES + ["nkgen.hip"]))
    gen_obj = objs.pop()
    tmp = LIB_NKPR + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB_NKPR)
    tmp = LIB_NKGEN + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, gen_obj])
    os.replace(tmp, LIB_NKGEN)
    return LIB_NKPR


def build_cython():
    """Cythonize _core.pyx and compile it with the C++ host sources against libnkpr.so."""
    from Cython.Build import cythonize  # noqa: F401  (import check)
    from Cython.Compiler import Main as cymain

    os.makedirs(BUILD, exist_ok=True)
    pyx = os.path.join(HERE, "_core.pyx")
    cpp = os.path.join(BUILD, "_core.cpp")
    opts = cymain.CompilationOptions(cymain.default_options, cplus=True, language_level=3,
                                     output_file=cpp)
    res = cymain.compile(pyx, opts)
    if res.num_errors:
        raise RuntimeError("cython failed on _core.pyx")
    py_inc = sysconfig.get_paths()["include"]
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_core" + ext)
    import numpy
    cxx = os.environ.get("CXX", "g++")
    flags = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
             "-I" + py_inc, "-I" + INCLUDE, "-I" + CSRC, "-I" + numpy.get_include(),
             # the HIP runtime API (host code only: device copies of CSR matrices)
             "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INCLUDE]
    objs = []
    for src in [cpp] + [os.path.join(CSRC, s) for s in CXX_SOURCES]:
        obj = os.path.join(BUILD, os.path.basename(src).replace(".cpp", ".host.o"))
        _run([cxx] + flags + ["-c", src, "-o", obj])
        objs.append(obj)
    tmp = out + ".tmp"
    _run([cxx, "-shared", "-fopenmp", "-o", tmp] + objs +
         ["-L" + HERE, "-lnkpr", "-Wl,-rpath,$ORIGIN", "-L" + ROCM_LIB, "-lamdhip64",
          "-Wl,-rpath," + ROCM_LIB])
    os.replace(tmp, out)
    return out


def build_all():
    build_hip()
    build_cython()


if __name__ == "__main__":
    build_all()
    print("built", LIB_NKPR, file=sys.stderr)
