"""Build recipe for the native libraries (run by __graft_entry__.build()).

libdft_hip.so  -- gfx950 kernels + C-ABI (dft_amd/csrc/*.hip, include/dft_hip.h)
libdft_core.so -- C++ host layer keeping the reference's solver interfaces
                  (dft_amd/host/*.cpp, include/dft_molecule.h, include/dft_core.h)

Both land in dft_amd/lib/ (in-tree, so they travel with the repository
snapshot).  Objects are compiled in parallel and linked with -Wl,-rpath so
libdft_core.so finds libdft_hip.so next to it.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
    # packed FP32 instructions stay off (see the kernels' FP32 paths)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
             "-I", os.path.join(HERE, "host"), "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    return r


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    cmds = []
    objs_hip, objs_host = [], []
    for s in hip_srcs:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs_hip.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    for s in host_srcs:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs_host.append(o)
        cmds.append(["g++", *CXX_FLAGS, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    hip_so = os.path.join(LIB, "libdft_hip.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", hip_so, *objs_hip,
          "-L", os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    out = [hip_so]
    if objs_host:
        core_so = os.path.join(LIB, "libdft_core.so")
        _run(["g++", "-shared", "-fPIC", "-o", core_so, *objs_host, "-L", LIB, "-ldft_hip",
              "-Wl,-rpath,$ORIGIN", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
              "-Wl,-rpath," + os.path.join(ROCM, "lib")])
        out.append(core_so)
        # test programs written against the reference-shaped C++ API
        tdir = os.path.join(ROOT, "tests", "cpp")
        for src in sorted(glob.glob(os.path.join(tdir, "*.cpp"))):
            os.makedirs(os.path.join(tdir, "bin"), exist_ok=True)
            exe = os.path.join(tdir, "bin", os.path.splitext(os.path.basename(src))[0])
            _run(["g++", "-std=c++17", "-O2", src, "-I", os.path.join(ROOT, "include"), "-L", LIB, "-ldft_core",
                  "-ldft_hip", "-Wl,-rpath,$ORIGIN/../../../dft_amd/lib", "-o", exe])
            out.append(exe)
    if verbose:
        print("built", *out)
    return out


if __name__ == "__main__":
    build(verbose=True)
