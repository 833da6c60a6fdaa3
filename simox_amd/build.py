"""Build the in-tree HIP library (gfx950) and the Fortran host pieces.

    python -m simox_amd.build          # library only
    python -m simox_amd.build --all    # library + Fortran drop-in + host driver
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("SIMOX_FC", "/opt/rocm/bin/amdflang")
LIB = os.path.join(HERE, "libsimox_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "simox.hip")]
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_lib(out=LIB):
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *SOURCES])
    os.replace(tmp, out)
    return out


def build_fortran():
    """Fortran iso_c_binding module, the drop-in FORCE/PME/... shim and the host
    driver program (simox_amd/fortran)."""
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    srcs = ["simox_datafmt.f90", "simox_gpu_iface.f90", "simox_dropin.f90", "simox_md.f90"]
    objs = []
    for s in srcs:
        o = os.path.join(bdir, s.replace(".f90", ".o"))
        _run([FC, "-O2", "-c", os.path.join(fdir, s), "-o", o, "-module-dir", bdir], cwd=bdir)
        objs.append(o)
    exe = os.path.join(fdir, "simox_md")
    _run([FC, "-O2", "-o", exe, *objs, "-L" + HERE, "-lsimox_gpu",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE], cwd=bdir)
    return exe


if __name__ == "__main__":
    print(build_lib())
    if "--all" in sys.argv:
        print(build_fortran())
