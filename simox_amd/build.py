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
    """The iso_c_binding module and the Fortran host program simox_md
    (simox_amd/fortran/_build, executable simox_amd/fortran/simox_md)."""
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in ("simox_gpu.f90", "simox_oxid.f90", "simox_md.f90"):
        o = os.path.join(bdir, src.replace(".f90", ".o"))
        # simox_oxid: no FMA contraction, as the reference build compiles its own
        fl = ["-ffp-contract=off"] if src == "simox_oxid.f90" else []
        _run([FC, "-O2", *fl, "-c", os.path.join(fdir, src), "-o", o, "-module-dir", bdir],
             cwd=bdir)
        objs.append(o)
    exe = os.path.join(fdir, "simox_md")
    _run([FC, "-O2", "-o", exe, *objs, "-L" + HERE, "-lsimox_gpu", "-Wl,-rpath," + HERE,
          "-Wl,-rpath,$ORIGIN/.."], cwd=bdir)
    return exe


def build_dropin_demo(ref="/root/reference"):
    """The reference's own program (main.f90 and its other units) with
    force.f90 / ctip.f90 / cellsort.f90 replaced by simox_dropin.f90: SIMOX
    running its hot path on the GPU, FFTE no longer needed.  Built only where the reference
    sources are present; the executable is simox_amd/fortran/_build/simox_dropin_demo."""
    if not os.path.isdir(ref):
        return None
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build", "dropin")
    os.makedirs(bdir, exist_ok=True)
    objs = {}
    for src in [os.path.join(ref, f + ".f90") for f in ("datafmt", "eam", "vverlet", "util")] + \
            [os.path.join(fdir, "simox_gpu.f90"), os.path.join(fdir, "simox_dropin.f90"),
             os.path.join(ref, "main.f90")]:
        o = os.path.join(bdir, os.path.basename(src).replace(".f90", ".o"))
        _run([FC, "-O2", "-c", src, "-o", o, "-module-dir", bdir], cwd=bdir)
        objs[os.path.basename(src)[:-4]] = o
    # simox_dropin.o ahead of vverlet.o: its VV_NVE1 / VV_NVE2 are the
    # definitions the link keeps (first definition wins), the rest of
    # vverlet.f90 (VINIT, thermostats, the reflecting half) stays the reference's
    order = ["datafmt", "eam", "simox_gpu", "simox_dropin", "vverlet", "util", "main"]
    exe = os.path.join(fdir, "_build", "simox_dropin_demo")
    _run([FC, "-O2", "-o", exe, *[objs[k] for k in order], "-Wl,--allow-multiple-definition",
          "-L" + HERE, "-lsimox_gpu", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../.."], cwd=bdir)
    return exe


if __name__ == "__main__":
    print(build_lib())
    if "--all" in sys.argv:
        print(build_fortran())
        print(build_dropin_demo())
