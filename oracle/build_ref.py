"""Build the oracle's native pieces (test / baseline infrastructure only).

  oracle/_cpu/libpb3d_cpu.so  - cpu_kernels.c (gcc -O3 -fopenmp): the reference's bin-integral loop
                                structure, timed by bench.py's cpu_baseline leg.
  oracle/_ref/                - the reference's own utils.pyx compiled with Cython + gcc from
                                /root/reference when that tree is present (this container only);
                                used by tests to cross-check the oracle's kernel restatement.
Outputs are git-ignored.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PYX = "/root/reference/PolyBin3D/cython/utils.pyx"


def build_cpu():
    out = os.path.join(HERE, "_cpu")
    os.makedirs(out, exist_ok=True)
    lib = os.path.join(out, "libpb3d_cpu.so")
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-o", lib, os.path.join(HERE, "cpu_kernels.c")]
    subprocess.run(cmd, check=True)
    return lib


def build_ref():
    """Cythonize + compile the reference utils.pyx into oracle/_ref/ (skipped when absent)."""
    if not os.path.exists(REF_PYX):
        return None
    try:
        import numpy as np
        from Cython.Build import cythonize  # noqa: F401
    except ImportError:
        return None
    out = os.path.join(HERE, "_ref")
    os.makedirs(out, exist_ok=True)
    c_file = os.path.join(out, "utils.c")
    subprocess.run([sys.executable, "-m", "cython", "-3", "-o", c_file, REF_PYX], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    so = os.path.join(out, "utils" + ext)
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-I", sysconfig.get_paths()["include"],
           "-I", np.get_include(), "-o", so, c_file, "-lm"]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT)
    return so


def build():
    build_cpu()
    try:
        build_ref()
    except subprocess.CalledProcessError as e:  # the reference build is optional
        sys.stderr.write("oracle/_ref build skipped: %s\n" % e)


if __name__ == "__main__":
    build()
