"""Build libpb3d.so (the HIP kernels + C ABI) in-tree with hipcc for gfx950.

Usage: python -m polybin3d_amd.build   (or polybin3d_amd.build.build())
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpb3d.so")
SOURCES = ["pb3d_api.hip", "pb3d_fft.hip", "pb3d_fftgen.hip", "pb3d_maps.hip", "pb3d_bins.hip", "pb3d_triplet.hip", "pb3d_pairgemm.hip", "pb3d_fisher.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 (v_pk_*_f32) disabled: hand-checked scalar FP32 only
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose=True, jobs=None, out=None, defines=()):
    """Compile every source to an object in parallel, then link libpb3d.so (or `out`, with extra
    -D `defines`, for variant builds)."""
    lib = out or LIB
    objdir = os.path.join(HERE, "_build") if out is None else os.path.join(HERE, "_build", "variant_" + os.path.basename(out).replace(".", "_"))
    os.makedirs(objdir, exist_ok=True)
    t0 = time.time()
    # hipcc reads a source once per offload pass (device, then host): a source edited during the
    # build would give a library whose host launch stubs and device kernels disagree
    watched = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "..", "include", "pb3d.h")]
    stamp = {f: os.stat(f).st_mtime_ns for f in watched}
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, *["-D" + d for d in defines], "-I", os.path.join(HERE, "..", "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write("hipcc failed for %s:\n%s\n" % (src, out.decode(errors="replace")))
        elif verbose and out.strip():
            sys.stderr.write(out.decode(errors="replace")[-2000:])
    if failed:
        raise RuntimeError("libpb3d build failed")
    changed = [f for f in watched if os.stat(f).st_mtime_ns != stamp[f]]
    if changed:
        raise RuntimeError("sources changed during the build (%s): build again" % ", ".join(map(os.path.basename, changed)))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("libpb3d link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, lib)
    if verbose:
        print("built %s in %.1f s" % (lib, time.time() - t0))
    return lib


if __name__ == "__main__":
    build()
