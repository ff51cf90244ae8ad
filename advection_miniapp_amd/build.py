"""Builds libdgops.so (HIP for gfx950) in-tree.

    python -m advection_miniapp_amd.build [-j N] [--force]

Compiles every translation unit of csrc/ (and the generated kernel
instances) with hipcc into _build/, then links advection_miniapp_amd/
libdgops.so. Object files are rebuilt when a source or header changed.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgops.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 instructions stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{CSRC}", f"-I{INCLUDE}",
]


def sources():
    sys.path.insert(0, CSRC)
    import gen_instances  # noqa: E402
    gen = gen_instances.main()
    own = sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                 glob.glob(os.path.join(CSRC, "*.cpp")))
    return own + gen


def headers():
    return (glob.glob(os.path.join(CSRC, "*.hpp")) +
            glob.glob(os.path.join(INCLUDE, "*.h")))


def obj_name(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(BUILD, rel + ".o")


def compile_one(src, force, hdr_time):
    obj = obj_name(src)
    if (not force and os.path.exists(obj)
            and os.path.getmtime(obj) >= max(os.path.getmtime(src), hdr_time)):
        return obj, None
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(jobs=None, force=False, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    hdr_time = max(os.path.getmtime(h) for h in headers())
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs, errors = [], []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = {ex.submit(compile_one, s, force, hdr_time): s for s in srcs}
        for f in cf.as_completed(futs):
            obj, err = f.result()
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("compilation failed:\n" + "\n".join(errors[:3]))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + \
        sorted(objs)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} objects")
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    build(a.j, a.force)


if __name__ == "__main__":
    main()
