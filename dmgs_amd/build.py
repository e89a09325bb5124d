"""Build the gfx950 HIP library ``dmgs_amd/lib/libdmgs_hip.so`` in-tree.

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 only and linked
into one shared library exposing the C-ABI of ``include/dmgs.h``.

    python -m dmgs_amd.build            # build (always recompiles)
    python -m dmgs_amd.build --asan     # host-side AddressSanitizer build (CPU checks only)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIBNAME = "libdmgs_hip.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # hardware float atomics on coarse-grained device memory
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build dmgs_amd")
    return exe


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# AddressSanitizer on the host code only (argument checks, workspace layout,
# view preparation); the device code is built as usual
ASAN_HOST_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
ASAN_DIR = os.path.join(LIBDIR, "asan")


def _compile(src, verbose, objdir=OBJDIR, extra=()):
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + ["-I", CSRC, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


def asan_library_path():
    return os.path.join(ASAN_DIR, LIBNAME)


def build(verbose=False, asan=False):
    """Compile every HIP source for gfx950 and link the shared library
    (asan: a second build with AddressSanitizer on the host code, into
    dmgs_amd/lib/asan/, loaded only by the CPU tests that check the C-ABI's
    host side under the sanitizer)."""
    objdir = os.path.join(ASAN_DIR, "obj") if asan else OBJDIR
    extra = ASAN_HOST_FLAGS if asan else []
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, objdir, extra), srcs))
    out = asan_library_path() if asan else library_path()
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC"] + (["-Xarch_host", "-fsanitize=address"] if asan
                                                                      else []) + ["-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, asan="--asan" in sys.argv))
