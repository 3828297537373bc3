"""Plugin build helpers (reference: system/tmine/utils.py:1-53).

The worker plugin is compiled with hipcc for gfx950 (the reference runs nvcc
-arch=sm_86).  Compiled plugins are cached by a hash of their source, so a
query shape seen before loads without recompiling.
"""
import hashlib
import logging
import os
import re
import shutil
import subprocess

from . import locations

log = logging.getLogger("everest.tmine")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EVEREST_OFFLOAD_ARCH", "gfx950")


def libNameToLibFile(name):
    return "lib" + name + ".so"


def libName(filename):
    return re.match(r".*lib(\w+)\.so", filename).group(1)


def libFileToLibName(filename):
    return re.match(r".*lib(\w+)\.so", filename).group(1)


def compile_command(src, out):
    inc = []
    for d in locations.includeDirs():
        inc += ["-I", d]
    lib = locations.buildDir()
    return ([HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + inc +
            [src, "-o", out, "-L" + lib, "-leverest_core", "-ltm_hip", "-Wl,-rpath," + lib])


def compileModule(file, out):
    """Compile one worker source into a shared library; raises on failure."""
    cmd = compile_command(file, out)
    log.debug("build Lib: %s", " ".join(cmd))
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("plugin build failed:\n" + p.stdout)
    return out


def headers_digest():
    """Hash of every header a worker includes: a kernel or ABI change makes
    cached plugins stale even when the generated source is the same."""
    h = hashlib.sha256()
    for d in locations.includeDirs():
        for root, _, files in sorted(os.walk(d)):
            for f in sorted(files):
                if f.endswith((".h", ".cuh", ".hpp")):
                    with open(os.path.join(root, f), "rb") as fh:
                        h.update(f.encode() + fh.read())
    return h.hexdigest()


def compilePlugin(source=None, out=None):
    """Compile the generated worker (cached by source hash) into libDYN.so."""
    source = source or locations.dynLibSource()
    out = out or locations.dynLibPath()
    createDirIfNotExist(os.path.dirname(out))
    createDirIfNotExist(locations.pluginCacheRoot())
    with open(source, "rb") as f:
        key = hashlib.sha256(f.read() + ARCH.encode() + headers_digest().encode()).hexdigest()[:24]
    cached = os.path.join(locations.pluginCacheRoot(), "lib%s.so" % key)
    if not os.path.exists(cached):
        tmp = cached + ".tmp%d" % os.getpid()
        compileModule(source, tmp)
        os.replace(tmp, cached)
    tmp = out + ".tmp%d" % os.getpid()
    shutil.copyfile(cached, tmp)
    os.replace(tmp, out)  # a new inode: a re-opened libDYN.so is the new worker
    return out


def allLibs(dir):
    files = os.listdir(dir)
    so_files = [os.path.join(dir, f) for f in files if f.endswith(".so")]
    return [libName(f) for f in so_files], so_files


def allTxts(dir):
    return [os.path.join(dir, f) for f in os.listdir(dir) if f.endswith(".txt")]


def createDirIfNotExist(dirpath):
    os.makedirs(dirpath, exist_ok=True)
