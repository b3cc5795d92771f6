"""Builds the library's device sources for the CPU (tests/emu/_build/driver):
each .hip file is rewritten by convert.py, then compiled with g++ against the
emulation headers in this directory and AddressSanitizer.  Test
infrastructure only: the product is the gfx950 build."""
import os
import pathlib
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
SRC = ROOT / "clustersm_amd" / "csrc"
OUT = HERE / "_build"
FLAGS = ["-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address", "-pthread", f"-I{HERE}", f"-I{ROOT / 'include'}", f"-I{OUT / 'gen'}",
         f"-I{SRC / 'device'}", "-Wno-unused-result"]


def build():
    sys.path.insert(0, str(HERE))
    import convert
    (OUT / "gen").mkdir(parents=True, exist_ok=True)
    (OUT / "obj").mkdir(parents=True, exist_ok=True)
    sources = []
    for f in sorted((SRC / "device").glob("*.hip")):
        dst = OUT / "gen" / (f.stem + ".cpp")
        convert.convert(str(f), str(dst))
        sources.append(dst)
    for f in sorted((SRC / "device").glob("*.hpp")):
        (OUT / "gen" / f.name).write_text(f.read_text())
    sources += sorted((SRC / "host").glob("*.cpp")) + [HERE / "emu_rt.cpp"]
    newest_dep = max(p.stat().st_mtime for p in [HERE / "hip" / "hip_runtime.h", HERE / "emu_rt.cpp"])

    def compile_one(src):
        obj = OUT / "obj" / (pathlib.Path(src).stem + ".o")
        if obj.exists() and obj.stat().st_mtime > max(pathlib.Path(src).stat().st_mtime, newest_dep):
            return obj
        subprocess.run(["g++", *FLAGS, "-c", str(src), "-o", str(obj)], check=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, sources))
    exe = OUT / "driver"
    subprocess.run(["g++", *FLAGS, "-o", str(exe), str(HERE / "driver.cpp"), *map(str, objs)], check=True)
    return exe


if __name__ == "__main__":
    print(build())
