"""Build libkrt_emu.so: the library's sources (artemis_amd/csrc) compiled for
the CPU under AddressSanitizer and UBSan against the HIP emulation in
tests/emu/hip/hip_runtime.h.  Test infrastructure only: the product is
libkrt.so; this copy exists so that a host run can check every memory access
the kernels make against the bounds of its allocation.

Each `kernel<<<grid, block, shmem, stream>>>(args);` launch is rewritten into
`krt_emu::launch("kernel", grid, block, shmem, stream,
krt_emu::bind_args([](auto &&...a) { kernel(a...); }, args));` -- the
arguments are evaluated and copied at the launch, as a real launch does;
the only other changes are that the two-level (instanced) path, which the
GPU build keeps switched off, is switched on (kTwoLevelEnabled), so that the
host run can exercise it, that the BVH build's top levels use 256-lane
workgroups as its lower ones do (kBuildWide), and that the top levels'
pre-passes start at 64 primitives instead of 65536 (kPreMinPrims).
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
# another checkout's sources can be built too (KRT_EMU_ROOT: its root; the
# library then goes to KRT_EMU_BUILD)
SRC_ROOT = Path(os.environ.get("KRT_EMU_ROOT", ROOT))
CSRC = SRC_ROOT / "artemis_amd" / "csrc"
BUILD = Path(os.environ.get("KRT_EMU_BUILD", HERE / "_build"))
LIB = BUILD / "libkrt_emu.so"
CLANG = "/opt/rocm/llvm/bin/clang++"

FLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-fno-omit-frame-pointer", "-ffp-contract=off",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-pthread",
         "-Wno-unused-command-line-argument", "-Wno-unknown-attributes", "-Wno-ignored-attributes",
         "-I", str(HERE)]


def asan_runtime() -> str:
    r = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if not Path(path).exists():
        raise RuntimeError("clang's AddressSanitizer runtime was not found")
    return path


_IDENT = re.compile(r"[A-Za-z0-9_:]")


def _callee_start(src: str, end: int) -> int:
    """Start of the kernel expression that ends just before `end` (an
    identifier, possibly qualified, possibly with template arguments)."""
    i = end
    while i > 0 and src[i - 1].isspace():
        i -= 1
    if src[i - 1] == ">":  # template arguments
        depth = 0
        while True:
            i -= 1
            c = src[i]
            if c == ">":
                depth += 1
            elif c == "<":
                depth -= 1
                if depth == 0:
                    break
    while i > 0 and _IDENT.match(src[i - 1]):
        i -= 1
    return i


def _balanced(src: str, start: int, open_: str, close: str) -> int:
    """Index just past the bracket that closes the one at `start`."""
    depth = 0
    for i in range(start, len(src)):
        if src[i] == open_:
            depth += 1
        elif src[i] == close:
            depth -= 1
            if depth == 0:
                return i + 1
    raise ValueError("unbalanced brackets")


def rewrite_launches(src: str) -> tuple[str, int]:
    out, pos, n = [], 0, 0
    while True:
        k = src.find("<<<", pos)
        if k < 0:
            out.append(src[pos:])
            return "".join(out), n
        c0 = _callee_start(src, k)
        callee = src[c0:k].strip()
        e = src.find(">>>", k)
        cfg = src[k + 3:e]
        p = e + 3
        while src[p].isspace():
            p += 1
        assert src[p] == "(", f"launch of {callee} without arguments"
        q = _balanced(src, p, "(", ")")
        args = src[p + 1:q - 1]
        parts, depth, cur = [], 0, ""
        for ch in cfg:  # split the launch configuration at top-level commas
            if ch in "([{":
                depth += 1
            elif ch in ")]}":
                depth -= 1
            if ch == "," and depth == 0:
                parts.append(cur)
                cur = ""
            else:
                cur += ch
        parts.append(cur)
        while len(parts) < 4:
            parts.append("0" if len(parts) == 2 else "nullptr")
        g, b, s, st = (x.strip() for x in parts)
        name = re.sub(r"\s+", " ", callee)
        out.append(src[pos:c0])
        sep = ", " if args.strip() else ""
        out.append(f'krt_emu::launch("{name}", dim3({g}), dim3({b}), (size_t)({s}), ({st}), '
                   f"krt_emu::bind_args([](auto &&...a) {{ {callee}(a...); }}{sep}{args}))")
        pos = q
        n += 1


def prepare() -> list[Path]:
    # the sources include "../../include/krt.h": keep them two levels below
    # the directory that holds include/
    src_dir = BUILD / "tree" / "src"
    if src_dir.exists():
        shutil.rmtree(src_dir)
    src_dir.mkdir(parents=True)
    units = []
    for f in sorted(CSRC.iterdir()):
        if f.suffix not in (".hip", ".h", ".cpp"):
            continue
        text, _ = rewrite_launches(f.read_text())
        # AMDGPU-only function attributes have no meaning on the host
        text = re.sub(r"__attribute__\(\(amdgpu_[a-z_]+\([^)]*\)\)\)", "", text)
        if f.name == "krt_bvh.hip":
            flag = "constexpr bool kTwoLevelEnabled = false;"
            assert flag in text, "kTwoLevelEnabled not found"
            text = text.replace(flag, "constexpr bool kTwoLevelEnabled = true;")
            # the build's result does not depend on its workgroup width (the
            # GPU tests compare both widths' trees with the oracle); one
            # emulated lane is a fiber, so the host run keeps the narrow one
            wide = "constexpr int kBuildWide = 1024;"
            assert wide in text, "kBuildWide not found"
            text = text.replace(wide, "constexpr int kBuildWide = 256;")
            # the top levels' multi-workgroup pre-passes run from 64 primitives
            # on, so that the host run's small scenes take them too
            pre = "constexpr uint32_t kPreMinPrims = 1u << 16;"
            assert pre in text, "kPreMinPrims not found"
            text = text.replace(pre, "constexpr uint32_t kPreMinPrims = 1u << 6;")
        (src_dir / f.name).write_text(text)
        if f.suffix in (".hip", ".cpp"):
            units.append(src_dir / f.name)
    inc = BUILD / "include"
    inc.mkdir(exist_ok=True)
    shutil.copy(SRC_ROOT / "include" / "krt.h", inc / "krt.h")
    return units


def _compile(unit: Path) -> Path:
    obj = BUILD / (unit.stem + ".o")
    cmd = [CLANG, *FLAGS, "-x", "c++", "-c", str(unit), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu compile failed for {unit.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> Path:
    units = prepare()
    runtime = HERE / "emu_runtime.cpp"
    jobs = min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, [*units, runtime]))
    cmd = [CLANG, *FLAGS, "-shared", "-o", str(LIB), *map(str, objs), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu link failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
