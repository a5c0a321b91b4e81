"""The library's kernels and host code under AddressSanitizer and UBSan on the
CPU (tests/emu: the sources compiled by clang for the host against a HIP
emulation whose lanes are fibers, every device allocation filled with
random bytes).  Each case runs the Python API over that build in a child
process with the sanitizer runtime preloaded and checks the results against
the oracle bit for bit; any invalid access, use of freed memory, undefined
behaviour, or a result that depends on memory nothing wrote, fails it.

* attachments: path tracer, ReSTIR DI and ReSTIR GI under a moving camera,
  the editor's G-buffer ReSTIR, a served request, the cached path tracer;
* envmap: the path tracer and ReSTIR DI with the armada's environment map;
* textures: material textures and normal maps in the path tracer, ReSTIR DI
  under a moving camera and a served request;
* refit: moved vertices, the bottom-up refit of both trees and the lights;
* mlp: the radiance-cache network's inference and gradients (emulated MFMA);
* instancing: the sequence of tests/test_gpu_instancing.py on the two-level
  path (which the GPU build keeps switched off): per-mesh and instance trees,
  closest hits and occlusion, path-tracer frames on both tree layouts,
  ReSTIR DI reservoirs, and the refusals of the flat-only attachments.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EMU = ROOT / "tests" / "emu"


@pytest.fixture(scope="module")
def emu_env():
    sys.path.insert(0, str(EMU))
    try:
        import build as emu_build
    finally:
        sys.path.remove(str(EMU))
    if not Path(emu_build.CLANG).exists():
        pytest.skip("clang++ of ROCm not found")
    emu_build.build()
    env = dict(os.environ)
    env["LD_PRELOAD"] = emu_build.asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["KRT_EMU_FILL"] = "random"
    return env


@pytest.mark.parametrize("case", ["attachments", "envmap", "textures", "refit", "mlp", "instancing"])
def test_sanitized_case(emu_env, case):
    r = subprocess.run([sys.executable, str(EMU / "run_case.py"), case], env=emu_env, capture_output=True,
                       text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0, f"case {case} failed:\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    assert f"case {case}: ok" in r.stdout
