"""include/krt_armada.hpp compiles against include/krt.h and links to libkrt.so
(host compiler only; nothing is launched)."""
import shutil
import subprocess
import textwrap
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_cpp_drop_in_layer_compiles_and_links(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    lib = ROOT / "artemis_amd" / "libkrt.so"
    if not lib.exists():
        pytest.fail("libkrt.so is not built")
    src = tmp_path / "use.cpp"
    src.write_text(textwrap.dedent("""
        #include "krt_armada.hpp"
        #include <cstdio>
        int main(int argc, char **) {
            if (argc > 5) {  // never taken: only proves the calls link
                krt_cpp::Context ctx(0);
                krt_cpp::Accelerator sys(ctx);
                krt_scene_desc d{};
                krt_handle h = sys.update(d);
                krt_cpp::ArmadaRTX a(sys, {64, 64});
                a.attach("pt", std::make_shared<krt_cpp::PathTracer>(ctx));
                a.set_option("seed", 3);
                a.set_depth(5);
                krt_cpp::Camera cam{};
                a.render(h, cam, true);
                uint32_t *ldr = nullptr;
                krt_cpp::hdr_to_ldr(ctx, a.color_buffer(), ldr, 64, 64, krt_cpp::eTonemappingSRGB);
                krt_attachment_ret r = krt_load_attachment(ctx.get(), "restir");
                std::printf("%s\\n", r.name);
            }
            std::printf("%s|\\n", krt_last_error());
            return 0;
        }
    """))
    exe = tmp_path / "use"
    r = subprocess.run([gxx, "-std=c++17", f"-I{ROOT / 'include'}", str(src), str(lib),
                        f"-Wl,-rpath,{lib.parent}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "|"
