"""Device code writes memory through vector instructions only: no scalar
memory store, scalar atomic or scalar-cache write-back / discard appears in
the gfx950 code of the engine.  CPU only (compiles to assembly); this file is
not needed on GPU hosts."""
import os
import re
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_no_scalar_memory_writes(native):
    from openems_amd import build as B
    src = os.path.join(ROOT, "openems_amd", "csrc", "engine", "ems_engine.hip")
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "engine.s")
        subprocess.run([B.HIPCC] + B.device_flags() + ["--cuda-device-only", "-S", "-o", asm, src], check=True,
                       capture_output=True)
        text = open(asm).read()
    assert "k_step" in text and "k_volt" in text
    bad = re.compile(r"^\s*s_(store|atomic|dcache_wb|dcache_discard|buffer_store|scratch_store)\w*", re.M)
    hits = bad.findall(text)
    assert not hits, hits[:5]
