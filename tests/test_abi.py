"""CPU checks of the drop-in boundary: the library loads, exports every
symbol the headers declare, and the FF struct layout of include/forcefield.h
is what the ctypes binding and C callers see."""
import ctypes
import os
import re
import subprocess

import pytest

from helpers import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("forcefield.h", "msmpb_hip.h", "msm_bridge.h")]
PME_HEADERS = [os.path.join(ROOT, "include", h) for h in ("pme.h", "pme_hip.h")]


def declared_functions(headers=HEADERS):
    names = set()
    for h in headers:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            prev = text[:m.start()].rstrip()
            if name in ("if", "sizeof", "__typeof__", "double", "int", "void", "char", "bool") \
                    or prev.endswith("("):
                continue
            if re.search(r"(typedef|struct|enum)\s*$", prev):
                continue
            names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from msmpb_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.EXPORTS)


def test_ff_struct_layout_matches_c(tmp_path):
    from msmpb_amd._lib import FFStruct
    fields = [f[0] for f in FFStruct._fields_]
    src = tmp_path / "off.c"
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(FF, {f}));' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "forcefield.h"\n'
                   f"int main(void) {{\n{body}\n  printf(\"sizeof %zu\\n\", sizeof(FF));\n}}\n")
    exe = tmp_path / "off"
    subprocess.check_call(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for f in fields:
        assert int(out[f]) == getattr(FFStruct, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(FFStruct)


def test_reference_header_layout_offsets():
    """Offsets fixed by code/forcefield.h:22-69 on LP64 (pointers 8 bytes,
    the three FFTW members pointer-sized)."""
    from msmpb_amd._lib import FFStruct
    assert FFStruct.q.offset == 8
    assert FFStruct.A.offset == 16 and FFStruct.Ai.offset == 88
    assert FFStruct.detA.offset == 160
    assert FFStruct.orderAcc.offset == 208 and FFStruct.topGridDim.offset == 216
    assert FFStruct.nlist.offset == 232 and FFStruct.nlist_interaction_count.offset == 244
    assert FFStruct.FFT.offset == 248 and FFStruct.fftw_in.offset == 256
    assert FFStruct.tau.offset == 280 and FFStruct.nLim.offset == 328
    assert FFStruct.cutoff.offset == 344 and FFStruct.coeff1.offset == 384
    assert FFStruct.time_partcl2partcl.offset == 400 and FFStruct.time_padding.offset == 752
    assert ctypes.sizeof(FFStruct) == 832


def test_cli_built_and_links():
    cli = os.path.join(ROOT, "msmpb_amd", "bin", "msmpb")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: msmpb" in r.stderr


def test_pme_library_exports_every_declared_symbol():
    """libpme_hip.so implements include/pme.h (the reference's code/pme.h)
    and pme_hip.h; FF_last_error comes from libmsmpb_hip.so."""
    from msmpb_amd import _lib
    P = _lib.pme_lib()
    names = declared_functions(PME_HEADERS)
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(P, n)]
    assert not missing, missing
    assert set(names) - {"FF_last_error"} <= set(_lib.PME_EXPORTS)
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.PME_LIB_PATH], text=True)
    defined = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(_lib.PME_EXPORTS) <= defined


def test_pme_struct_layout_matches_c(tmp_path):
    from msmpb_amd._lib import PMEStruct
    fields = [f[0] for f in PMEStruct._fields_]
    src = tmp_path / "off.c"
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(FF, {f}));' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "pme.h"\n'
                   f"int main(void) {{\n{body}\n  printf(\"sizeof %zu\\n\", sizeof(FF));\n}}\n")
    exe = tmp_path / "off"
    subprocess.check_call(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for f in fields:
        assert int(out[f]) == getattr(PMEStruct, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(PMEStruct)


def test_pme_reference_header_layout_offsets():
    """Offsets fixed by code/pme.h:15-44 on LP64 (pointers 8 bytes, the
    three FFTW members pointer-sized)."""
    from msmpb_amd._lib import PMEStruct
    assert PMEStruct.A.offset == 16 and PMEStruct.detA.offset == 160
    assert PMEStruct.cutoff.offset == 168 and PMEStruct.orderAcc.offset == 176
    assert PMEStruct.topGridDim.offset == 180 and PMEStruct.nlist.offset == 192
    assert PMEStruct.nlist_interaction_count.offset == 204 and PMEStruct.tolDir.offset == 208
    assert PMEStruct.fftw_in.offset == 216 and PMEStruct.beta.offset == 240
    assert PMEStruct.khat.offset == 280 and PMEStruct.coeff1.offset == 288
    assert PMEStruct.time_partcl2partcl.offset == 304
    assert ctypes.sizeof(PMEStruct) == 352


def test_pme_cli_built_and_links():
    cli = os.path.join(ROOT, "msmpb_amd", "bin", "pme")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: pme" in r.stderr
