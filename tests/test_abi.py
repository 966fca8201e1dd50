"""The C-ABI boundary without a GPU: the library loads, exports every function
include/emmi_mgfd.h declares, and the Python carriers have exactly the struct
layouts of that header (checked by a probe gcc compiles from the header).
When the reference's headers are present, the carriers are also checked to be
layout-identical to include/emf.h, include/acq.h and include/fwi.h there."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

from libemmi_mgfd_amd import carriers, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "emmi_mgfd.h")
REF_INCLUDE = "/root/reference/include"


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = []
    for m in re.finditer(r"^[ \t]*(?:const[ \t]+)?[A-Za-z_][\w]*[ \t\*]+\**[ \t]*([A-Za-z_]\w*)[ \t]*\(", text, re.M):
        if m.group(1) not in ("if", "while", "return", "sizeof"):
            names.append(m.group(1))
    return sorted(set(names))


def exported_functions():
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_library_loads_without_gpu():
    assert native.lib() is not None


def test_every_declared_function_is_exported():
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [n for n in decl if n not in exported_functions()]
    assert not missing, missing


def test_reference_entry_points_present():
    """The symbols do_modelling/fg_fwi bind (src/do_modelling.c:14-26,
    src/fg_fwi.c:19-41)."""
    want = ["gmg_init", "gmg_apply", "gmg_free", "inject_source", "inject_adj_source", "extract_emf_data",
            "extract_electric_field", "extract_magnetic_field", "build_gradient", "regrid_init", "regrid_free"]
    ex = exported_functions()
    assert [w for w in want if w not in ex] == []


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HDR
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m));
#define S(T) printf(#T " sizeof %zu\n", sizeof(T));
int main(void) {
BODY
  return 0;
}
"""


def probe_layout(header, structs, cflags=()):
    lines = []
    for tname, ctype in structs:
        lines.append(f"  S({tname})")
        for fname, _ in ctype._fields_:
            lines.append(f"  F({tname}, {fname})")
    src = PROBE.replace("BODY", "\n".join(lines))
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=gnu11", f'-DHDR="{header}"', *cflags, c, "-o", exe], check=True,
                       capture_output=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        t, m, v = line.split()
        res[(t, m)] = int(v)
    return res


STRUCTS = [("emf_t", carriers.Emf), ("acq_t", carriers.Acq), ("fwi_t", carriers.Fwi)]


def test_carriers_match_header_layout():
    lay = probe_layout(HEADER, STRUCTS)
    for tname, ctype in STRUCTS:
        assert lay[(tname, "sizeof")] == C.sizeof(ctype), tname
        for fname, _ in ctype._fields_:
            assert lay[(tname, fname)] == getattr(ctype, fname).offset, (tname, fname)


def test_params_and_stats_layout():
    lay = probe_layout(HEADER, [("emmi_params_t", native.Params), ("emmi_stats_t", native.Stats)])
    for tname, ctype in (("emmi_params_t", native.Params), ("emmi_stats_t", native.Stats)):
        assert lay[(tname, "sizeof")] == C.sizeof(ctype)
        for fname, _ in ctype._fields_:
            assert lay[(tname, fname)] == getattr(ctype, fname).offset, (tname, fname)


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present")
def test_carriers_layout_identical_to_reference_headers():
    ours = probe_layout(HEADER, STRUCTS)
    for hdr, tname, ctype in (("emf.h", "emf_t", carriers.Emf), ("acq.h", "acq_t", carriers.Acq),
                              ("fwi.h", "fwi_t", carriers.Fwi)):
        theirs = probe_layout(os.path.join(REF_INCLUDE, hdr), [(tname, ctype)], ["-include", "complex.h"])
        assert theirs[(tname, "sizeof")] == ours[(tname, "sizeof")], tname
        for fname, _ in ctype._fields_:
            assert theirs[(tname, fname)] == ours[(tname, fname)], (tname, fname)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_missing_library_fails_loudly(tmp_path):
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from libemmi_mgfd_amd import native\n"
            "native.LIB_PATH = %r\n"
            "try:\n    native.lib()\nexcept native.NativeError as e:\n    print('refused', e)\n") % (
        ROOT, str(tmp_path / "absent.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert "refused" in r.stdout
