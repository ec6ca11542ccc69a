"""The C-ABI library loads and exports every entry point that
include/canvaz_film.h declares; host-only entry points answer without a
device; device entry points fail with an error (not a crash) when no HIP
device is present."""
import ctypes as C
import os
import re
import subprocess

import pytest

import canvaz_film as czf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include"))) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(cz_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 40
    czf.lib()  # loads torch's HIP runtime first, as every user of the mirror does
    lib = C.CDLL(czf.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", czf.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (cz_\w+)", out))
    assert set(names) <= exported


def test_host_entry_points_without_device():
    s = czf.default_settings()
    L = czf.lib()
    d = czf.default_settings()
    L.cz_settings_default(C.byref(d))
    assert bytes(d) == bytes(s)
    assert L.cz_version().decode().startswith("canvaz-film")


def test_create_without_device_is_an_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    ptr = C.c_void_p()
    rc = czf.lib().cz_film_create(64, 64, None, None, None, C.byref(ptr))
    assert rc < 0
    assert czf.lib().cz_last_error()
    assert not ptr.value


def test_library_shares_torchs_hip_runtime():
    """One HIP runtime per process: the mirror loads torch's before the
    library, which then binds to it instead of mapping a second copy (a
    second runtime sees no device once the first has opened it)."""
    pytest.importorskip("torch")
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import canvaz_film as czf; czf.lib(); "
            "m = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}; print(len(m))"
            % os.path.join(ROOT, "paint-sim-3d_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and
    field offsets (a field added on one side only shifts everything after it)."""
    import cz_types as T
    pairs = {"cz_params": T.Params, "cz_deposit": T.Deposit, "cz_contact": T.Contact, "cz_knife": T.Knife,
             "cz_scrap": T.Scrap, "cz_settings": T.Settings, "cz_overrides": T.Overrides, "cz_stats": T.Stats,
             "cz_mass_report": T.MassReport, "cz_mass_totals": T.MassTotals, "cz_host_state": T.HostState,
             "cz_scene": T.Scene, "cz_request": T.Request, "cz_result": T.Result,
             "cz_server_config": T.ServerConfig, "cz_server_stats": T.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "canvaz_film.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'  printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, v = ln.split()
            got[(a, b)] = int(v)
    for cname, py in pairs.items():
        assert got[(cname, "__sizeof__")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
