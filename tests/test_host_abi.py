"""Host side of the C ABI, without a GPU.

* libaem.so loads and exports every entry point that include/aem.h declares,
  and the ctypes table in alpineem_amd/_lib.py binds exactly that set, each
  with the declared argument count, widths and pointer positions;
* the ctypes mirrors of aem_plan_view / aem_port_view have the C layout
  (offsets printed by a program compiled against include/aem_plan.h);
* deck errors come back as status codes with the reference's STOP text
  (main_fdtd/fdtd_solver.f90:809-829, 1116-1160) instead of ending the process.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from alpineem_amd import _lib
from alpineem_amd.solver import Plan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(aem_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    names = _declared("aem.h")
    assert len(names) >= 20, names
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert bound == _declared("aem.h")


_SCALARS = {"int32_t": C.c_int32, "int64_t": C.c_int64, "float": C.c_float, "double": C.c_double,
            "aem_status": C.c_int}
_OPAQUE = ("aem_plan", "aem_sim", "aem_batch", "aem_sim_options", "aem_plan_view", "aem_plan_options")


def _prototypes(header):
    """name -> (return type, [parameter declarations]) of every aem_ function."""
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"\s+", " ", text)
    out = {}
    for ret, name, params in re.findall(r"([A-Za-z_][\w ]*?\s*\**)\s*\b(aem_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ps = [p.strip() for p in params.split(",")]
        out[name] = (ret.strip(), [] if ps == ["void"] else ps)
    return out


def _ctype_ok(decl, ct):
    """Whether ctypes type ct passes a C value of declaration decl."""
    stars = decl.count("*")
    base = re.sub(r"\bconst\b", "", decl.split("*")[0]).split()
    base = base[0] if base else ""
    if stars == 0:
        if decl in ("void",):
            return ct is None
        return ct is _SCALARS[base]
    if stars == 1 and base == "char":
        return ct is C.c_char_p
    if stars >= 2:
        return ct in (C.c_void_p, C.POINTER(C.c_void_p))
    if base in _OPAQUE:
        return ct is C.c_void_p
    return ct is C.c_void_p or ct is C.POINTER(_SCALARS[base])


def test_ctypes_arguments_match_prototypes():
    """Every bound entry point passes as many arguments as include/aem.h
    declares, each of the declared width (int32/int64/float/double), pointer
    or not; the return types agree too."""
    protos = _prototypes("aem.h")
    bad = []
    for name, res, args in _lib.SIGNATURES:
        ret, params = protos[name]
        if len(params) != len(args):
            bad.append(f"{name}: {len(params)} parameters declared, {len(args)} bound")
            continue
        for p, a in zip(params, args):
            decl = re.sub(r"\s*\b\w+$", "", p) if not p.endswith("*") else p   # drop the parameter name
            if not _ctype_ok(decl, a):
                bad.append(f"{name}: '{p}' bound as {a}")
        if not _ctype_ok(ret, res):
            bad.append(f"{name}: returns '{ret}', bound as {res}")
    assert not bad, bad


def test_version_and_error_strings():
    lib = _lib.lib()
    assert lib.aem_version().decode()
    st = lib.aem_run(None, 1, 1)
    assert st == 4                                   # AEM_ERR_STATE
    assert "null" in lib.aem_last_error().decode()


def _c_offsets(tmp_path, struct, fields):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "aem_plan.h"', "int main(void) {",
           f'  printf("sizeof %zu\\n", sizeof({struct}));']
    src += [f'  printf("{f} %zu\\n", offsetof({struct}, {f}));' for f in fields]
    src += ["  return 0;", "}"]
    c_file = tmp_path / "offs.c"
    c_file.write_text("\n".join(src))
    exe = tmp_path / "offs"
    subprocess.run([cc, "-I", INCLUDE, str(c_file), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {a: int(b) for a, b in (line.split() for line in out.splitlines())}


@pytest.mark.parametrize("struct,cls", [("aem_plan_view", _lib.PlanView), ("aem_port_view", _lib.PortView),
                                        ("aem_plan_options", _lib.PlanOptions), ("aem_spice_view", _lib.SpiceView)])
def test_struct_layout_matches_header(tmp_path, struct, cls):
    fields = [f[0] for f in cls._fields_]
    offs = _c_offsets(tmp_path, struct, fields)
    assert offs["sizeof"] == C.sizeof(cls)
    for f in fields:
        assert offs[f] == getattr(cls, f).offset, f


def _deck(golden_dir, name):
    return open(os.path.join(golden_dir, "decks", name + ".txt")).read().splitlines()


def _expect_input_error(lines, fragment):
    with pytest.raises(_lib.AemError) as e:
        Plan.from_text("\n".join(lines) + "\n")
    assert e.value.status == 1                      # AEM_ERR_INPUT
    assert fragment in str(e.value)


def test_port_on_first_cube_is_rejected(golden_dir):
    lines = _deck(golden_dir, "dipole36")
    assert lines[39] == "18,18,18"                   # port location of the basic port
    lines[39] = "1,18,18"
    _expect_input_error(lines, "INTERNAL LUMPED PORT AT A FIRST CUBE")


def test_oblique_plane_wave_with_pbc_is_rejected(golden_dir):
    lines = _deck(golden_dir, "drude_slab")
    assert lines[6] == "1,1,1,1,0,0" and lines[10] == "0,0,270"
    lines[10] = "30,0,270"                           # theta = 30 deg with periodic x and y
    _expect_input_error(lines, "OBLIQUE ANGLE W/ RESPECT TO PBC")


def test_bad_number_and_truncated_deck(golden_dir):
    lines = _deck(golden_dir, "dipole36")
    bad = list(lines)
    bad[2] = "two hundred"
    _expect_input_error(bad, "bad integer")
    _expect_input_error(lines[:20], "end of deck")


def test_features_outside_the_path_are_refused(golden_dir):
    lines = _deck(golden_dir, "dipole36")
    assert lines[-1] == "0"                          # number of SPICE ports
    spice = lines[:-1] + ["1", "net.cir", "gridded", "-x", "20,18,18", "4,4", "feed.bin", "0", "out", "I", "0"]
    with pytest.raises(_lib.AemError) as e:
        Plan.from_text("\n".join(spice) + "\n")
    assert e.value.status == 5                      # AEM_ERR_UNSUPPORTED
    assert "gridded SPICE" in str(e.value)
    _expect_input_error(lines[:-1] + ["1"], "end of deck")    # a SPICE section cut short
    assert lines[7] == "none"                        # IGP flag
    # an infinite ground plane (f90:1402-1445): the -z face of the TF/SF box
    # moves to the plane, loses its corrections and far-field currents, and
    # that side's CPML shrinks to one surface
    igp = lines[:7] + ["IGP", "-z", "14"] + lines[8:]
    v = Plan.from_text("\n".join(igp) + "\n").view
    assert v.mirror_type == 5 and list(v.wall) == [1, 1, 1, 1, 0, 1]
    assert v.tf_lo[2] == 15 and v.npml[4] == 1 and v.ff_lo[2] == 16
    assert list(v.use_m) == [0, 0, 1] and v.off_m[2] == 30.0
    assert list(v.sgn_M) == [1, 1, -1] and list(v.sgn_J) == [-1, -1, 1]


def test_missing_deck_file_is_an_io_error(tmp_path):
    with pytest.raises(_lib.AemError) as e:
        Plan.from_inputs(str(tmp_path / "no_such_inputs.txt"))
    assert e.value.status == 2                      # AEM_ERR_IO
