"""The C-ABI library builds for gfx950, loads, and exports every symbol that
include/clarabel_hip.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

from helpers import ROOT


def test_header_and_export_list_agree():
    from cuclarabel_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "clarabel_hip.h")).read()
    declared = set(re.findall(r"\b(clh_[A-Za-z_]+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)


def test_library_exports_every_symbol():
    from cuclarabel_amd import _lib
    if not os.path.exists(_lib.DEFAULT_LIB):
        from cuclarabel_amd import build
        build.build()
    L = ctypes.CDLL(_lib.DEFAULT_LIB)
    for sym in _lib.EXPORTS:
        assert hasattr(L, sym), sym


def test_settings_struct_size_matches_header_defaults():
    import cuclarabel_amd as ca
    from cuclarabel_amd import _lib
    L = _lib.bind(_lib.DEFAULT_LIB)
    s = _lib.ClhSettings()
    assert L.clh_default_settings(ctypes.byref(s)) == 0
    d = ca.Settings().to_c()
    for f, _ in s._fields_:
        if f == "schedule":  # nested: every field 0 (the built-in rules)
            for g, _ in s.schedule._fields_:
                assert getattr(s.schedule, g) == getattr(d.schedule, g) == 0, g
            continue
        assert getattr(s, f) == pytest.approx(getattr(d, f)), f


def test_julia_struct_mirrors_match_the_abi():
    """the Julia glue's struct mirrors (cuclarabel_amd/julia/clh_abi.jl) hold
    the same fields, in the same order and C types, as the ctypes mirrors the
    library is tested through"""
    import ctypes
    import os
    import re

    from cuclarabel_amd import _lib
    src = open(os.path.join(os.path.dirname(_lib.__file__), "julia", "clh_abi.jl")).read()
    jl = {m.group(1): re.findall(r"^\s+(\w+)::([\w{},]+)\s*$", m.group(2), re.M)
          for m in re.finditer(r"^struct (\w+)\n(.*?)^end", src, re.M | re.S)}
    ctype = {"Int32": ctypes.c_int32, "Int64": ctypes.c_int64, "Float64": ctypes.c_double}
    for name, cls in (("ClhSchedule", _lib.ClhSchedule), ("ClhSettings", _lib.ClhSettings),
                      ("ClhInfo", _lib.ClhInfo), ("ClhTuneReport", _lib.ClhTuneReport),
                      ("ClhLdlInfo", _lib.ClhLdlInfo)):
        fields = jl[name]
        assert [f for f, _ in fields] == [f for f, _ in cls._fields_], name
        for (f, t), (_, ct) in zip(fields, cls._fields_):
            if t.startswith("NTuple"):
                k, et = re.match(r"NTuple\{(\d+),(\w+)\}", t).groups()
                assert ct._length_ == int(k) and ct._type_ == ctype[et], (name, f)
            elif t in ctype:
                assert ct == ctype[t], (name, f, t)
            else:
                assert ct.__name__ == t, (name, f, t)
