"""The C-ABI library loads, exports every declared entry point, and the
ctypes structures have the C layout (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nfgs_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nfgs_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from nfgs_amd import _lib as L
    lib = L.load()
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes layer declares every one of them
    import inspect
    src = inspect.getsource(L)
    assert all(f'"{n}"' in src for n in names)


def test_library_was_built_from_these_sources():
    from nfgs_amd import _lib as L
    from nfgs_amd.build import source_digest
    assert L.load().nfgs_source_digest().decode() == source_digest()


def test_loader_refuses_a_library_built_from_other_sources(monkeypatch):
    """A library older than its sources (a failed rebuild) must not load."""
    from nfgs_amd import _lib as L
    from nfgs_amd import build
    monkeypatch.setattr(build, "source_digest", lambda: "0" * 64)
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(RuntimeError, match="built from other sources"):
        L.load()


def test_last_error_is_empty_string_initially():
    from nfgs_amd import _lib as L
    assert isinstance(L.load().nfgs_last_error(), bytes)


def test_plan_create_rejects_bad_descriptors_without_touching_the_gpu():
    from nfgs_amd import _lib as L
    lib = L.load()
    plan = C.c_void_p()
    bad = L.GridDesc(2, 8, 8, 1, 1, 1, L.NFGS_F32, 0.0, 1.0, 1e-20, 1, 1, 0, 0, 0, 0)
    assert lib.nfgs_plan_create(C.byref(bad), C.byref(plan)) == L.NFGS_ERR_ARG
    assert b">= 3" in lib.nfgs_last_error()
    bad = L.GridDesc(8, 8, 8, 1, 1, 1, L.NFGS_F32, 1.5, 1.0, 1e-20, 1, 1, 0, 0, 0, 0)
    assert lib.nfgs_plan_create(C.byref(bad), C.byref(plan)) == L.NFGS_ERR_ARG
    with pytest.raises(ValueError):
        L.check(L.NFGS_ERR_ARG, "x")
    with pytest.raises(RuntimeError):
        L.check(L.NFGS_ERR_STATE, "x")


def test_struct_layouts_match_the_header(tmp_path):
    from nfgs_amd import _lib as L
    prog = tmp_path / "layout.c"
    fields = {"nfgs_grid_desc": L.GridDesc, "nfgs_boundary": L.Boundary,
              "nfgs_workspace": L.Workspace}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, cls in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(prog), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for cname, cls in fields.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)
    # enum values used by the Python layer
    prog2 = tmp_path / "enums.c"
    prog2.write_text("\n".join(['#include <stdio.h>', f'#include "{HEADER}"', "int main(void){",
                                'printf("%d %d %d %d %d %d %d %d %d\\n", NFGS_NQ, '
                                'NFGS_GDIAG_COUNT, NFGS_Q_ZERO_MODE, NFGS_GD_BY_SUM, '
                                'NFGS_ERR_STATE, NFGS_SD_COUNT, NFGS_SC_COUNT, NFGS_SC_JXB, '
                                'NFGS_SC_DIFF_REL);', "return 0;}"]))
    subprocess.run(["gcc", str(prog2), "-o", str(tmp_path / "enums")], check=True)
    vals = subprocess.run([str(tmp_path / "enums")], capture_output=True, text=True).stdout.split()
    assert list(map(int, vals)) == [L.NQ, L.GDIAG_COUNT, L.Q_ZERO_MODE, L.GD["BY_SUM"],
                                    L.NFGS_ERR_STATE, L.SD_COUNT, L.SC_COUNT, L.SC_JXB,
                                    L.SC_DIFF_REL]
