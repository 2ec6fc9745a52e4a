"""The C-ABI library loads and exports every symbol include/qftkb.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qftkb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kb_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import qftkb
    assert os.path.exists(qftkb.LIB_PATH), "libqftkb.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(qftkb.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(qftkb.EXPORTED) == decl


def test_library_exports_nothing_else():
    """Kernel stubs, helpers and template instantiations stay out of the
    dynamic symbol table (linker version script)."""
    import shutil
    import subprocess
    import qftkb
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", qftkb.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = sorted(line.split()[-1] for line in out.splitlines() if line.strip())
    assert names == declared_symbols(), [n for n in names if n not in declared_symbols()][:10]


def test_create_rejects_bad_parameters_without_device():
    """Argument validation happens before any HIP call."""
    from qftkb._abi import KB_ERR_ARG, kb_params, load_library
    L = load_library()
    h = ctypes.c_void_p()
    for bad in [dict(Nx=0), dict(sdim=4), dict(NstepsinMemory=2), dict(NstepsinMemory=100000), dict(dt=0.0),
                dict(ONgroup=0)]:
        p = dict(Nx=4, sdim=3, NstepsinMemory=5, ONgroup=1, dt=0.1, Lambda=1.0, Mass=1.0, Mass2=1.0, pexp=1, device=0)
        p.update(bad)
        par = kb_params(p["Nx"], p["sdim"], p["NstepsinMemory"], p["ONgroup"], p["dt"], p["Lambda"], p["Mass"],
                        p["Mass2"], p["pexp"], p["device"])
        assert L.kb_create(ctypes.byref(par), ctypes.byref(h)) == KB_ERR_ARG, bad


def test_product_path_does_not_import_oracle():
    """The product package must not route through the oracle."""
    pkg = os.path.join(ROOT, "qftdynamicsm.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".jl")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.replace("Oracle-free", ""), os.path.join(dirpath, f)
